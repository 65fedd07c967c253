// Device-wide integer scans and a stable key-value radix sort used by plan
// construction (not by the solve).  Thin wrappers over hipCUB.
#include <hipcub/hipcub.hpp>

#include "rb_internal.h"

namespace rb {

namespace {
struct Temp {
    void *p = nullptr;
    size_t bytes = 0;
    rb_status ensure(size_t need) {
        if (need <= bytes) return RB_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (hipMalloc(&p, need) != hipSuccess) {
            (void)hipGetLastError();
            set_error(RB_ERR_NO_MEMORY, "hipMalloc(temp) failed");
            return RB_ERR_NO_MEMORY;
        }
        bytes = need;
        return RB_OK;
    }
    ~Temp() {
        if (p) (void)hipFree(p);
    }
};
}  // namespace

rb_status exclusive_scan_i32(const int *d_in, int *d_out, int n,
                             hipStream_t s) {
    if (n <= 0) return RB_OK;
    Temp t;
    size_t need = 0;
    RB_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, d_in, d_out, n, s));
    RB_TRY(t.ensure(need));
    RB_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, need, d_in, d_out, n, s));
    RB_HIP(hipStreamSynchronize(s));
    return RB_OK;
}

rb_status inclusive_max_i32(const int *d_in, int *d_out, int n,
                            hipStream_t s) {
    if (n <= 0) return RB_OK;
    Temp t;
    size_t need = 0;
    RB_HIP(hipcub::DeviceScan::InclusiveScan(nullptr, need, d_in, d_out,
                                             hipcub::Max(), n, s));
    RB_TRY(t.ensure(need));
    RB_HIP(hipcub::DeviceScan::InclusiveScan(t.p, need, d_in, d_out,
                                             hipcub::Max(), n, s));
    RB_HIP(hipStreamSynchronize(s));
    return RB_OK;
}

rb_status sort_pairs_i32(const int *keys_in, int *keys_out,
                         const int *vals_in, int *vals_out, int n, int bits,
                         hipStream_t s) {
    if (n <= 0) return RB_OK;
    if (bits < 1) bits = 1;
    if (bits > 32) bits = 32;
    Temp t;
    size_t need = 0;
    RB_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need, keys_in,
                                              keys_out, vals_in, vals_out, n,
                                              0, bits, s));
    RB_TRY(t.ensure(need));
    RB_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, need, keys_in, keys_out,
                                              vals_in, vals_out, n, 0, bits,
                                              s));
    RB_HIP(hipStreamSynchronize(s));
    return RB_OK;
}

}  // namespace rb
