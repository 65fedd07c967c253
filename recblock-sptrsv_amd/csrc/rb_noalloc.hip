// rb_solve_noalloc: rb_solve for callers that cannot allocate device memory
// (code under stream/graph capture, serving loops that forbid hipMalloc).
//
// The per-solve scratch of a solve -- the permuted b and x vectors and the
// triangular blocks' tickets of the blocked schedule, the slot-order x and
// b/diag buffers of the fused schedule -- comes from one file-scope device
// array of kScratchBytes instead of plan-owned or on-demand allocations.
// Each device the library runs on holds its own copy of the array until the
// process ends.  Its device address is looked up once per device with
// hipGetSymbolAddress (the array's name is only a host-side handle) and that
// address is what the kernels receive.
//
// Every call on a device shares the device's array and leaves it holding
// that call's data: each call re-initialises everything it reads (the
// prologues write b, the sentinel fill and the ticket reset run first).
// Two calls that overlap on one device -- on different streams or from
// different threads -- race on the array; callers must order them.
#include <mutex>

#include "rb_internal.h"

namespace rb {
namespace {

constexpr size_t kScratchBytes = 64ull << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

// device address of this device's scratch array (looked up on the first
// call on each device, with that device current)
rb_status scratch_ptr(int dev, void **out) {
    if (dev < 0 || dev >= kMaxDevices) {
        set_error(RB_ERR_INVALID_ARG, "device id out of range");
        return RB_ERR_INVALID_ARG;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return status_of_hip(e != hipSuccess ? e : hipErrorInvalidSymbol,
                                 "hipGetSymbolAddress(scratch)");
        }
        g_ptr[dev] = ptr;
    }
    *out = g_ptr[dev];
    return RB_OK;
}

}  // namespace
}  // namespace rb

using namespace rb;

extern "C" size_t rb_noalloc_limit(void) { return kScratchBytes; }

extern "C" rb_status rb_solve_noalloc(rb_plan *p, const double *d_b, int ldb,
                                      double *d_x, int ldx, int nrhs,
                                      void *stream) {
    if (!p || nrhs < 0 || (nrhs > 0 && (!d_b || !d_x)) ||
        (nrhs > 0 && (ldb < p->m || ldx < p->m))) {
        set_error(RB_ERR_INVALID_ARG, "invalid solve arguments");
        return RB_ERR_INVALID_ARG;
    }
    if (nrhs == 0 || p->m == 0) {
        set_error(RB_OK, "");
        return RB_OK;
    }
    const bool fused = p->schedule == RB_SCHED_FUSED;
    const size_t need = fused ? 16 * fused_slots(p) : blocked_ws_bytes(p);
    if (need > kScratchBytes) {
        set_error(RB_ERR_INVALID_ARG,
                  "rb_solve_noalloc: the plan needs " + std::to_string(need) +
                      " bytes of scratch per right-hand side, above the "
                      "fixed " + std::to_string(kScratchBytes) +
                      " bytes; use rb_solve");
        return RB_ERR_INVALID_ARG;
    }
    int dev = 0;
    RB_HIP(hipGetDevice(&dev));
    if (dev != p->device) {
        set_error(RB_ERR_INVALID_ARG, "the plan belongs to another device");
        return RB_ERR_INVALID_ARG;
    }
    void *ws = nullptr;
    RB_TRY(scratch_ptr(dev, &ws));
    const hipStream_t s = (hipStream_t)stream;
    p->mark_first = true;
    rowcap_for_call(p, s);
    rb_status st =
        fused ? solve_fused_ws(p, d_b, ldb, d_x, ldx, nrhs, s, (double *)ws,
                               kScratchBytes)
              : solve_blocked_ws(p, d_b, ldb, d_x, ldx, nrhs, s, ws,
                                 kScratchBytes);
    if (st == RB_OK) set_error(RB_OK, "");
    return st;
}
