// Fused schedule: the whole solve in one persistent launch (plus two
// fully parallel passes that move b into and x out of solve order).
//
// Rows are grouped into chains that each belong to one consumer wave.  A
// wave solves its rows in level order, one *step* (at most 64 rows of one
// level, one per lane) at a time.  Every off-diagonal entry of a row is read
// from one of three places:
//   previous step, lane l-1 / l / l+1   registers, moved across lanes by DPP
//   same wave, within the last R steps  the wave's LDS ring of recent x
//   anything else ("global")            x in slot order in device memory,
//                                       a sentinel NaN until solved
//
// Assignment (host, at plan build): levels are walked in order; a row joins
// the wave of its BFS parent (the dependency solved last) while that wave
// has room at this level, otherwise a new wave (up to W) or the least
// loaded one.  On grid-like matrices this keeps every row chain inside one
// wave and most entries on the DPP path, so a level costs a few register
// operations instead of a trip through the memory hierarchy.
//
// A solve is three launches:
//   prologue  writes b[orig] of every slot into the slot's step record;
//   k_fused   one workgroup per consumer wave, plus a resolver wave:
//     consumer  streams its step records (FIFO slot, 1/diag, b, coefficients,
//               ring sources) into registers several steps ahead; per step
//               x = (b - fifo - ring - dpp) / diag, written to the LDS ring
//               and, contiguously, to the slot-order x array;
//     resolver  polls the global dependencies in order, many loads in
//               flight, and hands the consumer one partial sum per
//               (step, lane) through an LDS FIFO;
//   epilogue  scatters the slot-order x to the caller's x.
// Keeping the gathers and scatters by row id out of the consumer matters:
// a wave's scattered access occupies the compute unit's memory pipeline
// for hundreds of cycles, which would stall the record stream of the
// dependency chain; in the parallel passes it costs bandwidth only.
// The slot-order x array is double buffered: a solve reads and writes one
// buffer and re-arms the published slots of the other with the sentinel.
// Every wait is bounded: on a timeout the kernel sets a status flag and
// stops instead of hanging.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "rb_internal.h"

namespace rb {

namespace {

constexpr int kR = 32;          // LDS x ring depth (steps), a power of two
constexpr int kZeroOff = kR * 64 * 8;  // byte offset of the ring's zero
constexpr int kWG = 128;        // consumer wave + resolver wave

// The FIFO is indexed by step: slot (s mod kQ, lane) holds the partial sum
// of lane's global entries at step s, or zero when it has none.  The
// resolver writes a step's groups into their slots, the consumer reads its
// lane's slot (no per-lane FIFO offset is stored anywhere) and writes zero
// back for the step kQ later.  The resolver may fill step s once the
// consumer has consumed step s - kQ; kQ is at least two blocks of steps.
constexpr int kQ = 64;
// a block-checked consumer reads a step's FIFO value this many steps early
constexpr int kFP = 3;
// per global entry: (step << kStepShift) | (lane << 1) | last-of-group
constexpr int kStepShift = 7;

// Per-lane record of one step, as 32-bit words; every double starts at an
// even word so it is an aligned register pair of the load that brings it:
//   NDPP 3: [0..1] coefficient of lane l (previous step), [2..3] of lane
//           l-1, [4..5] of lane l+1; NDPP 1 (2): [0..1] lane l, [2..3]
//           lane l-1 (l+1)
//   NRING ring sources (byte offsets into the ring), padded to even, then
//   NRING coefficients
// padded to whole 16-byte words (none at all for NDPP 0 without a ring);
// word k of step s for lane l lives at rec[(s * K + k) * 64 + l].  Every
// loaded word is read (a last word with two ints is loaded as 8 bytes): a
// loaded word the step does not use would free its registers while the
// load is in flight.
// Every coefficient is stored as -a_ij / a_ii, and the global entries'
// values as a_ij / a_ii, so a row is x_i = b_i/a_ii + sum(coef * x_j) - fifo:
// fused multiply-adds only, without a subtraction or a product by 1/diag
// on the chain from the previous step's x.  b_i/a_ii comes per solve from a
// separate slot-order array (one per right-hand side), written by the
// prologue.

// record prefetch distance D in steps for K words per lane; the records
// of D + 1 steps are held in registers (4K + 2 each), and a wave's step
// count is padded to a multiple of D + 1.  A step issues K + 2 vector
// memory operations (its record words, b/diag, the store of x); the ones
// in flight stay within the 63 the vector memory counter tracks
// (D * (K + 2) <= 60), and for K > 2 the register ring stays near 140
// registers so that three waves fit a SIMD (several right-hand sides).
// At these depths a step is not bound by memory latency but by the data
// its instructions move through the SIMD: on the 1500^2 Laplacian (K = 1)
// D = 15 ran 43 ns per step, D = 20 50 ns and D = 11 51 ns, and the block
// of D + 1 steps is also the grain at which one wave waits for another.
constexpr int prefetch_steps(int K) {
    return K == 0 ? 20 : K <= 2 ? 15 : K == 3 ? 10 : K == 4 ? 7
         : K == 5 ? 6 : K <= 7 ? 4 : K == 8 ? 3 : 2;
}
constexpr int block_steps(int K) { return prefetch_steps(K) + 1; }
static_assert(2 * block_steps(0) <= kQ, "FIFO too small for two blocks");

// DPP formats: 0 none; 3 lanes l, l-1, l+1 of the previous step; 1 lanes
// l, l-1 only; 2 lanes l, l+1 only (a grid's forward / backward solve
// needs one side: one DPP move pair and one FMA less on the chain)
constexpr int dpp_ints(int ndpp) { return ndpp == 3 ? 6 : ndpp ? 4 : 0; }

// record words per lane for a format
constexpr int words_of(int ndpp, int nring) {
    return (dpp_ints(ndpp) + ((nring + 1) & ~1) + 2 * nring + 3) / 4;
}

template <int NDPP, int NRING>
struct Fmt {
    static constexpr int kSrc = dpp_ints(NDPP);
    static constexpr int kVal = kSrc + ((NRING + 1) & ~1);
    static constexpr int kInts = kVal + 2 * NRING;
    static constexpr int K = words_of(NDPP, NRING);
    // the last word holds two ints only: it is loaded as 8 bytes, so no
    // loaded register is left unused (the allocator would reuse it while
    // the load is in flight and have to wait for the load)
    static constexpr bool kTail2 = kInts % 4 == 2;
    static_assert(K == (kInts + 3) / 4, "record width");
    static constexpr int D = prefetch_steps(K);
    static constexpr int NB = D + 1;  // record buffers, steps per block
    static_assert(2 * NB <= kQ, "FIFO too small for two blocks");
};

__device__ inline int lds_load(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// the same counter as a wave-uniform (scalar) value
__device__ inline int lds_load_u(const int *p) {
    return __builtin_amdgcn_readfirstlane(lds_load(p));
}
__device__ inline void lds_store(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void wg_release() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
}
__device__ inline void wg_acquire() {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
__device__ inline double ld_x(const double *p) {
    return __longlong_as_double((long long)__hip_atomic_load(
        (const unsigned long long *)p, __ATOMIC_RELAXED,
        __HIP_MEMORY_SCOPE_AGENT));
}
__device__ inline bool pending(double v) {
    return __double_as_longlong(v) == (long long)kPending;
}
__device__ inline int status_now(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double dbl(int lo, int hi) { return __hiloint2double(hi, lo); }

// Buffer resources: a wave-uniform base and a size in bytes.  A buffer
// access adds a scalar offset and a 32-bit lane offset in the memory unit
// (no 64-bit address arithmetic in vector registers), and one outside the
// size reads zero / is dropped.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
constexpr int kBufWord3 = 0x00020000;  // raw 32-bit buffer (CDNA)
constexpr int kSC1 = 16;               // cache policy: device-coherent (sc1)
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(
        const_cast<void *>(p), (short)0, (int)min(bytes, 0x7fffffffL), kBufWord3);
}

// T: the value type of the solve (FP64, or FP32 for the inner solve of the
// FP32 + FP64 refinement)
template <class T>
struct FusedArgsT {
    const int *step0;    // [W+1] first step of each wave
    const int *gent0;    // [W+1] first global entry of each wave
    const int *pub0;     // [W+1] first published slot of each wave
    const int4 *rec;     // [nsteps * K * 64] step records (b filled in)
    const int *gsrc;     // [nent] slot of the dependency
    const T *gval;       // [nent]
    const int *gmeta;    // [nent] group ordinal * 2 + last-of-group
    const int *pub;      // [npub] published slots, grouped by wave
    T *xq;               // [nr][nsteps * 64] x in slot order (this solve)
    T *xq_next;          // the other buffers, re-armed for the next solve
                         // (null: nothing to re-arm)
    const T *brec;       // [nr][nsteps * 64] b/diag in slot order
    long nslots;         // nsteps * 64: stride between right-hand sides
    int W;               // consumer waves per right-hand side
    int *status;
    unsigned spin;       // polls before a bounded wait gives up
    // optional trace: per workgroup 16 counters of wall-clock ticks
    // (10 ns): [1] consumer waits, [2] consumer total, [3] steps,
    // [7] resolver FIFO-space wait, [8] resolver poll, [9] batches,
    // [10] poll rounds
    unsigned long long *trace;
    // optional: the launch does nothing when *left == 0 (a refinement
    // whose columns have all converged)
    const int *left;
};
using FusedArgs = FusedArgsT<double>;

__device__ inline unsigned long long now() { return wall_clock64(); }

// lane l receives lane l-1's value (lane 0 receives 0)
__device__ inline double dpp_from_lower(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
// lane l receives lane l+1's value (lane 63 receives 0)
__device__ inline double dpp_from_upper(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

// Per value type: record coefficients, the LDS ring offsets, the DPP moves,
// the sentinel and the loads and stores of x and b/diag.  The FP32 solve
// reads the FP64 records and rounds each coefficient where it is used
// (off the dependency chain): a record with an FP32 coefficient in half of
// each pair would leave the other half of a loaded register pair unused,
// and the register allocator then reuses it while the load is in flight
// and waits for that load on every step.  Ring offsets are stored for
// 8-byte slots and halved for 4-byte ones.
template <class T> struct FV;
template <> struct FV<double> {
    using U = unsigned long long;
    static constexpr U kPend = kPending;
    __device__ static double coef(int lo, int hi) { return dbl(lo, hi); }
    __device__ static int ring_off(int o) { return o; }
    __device__ static double from_lower(double v) { return dpp_from_lower(v); }
    __device__ static double from_upper(double v) { return dpp_from_upper(v); }
    __device__ static bool pend(double v) { return pending(v); }
    __device__ static double ld(const double *p) { return ld_x(p); }
    __device__ static double fma(double a, double b, double c) {
        return __builtin_fma(a, b, c);
    }
    __device__ static double load(__amdgpu_buffer_rsrc_t r, int off, int soff) {
        return __builtin_bit_cast(double,
                                  __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
    }
    template <int AUX>
    __device__ static void store(double v, __amdgpu_buffer_rsrc_t r, int off,
                                 int soff) {
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, off,
                                              soff, AUX);
    }
};
template <> struct FV<float> {
    using U = unsigned int;
    static constexpr U kPend = kPending32;
    __device__ static float coef(int lo, int hi) { return (float)dbl(lo, hi); }
    __device__ static int ring_off(int o) { return o >> 1; }
    __device__ static float from_lower(float v) {
        return __int_as_float(
            __builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
    }
    __device__ static float from_upper(float v) {
        return __int_as_float(
            __builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
    }
    __device__ static bool pend(float v) { return __float_as_uint(v) == kPending32; }
    __device__ static float ld(const float *p) {
        return __uint_as_float(__hip_atomic_load((const unsigned *)p, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT));
    }
    __device__ static float fma(float a, float b, float c) {
        return __builtin_fmaf(a, b, c);
    }
    __device__ static float load(__amdgpu_buffer_rsrc_t r, int off, int soff) {
        return __builtin_bit_cast(float,
                                  __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
    }
    template <int AUX>
    __device__ static void store(float v, __amdgpu_buffer_rsrc_t r, int off,
                                 int soff) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off,
                                              soff, AUX);
    }
};

struct Shared {
    int ready;    // steps whose groups are all in the FIFO (resolver)
    int cons;     // steps consumed (consumer)
    int abort;
};

// Consumer wave.  The records of the next D steps are in flight into a
// ring of D + 1 register buffers: step s first issues the load of step
// s + D into the buffer step s - 1 has freed, then computes with its own.
// The load then never waits for this step's chain or store, which a
// wave's in-order issue would otherwise put in front of it.  The step body
// is instantiated once per position U of a block of D + 1 steps, so every
// buffer index is a compile-time constant and the ring stays in registers;
// a wave's step count is a multiple of D + 1, so blocks have no exit and
// the loop latch sees one register assignment of the ring.
// BLK: a step has no branch (a branch costs a wave about as much as the
// whole dependency chain of a step, taken or not): the FIFO is checked
// once per block, for every step of the block.  That is only safe when no
// block waits, directly or through other blocks, for a row of a later
// step of its own block -- the host checks the block wait graph for cycles
// and otherwise builds with BLK off, which checks the FIFO at every step.
template <int NDPP, int NRING, bool BLK, class T>
struct Consumer {
    using F = Fmt<NDPP, NRING>;
    static constexpr int K = F::K, D = F::D, NB = F::NB;
    static constexpr int kStep = 64 * (int)sizeof(T);  // bytes of x per step

    const FusedArgsT<T> &a;
    const int lane, ns, last;
    const int lane16, laneB;  // byte offsets of the lane's record / x
    // this wave's records, slot-order x and b/diag
    const __amdgpu_buffer_rsrc_t rec, xq, br;
    T *xring;
    T *fifo;
    Shared &sh;
    int4 rb[NB][K > 0 ? K : 1];  // ring of step records
    T bb[NB];                    // ring of b/diag
    int fifo_avail = 0;          // steps ready (scalar)
    T fq[NB];                    // FIFO values of the block's steps (BLK)
    T xprev = 0;
    bool ok = true;
    unsigned long long t_wait = 0;

    __device__ __forceinline__ Consumer(const FusedArgsT<T> &a_, int w, int r,
                                        int lane_, T *xring_, T *fifo_,
                                        Shared &sh_)
        : a(a_), lane(lane_), ns(a_.step0[w + 1] - a_.step0[w]),
          last(a_.step0[w + 1] - a_.step0[w] - 1),
          lane16(lane_ * 16), laneB(lane_ * (int)sizeof(T)),
          rec(buf_rsrc(a_.rec + (size_t)a_.step0[w] * K * 64, (long)ns * K * 1024)),
          xq(buf_rsrc(a_.xq + r * a_.nslots + (size_t)a_.step0[w] * 64,
                      (long)ns * kStep)),
          br(buf_rsrc(a_.brec + r * a_.nslots + (size_t)a_.step0[w] * 64,
                      (long)ns * kStep)),
          xring(xring_), fifo(fifo_), sh(sh_) {}

    // the lane's FIFO slot of step s; read, then cleared for step s + kQ
    __device__ __forceinline__ T take_fifo(int s) {
        T *p = fifo + (s & (kQ - 1)) * 64 + lane;
        const T v = *p;
        *p = T(0);
        return v;
    }

    __device__ __forceinline__ void load_rec(int B, int t) {
        const int tc = min(t, last);
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (F::kTail2 && k == K - 1) {
                const int2 v = __builtin_bit_cast(
                    int2, __builtin_amdgcn_raw_buffer_load_b64(rec, lane16 + k * 1024,
                                                               tc * K * 1024, 0));
                rb[B][k] = make_int4(v.x, v.y, 0, 0);
            } else {
                rb[B][k] = __builtin_bit_cast(
                    int4, __builtin_amdgcn_raw_buffer_load_b128(rec, lane16 + k * 1024,
                                                                 tc * K * 1024, 0));
            }
        }
        bb[B] = FV<T>::load(br, laneB, tc * kStep);
    }

    // records of steps 0 .. D-1, oldest first.  Each is followed by a store
    // through an empty buffer resource (no record is in range, so the
    // memory unit drops it), so that the loop is entered with the same mix
    // of loads and stores in flight as every later block has and the
    // compiler's wait counts at the loop head are those of the steady state.
    template <int U>
    __device__ __forceinline__ void prime() {
        if constexpr (U < D) {
            load_rec(U, U);
            FV<T>::template store<0>(T(0), buf_rsrc(a.xq, 0), laneB, 0);
            __builtin_amdgcn_sched_barrier(0);
            prime<U + 1>();
        }
    }

    // wait until the resolver has every group of the steps below `need`;
    // on a timeout or an abort elsewhere, mark the workgroup aborted (the
    // remaining steps of the block then run through without waiting)
    __device__ __forceinline__ void wait_fifo(int need) {
        const unsigned long long tw = a.trace ? now() : 0;
        unsigned spins = 0;
#pragma nounroll
        while ((fifo_avail = lds_load_u(&sh.ready)) < need) {
            if (++spins > a.spin || lds_load_u(&sh.abort)) {
                if (ok && lane == 0) {
                    atomicOr(a.status, 2);
                    lds_store(&sh.abort, 1);
                }
                ok = false;
                fifo_avail = 1 << 30;
                return;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        wg_acquire();
        if (a.trace) t_wait += now() - tw;
    }

    template <int U>
    __device__ __forceinline__ void step(int sb) {
        if constexpr (U < NB) {
            const int s = sb + U;
            // the record of step s + D goes into the buffer of step s - 1
            load_rec((U + D) % NB, s + D);
            __builtin_amdgcn_sched_barrier(0);
            int f[4 * (K > 0 ? K : 1)];
#pragma unroll
            for (int k = 0; k < K; k++) {
                f[4 * k] = rb[U][k].x;
                f[4 * k + 1] = rb[U][k].y;
                f[4 * k + 2] = rb[U][k].z;
                f[4 * k + 3] = rb[U][k].w;
            }
            T bv = bb[U];
            if constexpr (U == 0 && BLK) {
                // every step of the block must be ready; then the FIFO
                // values of its first kFP steps are read
                fifo_avail = __builtin_amdgcn_readfirstlane(fifo_avail);
                if (__builtin_expect(fifo_avail < sb + NB, 0)) wait_fifo(sb + NB);
#pragma unroll
                for (int j = 0; j < kFP && j < NB; j++) fq[j] = take_fifo(sb + j);
            }
            if constexpr (!BLK) {
                fifo_avail = __builtin_amdgcn_readfirstlane(fifo_avail);
                if (__builtin_expect(fifo_avail <= s, 0)) wait_fifo(s + 1);
            }
            // With BLK every step of the block is ready once the block has
            // started, so a step's FIFO value is read kFP steps early and
            // the LDS latency leaves the chain.
            T fv;
            if constexpr (BLK) {
                fv = fq[U];
                if constexpr (U + kFP < NB) fq[U + kFP] = take_fifo(s + kFP);
                // issue the read here, not behind this step's chain
                __builtin_amdgcn_sched_barrier(0);
            } else {
                fv = take_fifo(s);
            }
            // with BLK the FIFO value is at hand (read a step early) and
            // goes in first; otherwise it is subtracted last, so the wait
            // for its LDS read sits behind the chain
            if constexpr (BLK) bv -= fv;
#pragma unroll
            for (int j = 0; j < NRING; j++) {
                const T xv = *(const T *)((const char *)xring +
                                          FV<T>::ring_off(f[F::kSrc + j]));
                bv = FV<T>::fma(FV<T>::coef(f[F::kVal + 2 * j], f[F::kVal + 2 * j + 1]),
                                xv, bv);
            }
            // the chain that carries the previous step's x: one FMA per
            // neighbour (the first needs no DPP)
            T acc = bv;
            if constexpr (NDPP == 3) {
                const T xl = FV<T>::from_lower(xprev);
                const T xr = FV<T>::from_upper(xprev);
                acc = FV<T>::fma(FV<T>::coef(f[0], f[1]), xprev, acc);
                acc = FV<T>::fma(FV<T>::coef(f[2], f[3]), xl, acc);
                acc = FV<T>::fma(FV<T>::coef(f[4], f[5]), xr, acc);
            } else if constexpr (NDPP == 1 || NDPP == 2) {
                const T xn = NDPP == 1 ? FV<T>::from_lower(xprev) : FV<T>::from_upper(xprev);
                acc = FV<T>::fma(FV<T>::coef(f[0], f[1]), xprev, acc);
                acc = FV<T>::fma(FV<T>::coef(f[2], f[3]), xn, acc);
            }
            const T xv = BLK ? acc : acc - fv;
            xprev = xv;
            if constexpr (NRING > 0) xring[(s & (kR - 1)) * 64 + lane] = xv;
            // every slot goes out (the epilogue reads them); device-coherent
            // (as an agent-scope store) so that published slots are visible
            // to the other waves' resolvers
            FV<T>::template store<kSC1>(xv, xq, laneB, s * kStep);
            if constexpr (U == NB - 1) {
                // the block's FIFO slots are cleared: the resolver may
                // refill them
                wg_release();
                if (lane == 0) lds_store(&sh.cons, sb + NB);
            }
            step<U + 1>(sb);
        }
    }

    __device__ __forceinline__ void run(int w, int r) {
        const unsigned long long t0 = a.trace ? now() : 0;
        prime<0>();
        for (int sb = 0; sb < ns && ok; sb += NB) step<0>(sb);
        if (lane == 0) lds_store(&sh.cons, 1 << 30);
        if (a.trace && lane == 0 && r == 0) {
            unsigned long long *tr = a.trace + (size_t)w * 16;
            tr[1] = t_wait;
            tr[2] = now() - t0;
            tr[3] = ns;
        }
    }
};

__device__ inline unsigned long long rotr64(unsigned long long x, int r) {
    return r ? (x >> r) | (x << (64 - r)) : x;
}

// Resolver wave: polls the wave's global entries in consumption order
// ((step, lane) order), sums each group (the global entries of one row) in
// entry order, writes it to the row's FIFO slot, and publishes in `ready`
// the first step that still lacks a group.
// The entries in flight form a window of 64 consecutive ones starting at
// the first unpublished entry `base`: lane l holds the entries e0 + l +
// 64 k, so when a prefix of the window is published, exactly those lanes
// move on to their next entry (its metadata is already in registers) and
// poll it at once, while the others poll again.  Every round -- one trip to
// memory -- publishes whatever became ready during the last one; there is
// no pause at batch boundaries.  An entry whose step the FIFO cannot hold
// yet (kQ steps beyond the consumer) waits like a pending one.  A group
// longer than the window is summed across rounds (`carry`).
template <class T>
__device__ __forceinline__ void resolve(const FusedArgsT<T> &a, int w, int r,
                                        int lane, T *fifo, T *rtmp, int *gtmp,
                                        T *ctmp, Shared &sh) {
    const T *xq = a.xq + r * a.nslots;
    const int e0 = a.gent0[w], e1 = a.gent0[w + 1];
    constexpr int kAll = 1 << 30;
    // steps before the first group need nothing
    if (lane == 0)
        lds_store(&sh.ready, e0 < e1 ? a.gmeta[e0] >> kStepShift : kAll);
    // re-arm this wave's published slots of the next solve's buffer (no
    // wave of this launch reads that buffer; nothing to do when the caller
    // fills every slot with the sentinel before each launch), eight slots
    // per lane in flight: done one slot at a time this delayed the first
    // polls by ~10 us on the 1500^2 Laplacian
    if (a.xq_next) {
        typename FV<T>::U *nxt = (typename FV<T>::U *)(a.xq_next + r * a.nslots);
        const int p0 = a.pub0[w], p1 = a.pub0[w + 1];
        for (int i0 = p0; i0 < p1; i0 += 8 * 64) {
            int sl[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int i = i0 + j * 64 + lane;
                sl[j] = i < p1 ? a.pub[i] : -1;
            }
#pragma unroll
            for (int j = 0; j < 8; j++)
                if (sl[j] >= 0) nxt[sl[j]] = FV<T>::kPend;
        }
    }
    unsigned long long t_poll = 0, nb = 0, nround = 0;
    const unsigned long long t0 = a.trace ? now() : 0;
    int e = e0 + lane;  // this lane's entry
    int src = 0, gm = 0, src2 = 0, gm2 = 0;
    T gv = 0, gv2 = 0;  // metadata of e and of e + 64
    if (e < e1) {
        src = a.gsrc[e];
        gv = a.gval[e];
        gm = a.gmeta[e];
    }
    if (e + 64 < e1) {
        src2 = a.gsrc[e + 64];
        gv2 = a.gval[e + 64];
        gm2 = a.gmeta[e + 64];
    }
    T xv = e < e1 ? FV<T>::ld(&xq[src]) : T(0);
    int base = e0;
    T carry = 0;
    int carry_g = -1;  // key (step << 6 | lane) of the group carried over
    unsigned spins = 0;
    while (base < e1) {
        nround++;
        const int lim = lds_load_u(&sh.cons) + kQ;
        const bool live = e < e1;
        const bool pend = live && FV<T>::pend(xv);
        const int key = gm >> 1, stp = gm >> kStepShift;
        const bool hold = live && (pend || stp >= lim);
        const int rot = (base - e0) & 63;   // lane of entry `base`
        const int pos = (lane - rot) & 63;  // this lane's place in the window
        const int avail = min(64, e1 - base);
        // window bits: bit i is entry base + i
        const unsigned long long mh = rotr64(__ballot(hold), rot);
        const int fp = min(avail, mh ? (int)__ffsll((long long)mh) - 1 : 64);
        int hi;  // publish window entries [0, hi)
        if (fp >= avail) {
            hi = avail;
        } else {
            const unsigned long long me =
                rotr64(__ballot(live && (gm & 1)), rot) & ((1ull << fp) - 1);
            hi = me ? 64 - __clzll((long long)me) : 0;
        }
        bool adv = false;
        if (hi > 0) {
            nb++;
            spins = 0;
            rtmp[lane] = live && !pend ? gv * xv : T(0);
            gtmp[lane] = live ? gm : -2;
            const int kprev = __shfl(key, (lane + 63) & 63, 64);
            // group heads sum their entries in order; a group begun in an
            // earlier round starts from the carried sum
            if (pos < hi && (pos == 0 || kprev != key)) {
                T sum = (pos == 0 && carry_g == key) ? carry : T(0);
                int q = pos, l = lane;
                while (true) {
                    sum += rtmp[l];
                    if (q + 1 >= hi || (gtmp[(l + 1) & 63] >> 1) != key) break;
                    q++;
                    l = (l + 1) & 63;
                }
                if (gtmp[l] & 1)  // complete: into the row's slot
                    fifo[(stp & (kQ - 1)) * 64 + (key & 63)] = sum;
                else
                    *ctmp = sum;  // the window's last group continues
            }
            const int gl = __shfl(gm, (rot + hi - 1) & 63, 64);
            int rdy;
            if (gl & 1) {
                carry_g = -1;
                rdy = hi < avail ? __shfl(stp, (rot + hi) & 63, 64)
                    : base + hi >= e1 ? kAll
                    : __shfl(gm2, rot, 64) >> kStepShift;  // entry base + 64
            } else {
                carry = *ctmp;
                carry_g = gl >> 1;
                rdy = gl >> kStepShift;
            }
            wg_release();
            if (lane == 0) lds_store(&sh.ready, rdy);
            // the published lanes move on to their next entry
            adv = live && pos < hi;
            if (adv) {
                e += 64;
                src = src2;
                gv = gv2;
                gm = gm2;
            }
            base += hi;
        } else {
            ++spins;
            if (spins > a.spin || lds_load_u(&sh.abort) ||
                ((spins & 1023) == 0 && (status_now(a.status) & 7))) {
                if (lane == 0) {
                    atomicOr(a.status, 4);
                    lds_store(&sh.abort, 1);
                }
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        // new entries and pending ones poll; then the next metadata
        if ((adv && e < e1) || (!adv && pend)) xv = FV<T>::ld(&xq[src]);
        if (adv && e + 64 < e1) {
            src2 = a.gsrc[e + 64];
            gv2 = a.gval[e + 64];
            gm2 = a.gmeta[e + 64];
        }
        if (lds_load_u(&sh.abort)) break;
    }
    if (a.trace) t_poll = now() - t0;
    if (a.trace && lane == 0 && r == 0) {
        unsigned long long *tr = a.trace + (size_t)w * 16;
        tr[7] = 0;
        tr[8] = t_poll;
        tr[9] = nb;
        tr[10] = nround;
    }
}

// Waves per SIMD the register allocation aims at.  MULTI (launches of
// several right-hand sides, where co-resident workgroups decide how many
// columns one launch takes): 2 for the block-checked kernels (their LDS
// holds a compute unit to about four workgroups).  Single
// right-hand sides and per-step-checked kernels keep their registers (a
// tighter allocation costs the one wave ~2 % per step).
constexpr int occupancy_target(bool blk, bool multi) {
    return multi && blk ? 2 : 1;
}

// One workgroup per (consumer wave w, right-hand side r): blockIdx.x =
// r * W + w.  Right-hand sides share the records and nothing else.
template <int NDPP, int NRING, bool BLK, bool MULTI, class T>
__global__ void __launch_bounds__(kWG)
    __attribute__((amdgpu_waves_per_eu(occupancy_target(BLK, MULTI))))
    k_fused(FusedArgsT<T> a) {
    __shared__ T xring[NRING > 0 ? kR * 64 + 1 : 1];
    __shared__ T fifo[kQ * 64];
    __shared__ T rtmp[64];
    __shared__ int gtmp[64];
    __shared__ T ctmp;
    __shared__ Shared sh;

    if (a.left && *a.left == 0) return;  // every workgroup alike
    int w = blockIdx.x % a.W;
    const int r = blockIdx.x / a.W;
    if (gridDim.x == (unsigned)a.W) {
        // one column: workgroups go to the 8 XCDs round-robin (XCD =
        // blockIdx mod 8), so XCD x takes the consecutive waves from
        // x q + min(x, W mod 8), q = W / 8 (a bijection onto [0, W)): most
        // wave-to-next-wave hand-offs stay in one XCD's L2.  Laplacian
        // 1500^2, 24 waves: 216.2 -> 213.1 us.
        const int b = blockIdx.x, x = b & 7, k = b >> 3, q = a.W >> 3, rem = a.W & 7;
        w = x * q + min(x, rem) + k;
    }
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    for (int i = threadIdx.x; i < kQ * 64; i += kWG) fifo[i] = T(0);
    if (threadIdx.x == 0) {
        sh.ready = 0;
        sh.cons = 0;
        sh.abort = 0;
        if constexpr (NRING > 0) xring[kR * 64] = T(0);
    }
    __syncthreads();
    if (wave == 0) {
        Consumer<NDPP, NRING, BLK, T> c(a, w, r, lane, xring, fifo, sh);
        c.run(w, r);
    } else if (wave == 1) {
        resolve<T>(a, w, r, lane, fifo, rtmp, gtmp, &ctmp, sh);
    }
}

// b/diag into slot order and x out of it, one tile of kTile consecutive
// slots per workgroup and one right-hand side per blockIdx.y.  Inside a tile
// the slots are listed by ascending row (trow; padding lanes last, as -1),
// tperm holds each listed row's slot offset in the tile and tinvd its
// 1/diag.  The caller's b and x are then read and written in row order --
// a chain's rows over consecutive steps are runs of consecutive rows on
// stencil matrices -- and the slot-order side is a contiguous copy through
// LDS.
constexpr int kTile = 1024, kTileThreads = 256;
// tile offsets sort in the low 12 bits of a key and fit an unsigned short
static_assert(kTile <= 4096 && (kTile & (kTile - 1)) == 0, "kTile");

// (left: as for k_fused, nothing to do when *left == 0; b of type B is
// rounded to T first: the FP32 solves of the refinement read the FP64
// residual)
template <class T, class B = T>
__global__ void __launch_bounds__(kTileThreads)
    k_fused_prologue(const int *__restrict__ trow,
                     const unsigned short *__restrict__ tperm,
                     const T *__restrict__ tinvd, long nslots,
                     const B *__restrict__ b, long ldb, T *brec,
                     const int *__restrict__ left) {
    if (left && *left == 0) return;
    __shared__ T buf[kTile];
    const long base = (long)blockIdx.x * kTile;
    const int cnt = (int)min((long)kTile, nslots - base);
    const B *br = b + blockIdx.y * ldb;
#pragma unroll
    for (int k = 0; k < kTile / kTileThreads; k++) {
        const int i = k * kTileThreads + threadIdx.x;
        if (i < cnt) {
            const int row = trow[base + i];
            buf[tperm[base + i]] = row >= 0 ? T(br[row]) * tinvd[base + i] : T(0);
        }
    }
    __syncthreads();
    T *out = brec + blockIdx.y * nslots + base;
#pragma unroll
    for (int k = 0; k < kTile / kTileThreads; k++) {
        const int i = k * kTileThreads + threadIdx.x;
        if (i < cnt) out[i] = buf[i];
    }
}

template <class T>
__global__ void __launch_bounds__(kTileThreads)
    k_fused_epilogue(const int *__restrict__ trow,
                     const unsigned short *__restrict__ tperm, long nslots,
                     const T *__restrict__ xq, T *x, long ldx,
                     const int *status, const int *__restrict__ left) {
    if (left && *left == 0) return;
    __shared__ T buf[kTile];
    const long base = (long)blockIdx.x * kTile;
    const int cnt = (int)min((long)kTile, nslots - base);
    const T *in = xq + blockIdx.y * nslots + base;
    // after a watchdog trip every x of the plan's solves is a quiet NaN
    const bool tripped = (*status & 7) != 0;
#pragma unroll
    for (int k = 0; k < kTile / kTileThreads; k++) {
        const int i = k * kTileThreads + threadIdx.x;
        if (i < cnt) buf[i] = tripped ? T(__builtin_nan("")) : in[i];
    }
    __syncthreads();
    T *xr = x + blockIdx.y * ldx;
#pragma unroll
    for (int k = 0; k < kTile / kTileThreads; k++) {
        const int i = k * kTileThreads + threadIdx.x;
        if (i < cnt) {
            const int row = trow[base + i];
            if (row >= 0) xr[row] = buf[tperm[base + i]];
        }
    }
}

// The refinement's epilogue on the FP32 schedule: x (FP64, caller's order)
// += the FP32 correction of every column not yet converged (done[c] == 0),
// and every slot read goes back to the sentinel, so the buffers are ready
// for the next correction without a pass of their own.
__global__ void __launch_bounds__(kTileThreads)
    k_fused_epilogue_ir(const int *__restrict__ trow,
                        const unsigned short *__restrict__ tperm, long nslots,
                        float *xq, double *x, long ldx, const int *status,
                        const int *__restrict__ done,
                        const int *__restrict__ left) {
    if (left && *left == 0) return;
    __shared__ float buf[kTile];
    const long base = (long)blockIdx.x * kTile;
    const int cnt = (int)min((long)kTile, nslots - base);
    float *in = xq + blockIdx.y * nslots + base;
    const bool tripped = (*status & 7) != 0;
#pragma unroll
    for (int k = 0; k < kTile / kTileThreads; k++) {
        const int i = k * kTileThreads + threadIdx.x;
        if (i < cnt) {
            buf[i] = tripped ? __builtin_nanf("") : in[i];
            in[i] = __uint_as_float(kPending32);
        }
    }
    __syncthreads();
    if (done[blockIdx.y]) return;  // the whole workgroup
    double *xr = x + blockIdx.y * ldx;
#pragma unroll
    for (int k = 0; k < kTile / kTileThreads; k++) {
        const int i = k * kTileThreads + threadIdx.x;
        if (i < cnt) {
            const int row = trow[base + i];
            if (row >= 0) xr[row] += (double)buf[tperm[base + i]];
        }
    }
}

template <class T>
__global__ void k_fill_pending(T *p, long n, const int *__restrict__ left) {
    if (left && *left == 0) return;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ((typename FV<T>::U *)p)[i] = FV<T>::kPend;
}

// FP32 copies of the fused schedule's global coefficients and 1/diag
__global__ void k_to32(const double *__restrict__ in, float *out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

// The values of the fused tables, computed where the host left (entry e,
// permuted row r) index pairs (e = -1: zero): record coefficients
// -val[e] / diag[r] (val: the permuted CSR's values, invd: 1/diag in
// permuted order), one thread per lane-step record of K words ...
__global__ void k_fill_rec_values(int4 *rec, long nls, int K, int kSrc, int nring,
                                  int kVal, const double *__restrict__ val,
                                  const double *__restrict__ invd) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nls) return;
    const long st = t / 64;
    const int l = (int)(t % 64);
    int f[32];
    for (int k = 0; k < K; k++) {
        const int4 w = rec[(st * K + k) * 64 + l];
        f[4 * k] = w.x;
        f[4 * k + 1] = w.y;
        f[4 * k + 2] = w.z;
        f[4 * k + 3] = w.w;
    }
    auto fill = [&](int i) {
        const int e = f[i], r = f[i + 1];
        const double v = e < 0 ? 0.0 : -val[e] * invd[r];
        f[i] = __double2loint(v);
        f[i + 1] = __double2hiint(v);
    };
    for (int i = 0; i < kSrc; i += 2) fill(i);
    for (int j = 0; j < nring; j++) fill(kVal + 2 * j);
    for (int k = 0; k < K; k++)
        rec[(st * K + k) * 64 + l] = make_int4(f[4 * k], f[4 * k + 1], f[4 * k + 2],
                                               f[4 * k + 3]);
}

// ... global entries' values val[e] / diag[r] ...
__global__ void k_fill_gval(double *gval, long n, const double *__restrict__ val,
                            const double *__restrict__ invd) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int2 er = reinterpret_cast<const int2 *>(gval)[i];
    gval[i] = er.x < 0 ? 0.0 : val[er.x] * invd[er.y];
}

// The tile tables of the prologue and epilogue, one workgroup per tile of
// kTile slots: trow holds each slot's permuted row (-1: a padding lane) on
// entry; the tile's slots are sorted by (original row = perm[permuted row],
// padding last, then slot) and trow / tperm / tinvd receive, in that
// order, the original row, the slot's offset in the tile and 1/diag.
// Every read of trow happens before the tile's first write.
__global__ void __launch_bounds__(kTileThreads)
    k_tiles(int *trow, const int *__restrict__ perm, const double *__restrict__ invd,
            long nslots, unsigned short *tperm, double *tinvd) {
    __shared__ unsigned long long key[kTile];
    __shared__ int prow[kTile];
    const long base = (long)blockIdx.x * kTile;
    const int cnt = (int)min((long)kTile, nslots - base);
    for (int i = threadIdx.x; i < kTile; i += kTileThreads) {
        const int pr = i < cnt ? trow[base + i] : -1;
        prow[i] = pr;
        const unsigned long long o = pr >= 0 ? (unsigned)perm[pr] : 0xffffffffu;
        key[i] = o << 12 | (unsigned)i;  // slots past cnt sort last
    }
    __syncthreads();
    for (int k = 2; k <= kTile; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < kTile; i += kTileThreads) {
                const int ij = i ^ j;
                if (ij > i) {
                    const unsigned long long x = key[i], y = key[ij];
                    if ((x > y) == ((i & k) == 0)) {
                        key[i] = y;
                        key[ij] = x;
                    }
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < cnt; i += kTileThreads) {
        const unsigned long long kv = key[i];
        const int q = (int)(kv & (kTile - 1));
        const unsigned o = (unsigned)(kv >> 12);
        tperm[base + i] = (unsigned short)q;
        trow[base + i] = o == 0xffffffffu ? -1 : (int)o;
        tinvd[base + i] = prow[q] < 0 ? 0.0 : invd[prow[q]];
    }
}

}  // namespace

// ----------------------------------------------------------------- host

struct FusedPlan {
    int W = 0, NDPP = 0, NRING = 0, K = 0, nsteps = 0, nent = 0, npub = 0;
    bool blk = false;  // FIFO checked once per block (wait graph acyclic)
    unsigned epoch = 0;  // solves run; selects the slot-order x buffer
    int *d_step0 = nullptr, *d_gent0 = nullptr, *d_pub0 = nullptr;
    int4 *d_rec = nullptr;
    // slot order <-> row order, per tile of kTile slots (see the prologue)
    int *d_trow = nullptr;
    unsigned short *d_tperm = nullptr;
    double *d_tinvd = nullptr;
    int *d_gsrc = nullptr, *d_gmeta = nullptr, *d_pub = nullptr;
    double *d_gval = nullptr;
    // slot-order x: two buffers of nr_alloc right-hand sides each, and
    // which (buffer, right-hand side) regions hold the sentinel in every
    // published slot
    double *d_xq = nullptr;
    int nr_alloc = 0;
    std::vector<char> armed[2];
    double *d_brec = nullptr;  // b/diag in slot order, nr_alloc right-hand sides
    // FP32 solves (the refinement's inner solver), built on first use:
    // global entries' values and 1/diag in FP32 (the step records are
    // shared with FP64); slot-order x and b/diag for nr32 right-hand sides
    // (x refilled with the sentinel before every launch)
    float *d_gval32 = nullptr, *d_tinvd32 = nullptr;
    float *d_xq32 = nullptr, *d_brec32 = nullptr;
    int nr32 = 0;
    bool armed32 = false;  // every FP32 slot holds the sentinel
    int ncu = 0, per_cu = 0;   // CUs; k_fused workgroups resident per CU
    bool small = false;        // single-column solves are ordinary launches
    unsigned long long *d_trace = nullptr;  // RB_FUSED_TRACE=1
    double build_ms = 0.0;
    ~FusedPlan();
};

// std::vector storage whose elements are default-initialised (left
// unset for ints and doubles) when the vector is sized
template <class T>
struct NoInit : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInit<U>;
    };
    NoInit() = default;
    template <class U>
    NoInit(const NoInit<U> &) {}
    template <class U>
    void construct(U *p) {
        ::new ((void *)p) U;
    }
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
};

static int ring_class(int e) { return e == 0 ? 0 : (e <= 2 ? 2 : (e <= 4 ? 4 : 8)); }

static int words_for(int ndpp, int nring) { return words_of(ndpp, nring); }

static rb_status ensure_rhs(FusedPlan *F, int nr, hipStream_t s);
static void device_shape(FusedPlan *F);

// Can every block wait for all of its groups before its first step?  Block
// X depends on the blocks holding the rows its global entries read and on
// the previous block of its wave; the waits cannot close a cycle iff this
// graph is acyclic (Kahn's algorithm over it).
static bool wait_graph_acyclic(int nblk, const std::vector<int> &step0, int NB,
                               const std::vector<std::pair<int, int>> &edges) {
    std::vector<int> deps(nblk, 0), cnt(nblk + 1, 0);
    std::vector<char> first(nblk, 0);
    for (size_t w = 0; w + 1 < step0.size(); w++)
        if (step0[w] < step0[w + 1]) first[step0[w] / NB] = 1;
    for (const auto &e : edges) {
        if (e.first == e.second) return false;
        deps[e.first]++;
        cnt[e.second + 1]++;
    }
    for (int b = 0; b < nblk; b++) {
        if (!first[b]) {
            deps[b]++;
            cnt[b]++;  // b - 1 -> b
        }
    }
    for (int b = 0; b < nblk; b++) cnt[b + 1] += cnt[b];
    std::vector<int> rev(cnt[nblk]), fill(cnt.begin(), cnt.end() - 1);
    for (const auto &e : edges) rev[fill[e.second]++] = e.first;
    for (int b = 0; b < nblk; b++)
        if (!first[b]) rev[fill[b - 1]++] = b;
    std::vector<int> q;
    q.reserve(nblk);
    for (int b = 0; b < nblk; b++)
        if (deps[b] == 0) q.push_back(b);
    for (size_t h = 0; h < q.size(); h++)
        for (int k = cnt[q[h]]; k < cnt[q[h] + 1]; k++)
            if (--deps[rev[k]] == 0) q.push_back(rev[k]);
    return (int)q.size() == nblk;
}

template <class T>
using HostVec = std::vector<T, NoInit<T>>;

// Host partitioner and record builder.  rp/cc: permuted CSR structure
// (ascending columns); lp: BFS level boundaries.  The final permutation
// (the plan's levelItem) is fetched only once the schedule is known to be
// worth building; the values (d_cval and the plan's invd) never leave the
// device: the uploaded tables carry index pairs in their place, which the
// k_fill_* kernels turn into values.  Everything before the upload is host
// work on host data; go() is called before the first device call and
// returns false when the upload is not wanted any more.
rb_status build_fused_host(rb_plan *p, const HostVec<int> &rp,
                           const HostVec<int> &cc, const HostVec<int> &lp,
                           const double *d_cval, int Wmax, hipStream_t s,
                           const std::function<bool()> &go) {
    auto t0 = std::chrono::steady_clock::now();
    // phase times for RB_FUSED_TRACE=1
    std::vector<std::pair<const char *, double>> marks;
    auto mark = [&](const char *name) {
        marks.emplace_back(name, std::chrono::duration<double, std::milli>(
                                     std::chrono::steady_clock::now() - t0)
                                     .count());
    };
    const int m = p->m;
    const bool fwd = p->subst == 0;
    if (Wmax <= 0) {
        int maxw = 1;
        for (int L = 0; L < p->nlevel; L++) maxw = std::max(maxw, lp[L + 1] - lp[L]);
        Wmax = std::min(fused_max_waves(), (maxw + 63) / 64);
    }
    Wmax = std::max(1, std::min(Wmax, fused_max_waves()));
    auto qof = [&](int perm) { return fwd ? perm : m - 1 - perm; };
    auto pof = [&](int q) { return fwd ? q : m - 1 - q; };
    const int nlev = p->nlevel;

    // ---- wave assignment (chain following) and steps
    // every entry of these is written before it is read
    HostVec<int> wave((size_t)std::max(1, m)), stepw((size_t)std::max(1, m)),
        lanew((size_t)std::max(1, m)), par((size_t)std::max(1, m));
    // a row's parent: its latest dependency (rows split over threads)
    const int nth_h = std::max(1, std::min({16, (int)std::thread::hardware_concurrency(),
                                            (m + 65535) / 65536}));
    {
        std::vector<std::thread> pool;
        for (int t = 0; t < nth_h; t++)
            pool.emplace_back([&, t]() {
                const int q0 = (int)((long)m * t / nth_h);
                const int q1 = (int)((long)m * (t + 1) / nth_h);
                for (int q = q0; q < q1; q++) {
                    const int r = pof(q);
                    int best = -1;
                    for (int k = rp[r]; k < rp[r + 1]; k++) {
                        if (cc[k] == r) continue;
                        best = std::max(best, qof(cc[k]));
                    }
                    par[q] = best;
                }
            });
        for (auto &t : pool) t.join();
    }
    int W = 0;
    std::vector<int> cnt;           // per wave count at current level
    std::vector<int> stamp;
    std::vector<std::vector<int>> wave_rows;  // rows (q) per wave, in order
    int rr = 0;
    for (int L = 0; L < nlev; L++) {
        const int a = lp[L], b = lp[L + 1];
        const int width = b - a;
        const long per = 64L * std::max(1, Wmax);
        const int S = (int)((width + per - 1) / per);
        const int cap = 64 * std::max(1, S);
        for (int q = a; q < b; q++) {
            int pref = par[q] >= 0 ? wave[par[q]] : -1;
            int w = -1;
            if (pref >= 0 && stamp[pref] == L && cnt[pref] < cap) w = pref;
            else if (pref >= 0 && stamp[pref] != L) w = pref;
            if (w < 0) {
                if (W < Wmax) {
                    w = W++;
                    cnt.push_back(0);
                    stamp.push_back(-1);
                    wave_rows.emplace_back();
                } else {
                    for (int t = 0; t < W; t++) {
                        const int c = (rr + t) % W;
                        if (stamp[c] != L || cnt[c] < cap) {
                            w = c;
                            rr = (c + 1) % W;
                            break;
                        }
                    }
                }
            }
            if (stamp[w] != L) {
                stamp[w] = L;
                cnt[w] = 0;
            }
            cnt[w]++;
            wave[q] = w;
            wave_rows[w].push_back(q);
        }
    }
    // steps: a wave's rows of one level, 64 at a time; lane of the parent
    // when it is free
    HostVec<int> lvl((size_t)std::max(1, m));
    for (int L = 0; L < nlev; L++)
        for (int q = lp[L]; q < lp[L + 1]; q++) lvl[q] = L;
    std::vector<std::vector<std::array<int, 64>>> steps(W);
    // waves are independent here (a lane follows a parent of its own
    // wave): one thread per wave
    auto make_steps = [&](int w) {
        auto &rows = wave_rows[w];
        size_t i = 0;
        while (i < rows.size()) {
            const int L = lvl[rows[i]];
            size_t j = i;
            while (j < rows.size() && lvl[rows[j]] == L) j++;
            for (size_t c0 = i; c0 < j; c0 += 64) {
                std::array<int, 64> st;
                st.fill(-1);
                const size_t c1 = std::min(j, c0 + 64);
                std::vector<int> later;
                for (size_t k = c0; k < c1; k++) {
                    const int q = rows[k];
                    const int pq = par[q];
                    int ln = -1;
                    if (pq >= 0 && wave[pq] == w && st[lanew[pq]] < 0) ln = lanew[pq];
                    if (ln >= 0) {
                        st[ln] = q;
                        lanew[q] = ln;
                    } else {
                        later.push_back(q);
                    }
                }
                int free_l = 0;
                for (int q : later) {
                    while (st[free_l] >= 0) free_l++;
                    st[free_l] = q;
                    lanew[q] = free_l;
                }
                const int sidx = (int)steps[w].size();
                for (int l = 0; l < 64; l++)
                    if (st[l] >= 0) stepw[st[l]] = sidx;
                steps[w].push_back(st);
            }
            i = j;
        }
    };
    {
        const int nth = std::max(1, std::min({W, 16, (int)std::thread::hardware_concurrency()}));
        std::vector<std::thread> pool;
        std::atomic<int> next{0};
        for (int t = 0; t < nth; t++)
            pool.emplace_back([&]() {
                for (int w; (w = next.fetch_add(1)) < W;) make_steps(w);
            });
        for (auto &t : pool) t.join();
    }
    mark("waves+steps");
    // ---- entry classes: DPP (previous step, lane l-1 / l / l+1), ring
    // (same wave, within kR steps), global (everything else)
    auto is_dpp = [&](int q, int u) {
        return wave[u] == wave[q] && stepw[q] - stepw[u] == 1 &&
               std::abs(lanew[u] - lanew[q]) <= 1;
    };
    auto is_loc = [&](int q, int u) {
        return wave[u] == wave[q] && stepw[q] - stepw[u] <= kR;
    };
    // one pass over the entries, rows split over threads: counts of all,
    // DPP and same-wave-window entries, DPP entries by side (lane l-1 or
    // l+1 of the previous step), and the most window entries of a row
    // outside / including the DPP ones (which holds depends on the format)
    struct Tally {
        long all = 0, dpp = 0, loc = 0, lo = 0, up = 0;
        int emax_dpp = 0, emax_all = 0;
    };
    const int nth_c = std::max(1, std::min({16, (int)std::thread::hardware_concurrency(),
                                            (m + 65535) / 65536}));
    const bool may_skip = p->want_schedule == RB_SCHED_AUTO && p->tune_rhs <= 1;
    // Rows q = 0, stride, 2*stride, ... only (stride 1: every row).  A
    // sample of every 16th row settles a schedule that is clearly not
    // worth building (more than 35 % of the sampled entries cross waves)
    // without the full pass.
    auto run_tally = [&](int stride, std::vector<Tally> &tally) {
        std::vector<std::thread> pool;
        for (int t = 0; t < nth_c; t++)
            pool.emplace_back([&, t]() {
                Tally &c = tally[t];
                const int q0 = (int)((long)m * t / nth_c);
                const int q1 = (int)((long)m * (t + 1) / nth_c);
                for (int q = (q0 + stride - 1) / stride * stride; q < q1; q += stride) {
                    const int r = pof(q);
                    int k_dpp = 0, k_all = 0;
                    for (int e = rp[r]; e < rp[r + 1]; e++) {
                        if (cc[e] == r) continue;
                        const int u = qof(cc[e]);
                        const bool d = is_dpp(q, u), l = is_loc(q, u);
                        c.all++;
                        c.dpp += d;
                        c.loc += l;
                        if (d) {
                            c.lo += lanew[u] < lanew[q];
                            c.up += lanew[u] > lanew[q];
                        }
                        k_dpp += l && !d;
                        k_all += l;
                    }
                    c.emax_dpp = std::max(c.emax_dpp, k_dpp);
                    c.emax_all = std::max(c.emax_all, k_all);
                }
            });
        for (auto &t : pool) t.join();
    };
    if (may_skip && m >= (1 << 20)) {
        std::vector<Tally> sample(nth_c);
        run_tally(16, sample);
        long all = 0, loc = 0;
        for (const Tally &c : sample) {
            all += c.all;
            loc += c.loc;
        }
        if (100 * (all - loc) > 35 * all) {
            if (getenv("RB_FUSED_TRACE")) {
                mark("sampled classes");
                fprintf(stderr, "fused schedule not built: %ld of %ld sampled "
                        "entries cross waves; build (cumulative ms):", all - loc, all);
                for (auto &mk : marks) fprintf(stderr, " %s %.1f", mk.first, mk.second);
                fprintf(stderr, "\n");
            }
            return RB_OK;
        }
    }
    std::vector<Tally> tally(nth_c);
    run_tally(1, tally);
    Tally tot;
    for (const Tally &c : tally) {
        tot.all += c.all;
        tot.dpp += c.dpp;
        tot.loc += c.loc;
        tot.lo += c.lo;
        tot.up += c.up;
        tot.emax_dpp = std::max(tot.emax_dpp, c.emax_dpp);
        tot.emax_all = std::max(tot.emax_all, c.emax_all);
    }
    const long n_all = tot.all, n_dpp = tot.dpp, n_loc = tot.loc;
    // A schedule whose entries are mostly cross-wave runs at the resolver's
    // pace and loses to the blocked schedule on one right-hand side (the
    // tuner's measurements on banded and power-law factors: 3-12x slower).
    // When the caller leaves the choice to the tuner for single right-hand
    // sides, do not build it: its build is the most expensive part of the
    // plan on such factors (1.5 s of 1.7 s on the 4M-row power-law case).
    if (may_skip && 4 * (n_all - n_loc) > n_all) {
        if (getenv("RB_FUSED_TRACE")) {
            mark("classes");
            fprintf(stderr, "fused schedule not built: %ld of %ld entries cross "
                    "waves; build (cumulative ms):", n_all - n_loc, n_all);
            for (auto &mk : marks) fprintf(stderr, " %s %.1f", mk.first, mk.second);
            fprintf(stderr, "\n");
        }
        return RB_OK;
    }
    const long n_lo = tot.lo, n_up = tot.up;
    const int NDPP = !(n_dpp > 0 && 4 * n_dpp >= n_all) ? 0
                   : n_up == 0 ? 1 : n_lo == 0 ? 2 : 3;
    const int emax = NDPP ? tot.emax_dpp : tot.emax_all;
    const int NRING = ring_class(std::min(emax, 8));
    const int K = words_for(NDPP, NRING);
    mark("classes");
    // pad every wave to whole prefetch blocks with empty steps
    {
        const int Bp = block_steps(K);
        std::array<int, 64> empty;
        empty.fill(-1);
        for (int w = 0; w < W; w++)
            while (steps[w].size() % Bp) steps[w].push_back(empty);
    }
    const int kSrc = dpp_ints(NDPP), kVal = kSrc + ((NRING + 1) & ~1);
    // a global entry names its step in 24 bits
    for (int w = 0; w < W; w++)
        if (steps[w].size() >= (1u << 24)) return RB_OK;
    // ---- emit records
    std::vector<int> step0(W + 1, 0), gent0(W + 1, 0), pub0(W + 1, 0);
    for (int w = 0; w < W; w++) step0[w + 1] = step0[w] + (int)steps[w].size();
    const int NS = step0[W];
    auto slot_of = [&](int q) {
        return (step0[wave[q]] + stepw[q]) * 64 + lanew[q];
    };
    // every element of these is written by the emitting threads: no
    // serial zero-fill of ~100 MB before they start
    std::vector<int4, NoInit<int4>> rec((size_t)NS * K * 64);
    std::vector<int, NoInit<int>> srow((size_t)NS * 64);  // permuted row, -1
    const int NB = block_steps(K);
    // Per wave, in parallel on host threads (waves write disjoint record,
    // slot and 1/diag ranges; their global entries, sources and block
    // wait edges are kept per wave and joined in wave order below).
    struct WaveOut {
        std::vector<int> gsrc, gmeta, srcq;
        std::vector<int2> gval;  // (entry, row) of the value
        std::vector<std::pair<int, int>> edges;
    };
    std::vector<WaveOut> wout(W);
    auto emit_wave = [&](int w) {
        WaveOut &o = wout[w];
        std::vector<int> f(4 * std::max(K, 1));
        // a coefficient's place holds (entry, permuted row) until the
        // device computes -val[entry] / diag[row] there; entry -1: zero
        auto put = [&](int at, int e, int r) {
            f[at] = e;
            f[at + 1] = r;
        };
        for (int si = 0; si < (int)steps[w].size(); si++) {
            const auto &st = steps[w][si];
            const size_t sbase = (size_t)(step0[w] + si) * K * 64;
            for (int l = 0; l < 64; l++) {
                const int q = st[l];
                std::fill(f.begin(), f.end(), 0);
                for (int i = 0; i < kSrc; i += 2) f[i] = -1;
                for (int j = 0; j < NRING; j++) {
                    f[kSrc + j] = kZeroOff;
                    f[kVal + 2 * j] = -1;
                }
                bool hasg = false;
                if (q >= 0) {
                    const int r = pof(q);
                    int nl = 0;
                    for (int e = rp[r]; e < rp[r + 1]; e++) {
                        if (cc[e] == r) continue;
                        const int u = qof(cc[e]);
                        if (NDPP && is_dpp(q, u)) {
                            const int du = lanew[u] - l;  // -1, 0, +1
                            // lane l at [0..1]; the one side of NDPP 1 / 2
                            // at [2..3]; NDPP 3: l-1 at [2..3], l+1 at [4..5]
                            put(du == 0 ? 0 : (du < 0 || NDPP == 2) ? 2 : 4, e, r);
                        } else if (is_loc(q, u) && nl < NRING) {
                            f[kSrc + nl] = ((stepw[u] % kR) * 64 + lanew[u]) * 8;
                            put(kVal + 2 * nl, e, r);
                            nl++;
                        } else {
                            o.srcq.push_back(u);
                            o.edges.emplace_back((step0[w] + si) / NB,
                                                 slot_of(u) / 64 / NB);
                            o.gsrc.push_back(slot_of(u));
                            o.gval.push_back(make_int2(e, r));
                            o.gmeta.push_back(si << kStepShift | l << 1);
                            hasg = true;
                        }
                    }
                    if (hasg) o.gmeta.back() |= 1;  // last of the row's group
                    srow[(size_t)(step0[w] + si) * 64 + l] = r;
                } else {  // padding lane
                    srow[(size_t)(step0[w] + si) * 64 + l] = -1;
                }
                for (int k = 0; k < K; k++)
                    rec[sbase + (size_t)k * 64 + l] =
                        make_int4(f[4 * k], f[4 * k + 1], f[4 * k + 2], f[4 * k + 3]);
            }
        }
    };
    {
        const int nth = std::max(1, std::min({W, 16, (int)std::thread::hardware_concurrency()}));
        std::vector<std::thread> pool;
        std::atomic<int> next{0};
        for (int t = 0; t < nth; t++)
            pool.emplace_back([&]() {
                for (int w; (w = next.fetch_add(1)) < W;) emit_wave(w);
            });
        for (auto &t : pool) t.join();
    }
    mark("emit");
    std::vector<int> gsrc, gmeta;
    std::vector<int2> gval;
    std::vector<char> needpub(m, 0);
    // (waiting block, producing block) of every global entry; waves start
    // at multiples of NB, so block = global step / NB
    std::vector<std::pair<int, int>> wait_edges;
    for (int w = 0; w < W; w++) {
        const WaveOut &o = wout[w];
        gent0[w] = (int)gsrc.size();
        gsrc.insert(gsrc.end(), o.gsrc.begin(), o.gsrc.end());
        gval.insert(gval.end(), o.gval.begin(), o.gval.end());
        gmeta.insert(gmeta.end(), o.gmeta.begin(), o.gmeta.end());
        wait_edges.insert(wait_edges.end(), o.edges.begin(), o.edges.end());
        for (int u : o.srcq) needpub[u] = 1;
        wout[w] = WaveOut();
    }
    gent0[W] = (int)gsrc.size();
    const bool blk = wait_graph_acyclic(NS / NB, step0, NB, wait_edges);
    // rows read by other waves (or beyond the ring window) publish their x
    std::vector<std::vector<int>> pubw(W);
    for (int q = 0; q < m; q++)
        if (needpub[q]) {
            const int sl = slot_of(q);
            pubw[wave[q]].push_back(sl);
        }
    std::vector<int> pub;
    for (int w = 0; w < W; w++) {
        pub0[w] = (int)pub.size();
        pub.insert(pub.end(), pubw[w].begin(), pubw[w].end());
    }
    pub0[W] = (int)pub.size();
    if (gsrc.empty()) {
        gsrc.push_back(0);
        gval.push_back(make_int2(-1, 0));
        gmeta.push_back(1);
    }
    if (pub.empty()) pub.push_back(0);
    mark("join+pub");
    if (go && !go()) return RB_OK;
    // ---- upload
    std::unique_ptr<FusedPlan> owner(new FusedPlan());
    FusedPlan *F = owner.get();
    F->W = W;
    F->NDPP = NDPP;
    F->NRING = NRING;
    F->K = K;
    F->blk = blk;
    F->nsteps = NS;
    F->nent = gent0[W];
    F->npub = pub0[W];
    Arena &A = p->arena;
    const size_t nslots = (size_t)NS * 64;
    RB_TRY(A.alloc(&F->d_step0, W + 1));
    RB_TRY(A.alloc(&F->d_gent0, W + 1));
    RB_TRY(A.alloc(&F->d_pub0, W + 1));
    RB_TRY(A.alloc(&F->d_rec, rec.size()));
    RB_TRY(A.alloc(&F->d_trow, nslots));
    RB_TRY(A.alloc(&F->d_tperm, nslots));
    RB_TRY(A.alloc(&F->d_tinvd, nslots));
    RB_TRY(A.alloc(&F->d_gsrc, gsrc.size()));
    RB_TRY(A.alloc(&F->d_gval, gval.size()));
    RB_TRY(A.alloc(&F->d_gmeta, gmeta.size()));
    RB_TRY(A.alloc(&F->d_pub, pub.size()));
    if (getenv("RB_FUSED_TRACE")) {
        fprintf(stderr,
                "fused schedule: %d waves, %d steps, %d words/lane, dpp %d, "
                "ring %d, global entries %d, block-level FIFO checks %d\n",
                W, NS, K, NDPP, NRING, gent0[W], (int)blk);
        RB_TRY(A.alloc(&F->d_trace, (size_t)W * 16));
        RB_HIP(hipMemsetAsync(F->d_trace, 0, sizeof(unsigned long long) * W * 16, s));
    }
    auto up = [&](void *d, const void *h, size_t bytes) -> rb_status {
        RB_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
        return RB_OK;
    };
    RB_TRY(up(F->d_step0, step0.data(), sizeof(int) * step0.size()));
    RB_TRY(up(F->d_gent0, gent0.data(), sizeof(int) * gent0.size()));
    RB_TRY(up(F->d_pub0, pub0.data(), sizeof(int) * pub0.size()));
    RB_TRY(up(F->d_rec, rec.data(), sizeof(int4) * rec.size()));
    // slot -> permuted row (or -1), turned into the tile tables below
    RB_TRY(up(F->d_trow, srow.data(), sizeof(int) * srow.size()));
    RB_TRY(up(F->d_gsrc, gsrc.data(), sizeof(int) * gsrc.size()));
    RB_TRY(up(F->d_gval, gval.data(), sizeof(int2) * gval.size()));
    RB_TRY(up(F->d_gmeta, gmeta.data(), sizeof(int) * gmeta.size()));
    RB_TRY(up(F->d_pub, pub.data(), sizeof(int) * pub.size()));
    // the index pairs uploaded above become values, in place
    {
        const long nls = (long)NS * 64, nent = (long)gval.size();
        if (K > 0)
            k_fill_rec_values<<<(int)((nls + 255) / 256), 256, 0, s>>>(
                F->d_rec, nls, K, kSrc, NRING, kVal, d_cval, p->d_invd);
        k_fill_gval<<<(int)((nent + 255) / 256), 256, 0, s>>>(F->d_gval, nent,
                                                             d_cval, p->d_invd);
        k_tiles<<<(int)((nls + kTile - 1) / kTile), kTileThreads, 0, s>>>(
            F->d_trow, p->d_levelItem, p->d_invd, nls, F->d_tperm, F->d_tinvd);
        RB_HIP(hipGetLastError());
    }
    RB_HIP(hipStreamSynchronize(s));
    auto t1 = std::chrono::steady_clock::now();
    F->build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    mark("upload");
    if (getenv("RB_FUSED_TRACE")) {
        fprintf(stderr, "fused build (cumulative ms):");
        for (auto &mk : marks) fprintf(stderr, " %s %.1f", mk.first, mk.second);
        fprintf(stderr, "\n");
    }
    device_shape(F);
    delete p->fused;
    p->fused = owner.release();
    return ensure_rhs(F, 1, s);
}

// Device part of the fused build: pull the permuted CSR's structure and
// the level boundaries the host partitioner needs first.
// A fused build in flight: the host passes run on `th` while the caller
// goes on; the upload waits for `go` (1: upload, -1: abandon).
struct FusedJob {
    std::thread th;
    std::mutex mu;
    std::condition_variable cv;
    int go = 0;
    rb_status st = RB_OK;
    std::string msg;
    HostVec<int> rp, cc, lp;
};

rb_status build_fused_start(rb_plan *p, const int *d_prptr, const int *d_ccol,
                            const double *d_cval, int Wmax, hipStream_t s,
                            FusedJob **out) {
    *out = nullptr;
    const int m = p->m, nnz = p->nnz;
    auto tp = std::chrono::steady_clock::now();
    std::unique_ptr<FusedJob> job(new FusedJob());
    job->rp.resize((size_t)m + 1);
    job->cc.resize((size_t)std::max(1, nnz));
    job->lp.resize((size_t)p->nlevel + 1);
    RB_HIP(hipMemcpyAsync(job->rp.data(), d_prptr, sizeof(int) * (m + 1),
                          hipMemcpyDeviceToHost, s));
    if (nnz > 0)
        RB_HIP(hipMemcpyAsync(job->cc.data(), d_ccol, sizeof(int) * nnz,
                              hipMemcpyDeviceToHost, s));
    RB_HIP(hipMemcpyAsync(job->lp.data(), p->d_levelPtr,
                          sizeof(int) * (p->nlevel + 1), hipMemcpyDeviceToHost,
                          s));
    RB_HIP(hipStreamSynchronize(s));
    if (getenv("RB_FUSED_TRACE"))
        fprintf(stderr, "fused build: host copies of the structure %.1f ms\n",
                std::chrono::duration<double, std::milli>(
                    std::chrono::steady_clock::now() - tp)
                    .count());
    int dev = 0;
    RB_HIP(hipGetDevice(&dev));
    FusedJob *J = job.get();
    J->th = std::thread([p, J, d_cval, Wmax, s, dev]() {
        // the upload runs here, on the caller's device and stream, while
        // the caller waits in build_fused_finish
        if (hipSetDevice(dev) != hipSuccess) {
            J->st = status_of_hip(hipGetLastError(), "hipSetDevice");
            J->msg = rb_last_error();
            return;
        }
        J->st = build_fused_host(p, J->rp, J->cc, J->lp, d_cval, Wmax, s, [J]() {
            std::unique_lock<std::mutex> lk(J->mu);
            J->cv.wait(lk, [J] { return J->go != 0; });
            return J->go > 0;
        });
        if (J->st != RB_OK) J->msg = rb_last_error();
    });
    *out = job.release();
    return RB_OK;
}

rb_status build_fused_finish(FusedJob *job, bool upload) {
    {
        std::lock_guard<std::mutex> lk(job->mu);
        job->go = upload ? 1 : -1;
    }
    job->cv.notify_all();
    job->th.join();
    const rb_status st = job->st;
    const std::string msg = job->msg;
    delete job;
    if (!upload) return RB_OK;  // the caller's own error stands
    if (st != RB_OK) set_error(st, msg);
    return st;
}

int fused_max_waves() {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess)
        return 256;
    return ncu > 0 ? ncu : 256;
}

template <int NDPP, int NRING, class T>
static const void *kernel_of(bool blk, bool multi) {
    if (!blk) return (const void *)k_fused<NDPP, NRING, false, false, T>;
    return multi ? (const void *)k_fused<NDPP, NRING, true, true, T>
                 : (const void *)k_fused<NDPP, NRING, true, false, T>;
}

// the kernel for F's record format; multi: a launch of several columns
template <class T = double>
static const void *pick_kernel(const FusedPlan *F, bool multi) {
    const int r = F->NRING;
    const bool b = F->blk;
    if (F->NDPP == 3)
        return r == 0 ? kernel_of<3, 0, T>(b, multi) : r == 2 ? kernel_of<3, 2, T>(b, multi)
             : r == 4 ? kernel_of<3, 4, T>(b, multi) : kernel_of<3, 8, T>(b, multi);
    if (F->NDPP == 1)
        return r == 0 ? kernel_of<1, 0, T>(b, multi) : r == 2 ? kernel_of<1, 2, T>(b, multi)
             : r == 4 ? kernel_of<1, 4, T>(b, multi) : kernel_of<1, 8, T>(b, multi);
    if (F->NDPP == 2)
        return r == 0 ? kernel_of<2, 0, T>(b, multi) : r == 2 ? kernel_of<2, 2, T>(b, multi)
             : r == 4 ? kernel_of<2, 4, T>(b, multi) : kernel_of<2, 8, T>(b, multi);
    return r == 0 ? kernel_of<0, 0, T>(b, multi) : r == 2 ? kernel_of<0, 2, T>(b, multi)
         : r == 4 ? kernel_of<0, 4, T>(b, multi) : kernel_of<0, 8, T>(b, multi);
}

// Slot-order x buffers for nr right-hand sides (grown on demand; a fresh
// allocation holds the sentinel everywhere).
static rb_status ensure_rhs(FusedPlan *F, int nr, hipStream_t s) {
    if (nr <= F->nr_alloc) return RB_OK;
    const size_t nslots = (size_t)F->nsteps * 64;
    RB_HIP(hipStreamSynchronize(s));
    if (F->d_xq) RB_HIP(hipFree(F->d_xq));
    if (F->d_brec) RB_HIP(hipFree(F->d_brec));
    F->d_xq = nullptr;
    F->d_brec = nullptr;
    F->nr_alloc = 0;
    if (hipMalloc(&F->d_xq, 2 * nr * nslots * sizeof(double)) != hipSuccess ||
        hipMalloc(&F->d_brec, nr * nslots * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        set_error(RB_ERR_NO_MEMORY, "hipMalloc failed (slot-order x buffers)");
        return RB_ERR_NO_MEMORY;
    }
    const long n = 2L * nr * (long)nslots;
    k_fill_pending<double><<<(int)((n + 255) / 256), 256, 0, s>>>(F->d_xq, n, nullptr);
    RB_HIP(hipGetLastError());
    F->nr_alloc = nr;
    F->armed[0].assign(nr, 1);
    F->armed[1].assign(nr, 1);
    return RB_OK;
}

// Launch form.  The consumers wait on each other, so all W * nr workgroups
// of a launch must be resident together.  A cooperative launch guarantees
// that, but costs about 18 us more per launch than an ordinary one on MI355X
// (single-column Laplacian 1500^2 solve: 0.291 ms cooperative, 0.273 ms
// ordinary).  Launches of at most ncu / 8 workgroups -- single-column solves
// of plans with few waves -- are ordinary launches: such a grid is resident
// once dispatched as long as the device's other resident work leaves it that
// many workgroup slots.  The cooperative (multi-column) launches size
// themselves to leave those slots to up to kSmallStreams other live plans
// whose solves may be ordinary launches on other streams.  Every wait in the
// kernel is bounded all the same.
constexpr int kSmallStreams = 3;
static std::atomic<int> g_small_plans{0};

static int small_grid(int ncu) { return std::max(1, ncu / 8); }

static void device_shape(FusedPlan *F) {
    if (F->ncu > 0) return;
    int dev = 0, ncu = 0, per = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, pick_kernel(F, true),
                                                     kWG, 0) != hipSuccess)
        per = 1;
    (void)hipGetLastError();
    F->ncu = std::max(1, ncu);
    F->per_cu = std::max(1, per);
    F->small = F->W <= small_grid(F->ncu);
    if (F->small) g_small_plans.fetch_add(1);
}

FusedPlan::~FusedPlan() {
    if (small) g_small_plans.fetch_sub(1);
    for (void *q : {(void *)d_xq, (void *)d_brec, (void *)d_gval32,
                    (void *)d_tinvd32, (void *)d_xq32, (void *)d_brec32})
        if (q) (void)hipFree(q);
}

// Right-hand sides that fit one cooperative launch (W workgroups each).
static int rhs_per_launch(FusedPlan *F) {
    device_shape(F);
    const int others = std::max(
        0, std::min(kSmallStreams, g_small_plans.load() - (F->small ? 1 : 0)));
    const int slots = F->per_cu * F->ncu - others * small_grid(F->ncu);
    return std::max(1, std::min(64, slots / std::max(1, F->W)));
}

// One launch group: nr right-hand sides (b, x with leading dimensions ldb,
// ldx) through prologue, k_fused and epilogue.  cur: nr slot-order x
// buffers whose published slots hold the sentinel; nxt: the buffers of the
// next solve, whose published slots the launch re-arms (may be null); brec:
// nr slot-order b/diag buffers.  Buffers of one right-hand side are nslots
// apart.
// T: FP64, or FP32 on the FP32 tables for the refinement (left: see
// k_fused): then b is the FP64 residual and x (FP64) receives x += the
// correction in its not converged columns (done) and cur is refilled with
// the sentinel.
template <class T>
static rb_status launch_group(rb_plan *p, const double *b, int ldb, double *x,
                              int ldx, int nr, T *cur, T *nxt, T *brec,
                              hipStream_t s, const int *left = nullptr,
                              const int *done = nullptr) {
    FusedPlan *F = p->fused;
    constexpr bool f32 = sizeof(T) == 4;
    const size_t nslots = (size_t)F->nsteps * 64;
    const dim3 tiles((unsigned)((nslots + kTile - 1) / kTile), (unsigned)nr);
    const T *tinvd, *gval;
    if constexpr (f32) {
        tinvd = F->d_tinvd32;
        gval = F->d_gval32;
    } else {
        tinvd = F->d_tinvd;
        gval = F->d_gval;
    }
    k_fused_prologue<T, double><<<tiles, kTileThreads, 0, s>>>(
        F->d_trow, F->d_tperm, tinvd, (long)nslots, b, ldb, brec, left);
    RB_HIP(hipGetLastError());
    FusedArgsT<T> args{F->d_step0, F->d_gent0, F->d_pub0, F->d_rec,
                       F->d_gsrc, gval, F->d_gmeta, F->d_pub, cur, nxt,
                       brec, (long)nslots, F->W, p->d_status,
                       p->spin_limit, F->d_trace, left};
    void *kargs[] = {(void *)&args};
    mark_start(p, s);
    const bool ordinary = F->W * nr <= small_grid(F->ncu);
    const void *kern = pick_kernel<T>(F, nr > 1);
    hipError_t e =
        ordinary
            ? hipLaunchKernel(kern, dim3(F->W * nr), dim3(kWG), kargs, 0, s)
            : hipLaunchCooperativeKernel(kern, dim3(F->W * nr), dim3(kWG),
                                         kargs, 0, s);
    if (e != hipSuccess)
        return status_of_hip(e, ordinary ? "hipLaunchKernel(k_fused)"
                                         : "hipLaunchCooperativeKernel(k_fused)");
    mark_end(p, s);
    if constexpr (f32)
        k_fused_epilogue_ir<<<tiles, kTileThreads, 0, s>>>(
            F->d_trow, F->d_tperm, (long)nslots, cur, x, ldx, p->d_status, done,
            left);
    else
        k_fused_epilogue<T><<<tiles, kTileThreads, 0, s>>>(
            F->d_trow, F->d_tperm, (long)nslots, cur, x, ldx, p->d_status, left);
    RB_HIP(hipGetLastError());
    return RB_OK;
}

rb_status solve_fused(rb_plan *p, const double *d_b, int ldb, double *d_x,
                      int ldx, int nrhs, hipStream_t s) {
    FusedPlan *F = p->fused;
    if (!F) {
        set_error(RB_ERR_INTERNAL, "fused schedule not built");
        return RB_ERR_INTERNAL;
    }
    if (p->m == 0) return RB_OK;
    const size_t nslots = (size_t)F->nsteps * 64;
    device_shape(F);
    const int cap = nrhs > 1 ? rhs_per_launch(F) : 1;
    RB_TRY(ensure_rhs(F, std::min(nrhs, cap), s));
    for (int k0 = 0; k0 < nrhs; k0 += cap) {
        const int nr = std::min(cap, nrhs - k0);
        const int cb = F->epoch & 1, nb = cb ^ 1;
        double *cur = F->d_xq + (size_t)cb * F->nr_alloc * nslots;
        double *nxt = F->d_xq + (size_t)nb * F->nr_alloc * nslots;
        // regions this launch polls must hold the sentinel
        for (int r = 0; r < nr; r++)
            if (!F->armed[cb][r]) {
                k_fill_pending<double><<<(int)((nslots + 255) / 256), 256, 0, s>>>(
                    cur + r * nslots, (long)nslots, nullptr);
                RB_HIP(hipGetLastError());
            }
        RB_TRY(launch_group<double>(p, d_b + (size_t)k0 * ldb, ldb,
                                    d_x + (size_t)k0 * ldx, ldx, nr, cur, nxt,
                                    F->d_brec, s));
        for (int r = 0; r < nr; r++) {
            F->armed[cb][r] = 0;
            F->armed[nb][r] = 1;
        }
        F->epoch++;
    }
    return RB_OK;
}

size_t fused_slots(const rb_plan *p) {
    return p->fused ? (size_t)p->fused->nsteps * 64 : 0;
}

rb_status solve_fused_ws(rb_plan *p, const double *d_b, int ldb, double *d_x,
                         int ldx, int nrhs, hipStream_t s, double *ws,
                         size_t ws_bytes) {
    FusedPlan *F = p->fused;
    if (!F) {
        set_error(RB_ERR_INTERNAL, "fused schedule not built");
        return RB_ERR_INTERNAL;
    }
    if (p->m == 0) return RB_OK;
    const size_t nslots = (size_t)F->nsteps * 64;
    const size_t per_rhs = 2 * nslots * sizeof(double);
    device_shape(F);
    const int fit = (int)std::min<size_t>(64, ws_bytes / per_rhs);
    if (fit < 1) {
        set_error(RB_ERR_INVALID_ARG, "workspace too small for one column");
        return RB_ERR_INVALID_ARG;
    }
    const int cap = std::min(fit, nrhs > 1 ? rhs_per_launch(F) : 1);
    for (int k0 = 0; k0 < nrhs; k0 += cap) {
        const int nr = std::min(cap, nrhs - k0);
        double *cur = ws;
        double *brec = ws + (size_t)nr * nslots;
        // the workspace keeps whatever the last call left: every slot of
        // every column goes back to the sentinel, and nothing is re-armed
        const long n = (long)nr * (long)nslots;
        k_fill_pending<double><<<(int)((n + 255) / 256), 256, 0, s>>>(cur, n, nullptr);
        RB_HIP(hipGetLastError());
        RB_TRY(launch_group<double>(p, d_b + (size_t)k0 * ldb, ldb,
                                    d_x + (size_t)k0 * ldx, ldx, nr, cur, nullptr,
                                    brec, s));
    }
    return RB_OK;
}

// FP32 tables (once) and buffers for nr right-hand sides.
static rb_status ensure32(rb_plan *p, int nr, hipStream_t s) {
    FusedPlan *F = p->fused;
    const size_t nslots = (size_t)F->nsteps * 64;
    if (!F->d_gval32) {
        const size_t nent = std::max(1, F->nent);
        if (hipMalloc(&F->d_gval32, sizeof(float) * nent) != hipSuccess ||
            hipMalloc(&F->d_tinvd32, sizeof(float) * nslots) != hipSuccess) {
            (void)hipGetLastError();
            if (F->d_gval32) (void)hipFree(F->d_gval32);
            F->d_gval32 = nullptr;  // built on the next call
            set_error(RB_ERR_NO_MEMORY, "hipMalloc failed (FP32 fused tables)");
            return RB_ERR_NO_MEMORY;
        }
        k_to32<<<(int)((nent + 255) / 256), 256, 0, s>>>(F->d_gval, F->d_gval32,
                                                        (long)nent);
        k_to32<<<(int)((nslots + 255) / 256), 256, 0, s>>>(F->d_tinvd, F->d_tinvd32,
                                                          (long)nslots);
        RB_HIP(hipGetLastError());
    }
    if (nr > F->nr32) {
        RB_HIP(hipStreamSynchronize(s));
        if (F->d_xq32) RB_HIP(hipFree(F->d_xq32));
        if (F->d_brec32) RB_HIP(hipFree(F->d_brec32));
        F->d_xq32 = F->d_brec32 = nullptr;
        F->nr32 = 0;
        if (hipMalloc(&F->d_xq32, sizeof(float) * nr * nslots) != hipSuccess ||
            hipMalloc(&F->d_brec32, sizeof(float) * nr * nslots) != hipSuccess) {
            (void)hipGetLastError();
            set_error(RB_ERR_NO_MEMORY, "hipMalloc failed (FP32 slot-order buffers)");
            return RB_ERR_NO_MEMORY;
        }
        F->nr32 = nr;
        F->armed32 = false;
    }
    if (!F->armed32) {  // every slot to the sentinel; the epilogues keep it so
        const long n = (long)F->nr32 * (long)nslots;
        k_fill_pending<float><<<(int)((n + 255) / 256), 256, 0, s>>>(F->d_xq32, n,
                                                                    nullptr);
        RB_HIP(hipGetLastError());
        F->armed32 = true;
    }
    return RB_OK;
}

rb_status solve_fused32(rb_plan *p, const double *d_r, int ldr, double *d_x,
                        int ldx, int nrhs, hipStream_t s, const int *left,
                        const int *done) {
    FusedPlan *F = p->fused;
    if (!F) {
        set_error(RB_ERR_INTERNAL, "fused schedule not built");
        return RB_ERR_INTERNAL;
    }
    if (p->m == 0) return RB_OK;
    device_shape(F);
    const int cap = nrhs > 1 ? rhs_per_launch(F) : 1;
    RB_TRY(ensure32(p, std::min(nrhs, cap), s));
    for (int k0 = 0; k0 < nrhs; k0 += cap) {
        const int nr = std::min(cap, nrhs - k0);
        const rb_status st = launch_group<float>(
            p, d_r + (size_t)k0 * ldr, ldr, d_x + (size_t)k0 * ldx, ldx, nr,
            F->d_xq32, nullptr, F->d_brec32, s, left, done + k0);
        if (st != RB_OK) {
            F->armed32 = false;  // an epilogue may not have run
            return st;
        }
    }
    return RB_OK;
}

void fused_rearm(rb_plan *p) {
    FusedPlan *F = p->fused;
    if (!F) return;
    for (int b = 0; b < 2; b++)
        for (size_t r = 0; r < F->armed[b].size(); r++) F->armed[b][r] = 0;
}

void fused_destroy(rb_plan *p) {
    delete p->fused;
    p->fused = nullptr;
}

}  // namespace rb

extern "C" rb_status rb_plan_trace(const rb_plan *p, unsigned long long *counters,
                                   int *nwaves) {
    using namespace rb;
    if (!p || !nwaves) {
        set_error(RB_ERR_INVALID_ARG, "null argument");
        return RB_ERR_INVALID_ARG;
    }
    *nwaves = p->fused ? p->fused->W : 0;
    if (!p->fused || !p->fused->d_trace) return RB_OK;
    if (counters) {
        RB_HIP(hipDeviceSynchronize());
        RB_HIP(hipMemcpy(counters, p->fused->d_trace,
                         sizeof(unsigned long long) * 16 * (size_t)p->fused->W,
                         hipMemcpyDeviceToHost));
    }
    return RB_OK;
}
