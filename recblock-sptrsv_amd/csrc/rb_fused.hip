// Fused schedule: the whole solve in one persistent launch.
//
// Rows are grouped into "chains" that each belong to one consumer wave.  A
// wave solves its rows in level order, one *step* (at most 64 rows of one
// level, one per lane) at a time.  A dependency on a row solved by the same
// wave within the last R steps is read from the wave's LDS ring of recent x
// values; every other dependency is a *global* one: it is read from the x
// array in solve order, where each slot holds a sentinel NaN until solved.
//
// Assignment (host, at plan build): levels are walked in order; a row joins
// the wave of its BFS parent (the dependency solved last) while that wave
// has room at this level, otherwise a new wave (up to W) or the least
// loaded one.  On grid-like matrices this keeps every row chain inside one
// wave, so the per-level hop is an LDS round trip instead of a trip through
// the memory hierarchy.
//
// One workgroup per consumer wave, plus helper waves:
//   feeders   copy the step records (row ids, 1/diag, LDS-class entries, b)
//             from HBM into an LDS ring with asynchronous LDS-DMA loads
//             (global_load_lds_dwordx4), several steps ahead;
//   resolver  polls the global dependencies in order, many loads in flight,
//             and hands the consumer one partial sum per (step, lane) through
//             an LDS FIFO;
//   consumer  x = (b - fifo - sum(val * xring[slot])) / diag, writes x to the
//             LDS ring, to the solve-order x array (for other waves) and to
//             the caller's x.
// Every wait is bounded: on a timeout the kernel sets a status flag and
// releases the waiting rows with NaN instead of hanging.
#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <numeric>
#include <vector>

#include "rb_internal.h"

namespace rb {

namespace {

typedef __attribute__((address_space(3))) void lds_void;

constexpr int kR = 32;          // LDS x ring depth (steps)
constexpr int kFF = 512;        // FIFO capacity (groups)
constexpr int kU = 4;           // resolver batches of 64 polled together
constexpr int kH = 4;           // steps per feeder chunk
constexpr unsigned kSpin = 1u << 22;
constexpr int kZeroSlot = kR * 64;

// s_waitcnt immediate that waits for vmcnt <= n only (gfx9 encoding)
constexpr int vm_imm(int n) {
    return (n & 0xF) | ((n >> 4) << 14) | (0x7 << 4) | (0xF << 8);
}

// record layout: ops of 16 bytes per lane
//   op0           {int q, int orig, double invd}
//   op1..         {int hg, int s0, int s1, int s2} {s3..s6} {s7, ...}
//   vals          {v0, v1} {v2, v3} ...
//   b (own array) {double b, 0}
template <int EC>
struct Layout {
    static constexpr int kSrcOps = EC <= 3 ? 1 : (EC <= 7 ? 2 : 3);
    static constexpr int kValOps = (EC + 1) / 2;
    static constexpr int kOps = 1 + kSrcOps + kValOps;  // without b
    static constexpr int kP = EC <= 2 ? 16 : 8;          // ring steps
    static constexpr int kNF = kP / kH;                  // feeder waves
};

__device__ inline int lds_load(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
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
__device__ inline void st_x(double *p, double v) {
    __hip_atomic_store((unsigned long long *)p,
                       (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool pending(double v) {
    return __double_as_longlong(v) == (long long)kPending;
}

struct FusedArgs {
    const int *step0;   // [W+1] first step of each wave
    const int *gent0;   // [W+1] first global entry of each wave
    const int4 *rec;    // [nsteps * ops * 64]
    const double2 *brec;// [nsteps * 64]
    const int *gsrc;    // [nent] solve-order index of the dependency
    const double *gval; // [nent]
    const int *gmeta;   // [nent] group ordinal * 2 + last-of-group
    double *xq;         // [m] x in solve order (sentinel until solved)
    double *x;          // caller's x (original order)
    int *status;
    // optional trace: per workgroup 16 counters of wall-clock ticks
    // (10 ns): consumer wait-record, wait-fifo, total, steps; feeder
    // wait-space, wait-dma, chunks; resolver wait-space, poll, batches,
    // rounds
    unsigned long long *trace;
};

__device__ inline unsigned long long now() { return wall_clock64(); }

// source codes of a record entry (non-negative: LDS ring slot)
constexpr int kPrevSame = -1;   // previous step, same lane (register)
constexpr int kPrevLower = -2;  // previous step, lane - 1 (DPP)
constexpr int kPrevUpper = -3;  // previous step, lane + 1 (DPP)
constexpr int kPad = -4;        // padding, contributes nothing
constexpr int kPublish = 1 << 30;  // hg flag: other waves read this x

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

// One lane's step record, read from an LDS ring slot (ops of 16 bytes).
template <int EC>
struct Rec {
    int q, orig, hg;
    double invd, b;
    int src[EC];
    double val[EC];
    __device__ void load(const int4 (*slot)[64], int lane) {
        using L = Layout<EC>;
        const int4 o0 = slot[0][lane];
        q = o0.x;
        orig = o0.y;
        invd = __hiloint2double(o0.w, o0.z);
#pragma unroll
        for (int k = 0; k < L::kSrcOps; k++) {
            const int4 o = slot[1 + k][lane];
            const int vv[4] = {o.x, o.y, o.z, o.w};
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int idx = k * 4 + t - 1;
                if (k == 0 && t == 0) hg = vv[0];
                else if (idx < EC) src[idx] = vv[t];
            }
        }
#pragma unroll
        for (int k = 0; k < L::kValOps; k++) {
            const int4 o = slot[1 + L::kSrcOps + k][lane];
            val[2 * k] = __hiloint2double(o.y, o.x);
            if (2 * k + 1 < EC) val[2 * k + 1] = __hiloint2double(o.w, o.z);
        }
        const int4 ob = slot[L::kOps][lane];
        b = __hiloint2double(ob.y, ob.x);
    }
};

struct Shared {
    int rec_ready[64];  // per feeder chunk: chunk index + 1 once loaded
    int cons;           // steps consumed
    int fifo_ready;     // groups resolved
    int fifo_cons;      // groups consumed
    int abort;
};

template <int EC>
__global__ void __launch_bounds__(64 * (2 + Layout<EC>::kNF))
k_fused(FusedArgs a) {
    using L = Layout<EC>;
    constexpr int P = L::kP, NF = L::kNF, K = L::kOps;
    __shared__ int4 ring[P][K + 1][64];   // +1: b
    __shared__ double xring[kR * 64 + 1];
    __shared__ double fifo[kFF];
    __shared__ Shared sh;
    __shared__ double rtmp[64];
    __shared__ int gtmp[64];
    __shared__ double ctmp;

    const int w = blockIdx.x;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int s0 = a.step0[w];
    const int ns = a.step0[w + 1] - s0;

    if (threadIdx.x == 0) {
        for (int i = 0; i < 64; i++) sh.rec_ready[i] = 0;
        sh.cons = 0;
        sh.fifo_ready = 0;
        sh.fifo_cons = 0;
        sh.abort = 0;
        xring[kZeroSlot] = 0.0;
    }
    __syncthreads();

    if (wave == 0) {
        // ------------------------------------------------------- consumer
        int fifo_pos = 0;
        int rec_upto = 0;     // steps whose records are known to be in LDS
        int fifo_avail = 0;   // groups known to be resolved
        double xprev = 0.0;   // this lane's x of the previous step
        unsigned long long t_rec = 0, t_fifo = 0, t_start = now();
        // wait until the records of step s are in the ring (checks whole
        // chunks, so most steps pay no LDS poll)
        auto wait_rec = [&](int s) -> bool {
            if (s < rec_upto) return true;
            unsigned long long tw = a.trace ? now() : 0;
            unsigned spins = 0;
            while (true) {
                const int c = rec_upto / kH;
                if (lds_load(&sh.rec_ready[c % 64]) == c + 1) {
                    rec_upto += kH;
                    if (s < rec_upto) break;
                    continue;
                }
                if (++spins > kSpin || lds_load(&sh.abort)) return false;
                __builtin_amdgcn_s_sleep(1);
            }
            wg_acquire();
            if (a.trace) t_rec += now() - tw;
            return true;
        };
        Rec<EC> cur, nxt;
        bool ok = ns > 0 ? wait_rec(0) : true;
        if (ns > 0 && ok) cur.load(ring[0], lane);
        for (int s = 0; s < ns && ok; s++) {
            // prefetch the next step's record while this one computes
            const bool pre = s + 1 < ns && s + 1 < rec_upto;
            if (pre) nxt.load(ring[(s + 1) % P], lane);
            const int has = cur.hg & 1;
            const int gofs = (cur.hg >> 1) & 0x7fff;
            const int gtot = __shfl(gofs + has, 63, 64);
            double acc = 0.0;
            if (gtot > 0) {
                if (fifo_avail < fifo_pos + gtot) {
                    unsigned long long tw = a.trace ? now() : 0;
                    unsigned spins = 0;
                    while ((fifo_avail = lds_load(&sh.fifo_ready)) <
                           fifo_pos + gtot) {
                        if (++spins > kSpin || lds_load(&sh.abort)) break;
                        __builtin_amdgcn_s_sleep(1);
                    }
                    if (spins > kSpin || lds_load(&sh.abort)) {
                        ok = false;
                        break;
                    }
                    wg_acquire();
                    if (a.trace) t_fifo += now() - tw;
                }
                if (has) acc = fifo[(fifo_pos + gofs) % kFF];
            }
            // previous step's x of this lane and of its neighbours
            const double xl = dpp_from_lower(xprev), xr = dpp_from_upper(xprev);
#pragma unroll
            for (int j = 0; j < EC; j++) {
                const int sj = cur.src[j];
                double xv;
                if (sj >= 0) xv = xring[sj];
                else xv = sj == kPrevSame ? xprev
                        : sj == kPrevLower ? xl
                        : sj == kPrevUpper ? xr : 0.0;
                acc += cur.val[j] * xv;
            }
            const double xv = (cur.b - acc) * cur.invd;
            xprev = xv;
            xring[(s % kR) * 64 + lane] = xv;
            if (cur.q >= 0) {
                if (cur.hg & kPublish) st_x(&a.xq[cur.q], xv);
                a.x[cur.orig] = xv;
            }
            fifo_pos += gtot;
            wg_release();
            if (lane == 0) {
                lds_store(&sh.fifo_cons, fifo_pos);
                lds_store(&sh.cons, s + 1);
            }
            if (s + 1 < ns) {
                if (!pre) {
                    if (!wait_rec(s + 1)) {
                        ok = false;
                        break;
                    }
                    nxt.load(ring[(s + 1) % P], lane);
                }
                cur = nxt;
            }
        }
        if (!ok && lane == 0) {
            atomicOr(a.status, 2);
            lds_store(&sh.abort, 1);
        }
        if (lane == 0) lds_store(&sh.cons, 1 << 30);
        if (a.trace && lane == 0) {
            unsigned long long *T = a.trace + (size_t)w * 16;
            T[0] = t_rec;
            T[1] = t_fifo;
            T[2] = now() - t_start;
            T[3] = ns;
        }
    } else if (wave <= NF) {
        // ------------------------------------------------------- feeders
        const int f = wave - 1;
        const int4 *rec = a.rec + (size_t)s0 * K * 64;
        const double2 *brec = a.brec + (size_t)s0 * 64;
        const int nchunks = (ns + kH - 1) / kH;
        unsigned long long t_space = 0, t_dma = 0, tw = 0;
        for (int c = f; c < nchunks; c += NF) {
            const int first = c * kH;
            if (a.trace) tw = now();
            // the ring holds P steps: wait until steps < first + kH - P are
            // consumed
            unsigned spins = 0;
            while (lds_load(&sh.cons) < first + kH - P) {
                if (++spins > kSpin || lds_load(&sh.abort)) break;
                __builtin_amdgcn_s_sleep(1);
            }
            if (spins > kSpin || lds_load(&sh.abort)) break;
            if (a.trace) {
                t_space += now() - tw;
                tw = now();
            }
            for (int t = 0; t < kH; t++) {
                const int s = first + t;
                if (s >= ns) break;
                const int slot = s % P;
#pragma unroll
                for (int k = 0; k < K; k++)
                    __builtin_amdgcn_global_load_lds(
                        (const void *)(rec + ((size_t)s * K + k) * 64 + lane),
                        (lds_void *)&ring[slot][k][0], 16, 0, 0);
                __builtin_amdgcn_global_load_lds(
                    (const void *)(brec + (size_t)s * 64 + lane),
                    (lds_void *)&ring[slot][K][0], 16, 0, 0);
            }
            __builtin_amdgcn_s_waitcnt(vm_imm(0));
            if (a.trace) t_dma += now() - tw;
            wg_release();
            if (lane == 0) lds_store(&sh.rec_ready[c % 64], c + 1);
        }
        if (a.trace && lane == 0 && f == 0) {
            unsigned long long *T = a.trace + (size_t)w * 16;
            T[4] = t_space;
            T[5] = t_dma;
            T[6] = nchunks;
        }
    } else {
        // ------------------------------------------------------- resolver
        const int e0 = a.gent0[w], e1 = a.gent0[w + 1];
        double carry = 0.0;
        int carry_g = -1;
        int done_groups = 0;
        unsigned long long t_rspace = 0, t_poll = 0, nb = 0, nround = 0, tw = 0;
        for (int kb = e0; kb < e1; kb += 64 * kU) {
            // metadata and first polls of kU batches of 64 entries at once,
            // so kU * 64 loads are in flight per round trip
            if (a.trace) tw = now();
            int srcu[kU], gmu[kU];
            double gvu[kU], xvu[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const int idx = kb + 64 * u + lane;
                srcu[u] = 0;
                gmu[u] = 0;
                gvu[u] = 0.0;
                if (idx < e1) {
                    srcu[u] = a.gsrc[idx];
                    gvu[u] = a.gval[idx];
                    gmu[u] = a.gmeta[idx];
                }
            }
            {
                // FIFO space for every group these batches may complete
                const int nt = min(64 * kU, e1 - kb);
                const int gmax = a.gmeta[kb + nt - 1] >> 1;
                unsigned spins = 0;
                while (gmax - lds_load(&sh.fifo_cons) >= kFF) {
                    if (++spins > kSpin || lds_load(&sh.abort)) break;
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            if (a.trace) {
                t_rspace += now() - tw;
                tw = now();
            }
#pragma unroll
            for (int u = 0; u < kU; u++)
                xvu[u] = kb + 64 * u + lane < e1 ? ld_x(&a.xq[srcu[u]]) : 0.0;
#pragma unroll
            for (int u = 0; u < kU; u++) {
            const int k0 = kb + 64 * u;
            if (k0 >= e1) break;
            nb++;
            const int n = min(64, e1 - k0);
            const bool act = lane < n;
            const int src = srcu[u], gm = gmu[u];
            const double gv = gvu[u];
            double xv = xvu[u];
            const int g = gm >> 1;
            unsigned spins = 0;
            bool pend = act && pending(xv);
            rtmp[lane] = act && !pend ? gv * xv : 0.0;
            gtmp[lane] = act ? gm : -2;
            const int gprev = __shfl_up(g, 1, 64);
            // Publish progressively: groups whose entries lie in the ready
            // prefix of the batch go out at once, so a pending entry of a
            // later step never holds back an earlier step (that could close
            // a wait cycle between waves).
            int done_lo = 0;  // lanes below are summed and published
            while (true) {
                const unsigned long long pm = __ballot(pend);
                const int fp = pm ? (int)__ffsll((long long)pm) - 1 : n;
                int hi;  // sum lanes [done_lo, hi)
                if (fp >= n) {
                    hi = n;
                } else {
                    const unsigned long long em = __ballot(
                        act && lane >= done_lo && lane < fp && (gtmp[lane] & 1));
                    hi = em ? 64 - __clzll((long long)em) : done_lo;
                }
                if (hi > done_lo) {
                    // group heads sum their entries in order; a group begun
                    // in an earlier batch starts from the carried sum
                    const bool head = act && lane >= done_lo && lane < hi &&
                                      (lane == done_lo || gprev != g);
                    if (head) {
                        double sum = (lane == 0 && carry_g == g) ? carry : 0.0;
                        int l = lane;
                        while (true) {
                            sum += rtmp[l];
                            if (l + 1 >= hi || (gtmp[l + 1] >> 1) != g) break;
                            l++;
                        }
                        if (gtmp[l] & 1) fifo[g % kFF] = sum;  // complete
                        else ctmp = sum;  // only the batch's last group
                    }
                    const int gl = __shfl(gm, hi - 1, 64);
                    if (gl & 1) {
                        done_groups = (gl >> 1) + 1;
                    } else {  // hi == n: the last group continues
                        carry = ctmp;
                        done_groups = gl >> 1;
                    }
                    carry_g = (gl & 1) ? -1 : (gl >> 1);
                    wg_release();
                    if (lane == 0) lds_store(&sh.fifo_ready, done_groups);
                    done_lo = hi;
                }
                if (fp >= n) break;
                nround++;
                if (++spins > kSpin || lds_load(&sh.abort)) {
                    if (lane == 0) {
                        atomicOr(a.status, 4);
                        lds_store(&sh.abort, 1);
                    }
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                if (pend) {
                    xv = ld_x(&a.xq[src]);
                    pend = pending(xv);
                    if (!pend) rtmp[lane] = gv * xv;
                }
            }
            if (lds_load(&sh.abort)) break;
            }
            if (a.trace) t_poll += now() - tw;
            if (lds_load(&sh.abort)) break;
        }
        if (a.trace && lane == 0) {
            unsigned long long *T = a.trace + (size_t)w * 16;
            T[7] = t_rspace;
            T[8] = t_poll;
            T[9] = nb;
            T[10] = nround;
        }
    }
}

__global__ void k_fused_prologue(const int *__restrict__ orig, long nslots,
                                 const double *__restrict__ b,
                                 double2 *brec, double *xq, int m) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nslots) {
        const int o = orig[i];
        brec[i] = make_double2(o >= 0 ? b[o] : 0.0, 0.0);
    }
    if (i < m) ((unsigned long long *)xq)[i] = kPending;
}

}  // namespace

// ----------------------------------------------------------------- host

struct FusedPlan {
    int W = 0, EC = 0, nsteps = 0, nent = 0;
    int *d_step0 = nullptr, *d_gent0 = nullptr, *d_orig = nullptr;
    int4 *d_rec = nullptr;
    double2 *d_brec = nullptr;
    int *d_gsrc = nullptr, *d_gmeta = nullptr;
    double *d_gval = nullptr;
    double *d_xq = nullptr;
    unsigned long long *d_trace = nullptr;  // RB_FUSED_TRACE=1
    double build_ms = 0.0;
};

static int ec_class(int e) { return e <= 2 ? 2 : (e <= 4 ? 4 : 8); }

static int ops_for(int ec) {
    switch (ec) {
    case 2: return Layout<2>::kOps;
    case 4: return Layout<4>::kOps;
    default: return Layout<8>::kOps;
    }
}

// Host partitioner and record builder.  rp/cc/cv: permuted CSR (ascending
// columns); lp: BFS level boundaries; lf: final permutation; invd: 1/diag.
rb_status build_fused_host(rb_plan *p, const std::vector<int> &rp,
                           const std::vector<int> &cc,
                           const std::vector<double> &cv,
                           const std::vector<int> &lp,
                           const std::vector<int> &lf,
                           const std::vector<double> &invd, int Wmax,
                           hipStream_t s) {
    auto t0 = std::chrono::steady_clock::now();
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
    std::vector<int> wave(m), stepw(m), lanew(m);
    std::vector<int> par(m, -1);
    for (int q = 0; q < m; q++) {
        const int r = pof(q);
        int best = -1;
        for (int k = rp[r]; k < rp[r + 1]; k++) {
            if (cc[k] == r) continue;
            best = std::max(best, qof(cc[k]));
        }
        par[q] = best;
    }
    int W = 0;
    std::vector<int> cnt;           // per wave count at current level
    std::vector<int> stamp;
    std::vector<int> nsteps_w;      // steps used so far per wave
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
                    nsteps_w.push_back(0);
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
    std::vector<int> lvl(m);
    for (int L = 0; L < nlev; L++)
        for (int q = lp[L]; q < lp[L + 1]; q++) lvl[q] = L;
    std::vector<std::vector<std::array<int, 64>>> steps(W);
    for (int w = 0; w < W; w++) {
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
    }
    // ---- classify entries, choose EC
    int emax = 0;
    std::vector<int> nloc(m, 0);
    for (int q = 0; q < m; q++) {
        const int r = pof(q);
        int k = 0;
        for (int e = rp[r]; e < rp[r + 1]; e++) {
            if (cc[e] == r) continue;
            const int u = qof(cc[e]);
            if (wave[u] == wave[q] && stepw[q] - stepw[u] <= kR) k++;
        }
        nloc[q] = k;
        emax = std::max(emax, k);
    }
    const int EC = ec_class(std::min(emax, 8));
    const int K = ops_for(EC);
    const int KS = EC <= 3 ? 1 : (EC <= 7 ? 2 : 3);
    // ---- emit records
    std::vector<int> step0(W + 1, 0), gent0(W + 1, 0);
    for (int w = 0; w < W; w++) step0[w + 1] = step0[w] + (int)steps[w].size();
    const int NS = step0[W];
    std::vector<int4> rec((size_t)NS * K * 64);
    std::vector<int> origv((size_t)NS * 64, -1);
    std::vector<int> gsrc, gmeta;
    std::vector<double> gval;
    std::vector<char> needpub(m, 0);
    for (int w = 0; w < W; w++) {
        gent0[w] = (int)gsrc.size();
        int gord = 0;
        for (int si = 0; si < (int)steps[w].size(); si++) {
            const auto &st = steps[w][si];
            const size_t sbase = (size_t)(step0[w] + si) * K * 64;
            int gofs = 0;
            for (int l = 0; l < 64; l++) {
                const int q = st[l];
                int4 *op = &rec[sbase];
                // defaults: inactive lane / padding
                int srcs[8];
                double vals[8];
                for (int j = 0; j < 8; j++) {
                    srcs[j] = kPad;
                    vals[j] = 0.0;
                }
                int hg = gofs << 1;
                if (q >= 0) {
                    const int r = pof(q);
                    const int orig = lf[r];
                    origv[(size_t)(step0[w] + si) * 64 + l] = orig;
                    int nl = 0;
                    bool hasg = false;
                    for (int e = rp[r]; e < rp[r + 1]; e++) {
                        if (cc[e] == r) continue;
                        const int u = qof(cc[e]);
                        const bool loc = wave[u] == w && si - stepw[u] <= kR;
                        if (loc && nl < EC) {
                            const int du = lanew[u] - l;
                            if (stepw[u] == si - 1 && du == 0) srcs[nl] = kPrevSame;
                            else if (stepw[u] == si - 1 && du == -1) srcs[nl] = kPrevLower;
                            else if (stepw[u] == si - 1 && du == 1) srcs[nl] = kPrevUpper;
                            else srcs[nl] = (stepw[u] % kR) * 64 + lanew[u];
                            vals[nl] = cv[e];
                            nl++;
                        } else {
                            needpub[u] = 1;
                            gsrc.push_back(u);
                            gval.push_back(cv[e]);
                            gmeta.push_back(gord * 2);
                            hasg = true;
                        }
                    }
                    if (hasg) {
                        gmeta.back() |= 1;
                        gord++;
                        hg = (gofs << 1) | 1;
                        gofs++;
                    }
                    const double d = invd[r];
                    long long dbits;
                    std::memcpy(&dbits, &d, 8);
                    op[0 * 64 + l] = make_int4(q, orig, (int)(dbits & 0xffffffff),
                                               (int)(dbits >> 32));
                } else {
                    op[0 * 64 + l] = make_int4(-1, 0, 0, 0);
                }
                // sources
                int flat[12];
                flat[0] = hg;
                for (int j = 0; j < 11; j++) flat[1 + j] = j < EC ? srcs[j] : kPad;
                for (int k = 0; k < KS; k++)
                    op[(1 + k) * 64 + l] = make_int4(flat[4 * k], flat[4 * k + 1],
                                                     flat[4 * k + 2], flat[4 * k + 3]);
                for (int k = 0; k < (EC + 1) / 2; k++) {
                    long long v0, v1;
                    std::memcpy(&v0, &vals[2 * k], 8);
                    std::memcpy(&v1, &vals[2 * k + 1], 8);
                    op[(1 + KS + k) * 64 + l] =
                        make_int4((int)(v0 & 0xffffffff), (int)(v0 >> 32),
                                  (int)(v1 & 0xffffffff), (int)(v1 >> 32));
                }
            }
        }
    }
    // rows read by other waves (or beyond the LDS window) publish their x
    for (int q = 0; q < m; q++)
        if (needpub[q])
            rec[((size_t)(step0[wave[q]] + stepw[q]) * K + 1) * 64 + lanew[q]].x |=
                kPublish;
    gent0[W] = (int)gsrc.size();
    if (gsrc.empty()) {
        gsrc.push_back(0);
        gval.push_back(0.0);
        gmeta.push_back(1);
    }
    // ---- upload
    FusedPlan *F = new FusedPlan();
    F->W = W;
    F->EC = EC;
    F->nsteps = NS;
    F->nent = gent0[W];
    Arena &A = p->arena;
    RB_TRY(A.alloc(&F->d_step0, W + 1));
    RB_TRY(A.alloc(&F->d_gent0, W + 1));
    RB_TRY(A.alloc(&F->d_rec, rec.size()));
    RB_TRY(A.alloc(&F->d_brec, (size_t)NS * 64));
    RB_TRY(A.alloc(&F->d_orig, (size_t)NS * 64));
    RB_TRY(A.alloc(&F->d_gsrc, gsrc.size()));
    RB_TRY(A.alloc(&F->d_gval, gval.size()));
    RB_TRY(A.alloc(&F->d_gmeta, gmeta.size()));
    RB_TRY(A.alloc(&F->d_xq, m));
    if (getenv("RB_FUSED_TRACE")) {
        RB_TRY(A.alloc(&F->d_trace, (size_t)W * 16));
        RB_HIP(hipMemsetAsync(F->d_trace, 0, sizeof(unsigned long long) * W * 16, s));
    }
    auto up = [&](void *d, const void *h, size_t bytes) -> rb_status {
        RB_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
        return RB_OK;
    };
    RB_TRY(up(F->d_step0, step0.data(), sizeof(int) * step0.size()));
    RB_TRY(up(F->d_gent0, gent0.data(), sizeof(int) * gent0.size()));
    RB_TRY(up(F->d_rec, rec.data(), sizeof(int4) * rec.size()));
    RB_TRY(up(F->d_orig, origv.data(), sizeof(int) * origv.size()));
    RB_TRY(up(F->d_gsrc, gsrc.data(), sizeof(int) * gsrc.size()));
    RB_TRY(up(F->d_gval, gval.data(), sizeof(double) * gval.size()));
    RB_TRY(up(F->d_gmeta, gmeta.data(), sizeof(int) * gmeta.size()));
    RB_HIP(hipStreamSynchronize(s));
    auto t1 = std::chrono::steady_clock::now();
    F->build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    delete p->fused;
    p->fused = F;
    return RB_OK;
}

// Device part of the fused build: pull the permuted CSR and the tables the
// host partitioner needs.
rb_status build_fused(rb_plan *p, const int *d_prptr, const int *d_ccol,
                      const double *d_cval, int Wmax, hipStream_t s) {
    const int m = p->m, nnz = p->nnz;
    std::vector<int> rp(m + 1), cc(nnz > 0 ? nnz : 1), lp(p->nlevel + 1),
        lf(m > 0 ? m : 1);
    std::vector<double> cv(nnz > 0 ? nnz : 1), invd(m > 0 ? m : 1);
    RB_HIP(hipMemcpyAsync(rp.data(), d_prptr, sizeof(int) * (m + 1),
                          hipMemcpyDeviceToHost, s));
    if (nnz > 0) {
        RB_HIP(hipMemcpyAsync(cc.data(), d_ccol, sizeof(int) * nnz,
                              hipMemcpyDeviceToHost, s));
        RB_HIP(hipMemcpyAsync(cv.data(), d_cval, sizeof(double) * nnz,
                              hipMemcpyDeviceToHost, s));
    }
    RB_HIP(hipMemcpyAsync(lp.data(), p->d_levelPtr,
                          sizeof(int) * (p->nlevel + 1), hipMemcpyDeviceToHost,
                          s));
    if (m > 0) {
        RB_HIP(hipMemcpyAsync(lf.data(), p->d_levelItem, sizeof(int) * m,
                              hipMemcpyDeviceToHost, s));
        RB_HIP(hipMemcpyAsync(invd.data(), p->d_invd, sizeof(double) * m,
                              hipMemcpyDeviceToHost, s));
    }
    RB_HIP(hipStreamSynchronize(s));
    return build_fused_host(p, rp, cc, cv, lp, lf, invd, Wmax, s);
}

int fused_max_waves() {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                              dev) != hipSuccess)
        return 256;
    return ncu > 0 ? ncu : 256;
}

template <int EC>
static rb_status launch_fused(FusedPlan *F, const FusedArgs &args,
                              hipStream_t s) {
    void *kargs[] = {(void *)&args};
    const int threads = 64 * (2 + Layout<EC>::kNF);
    hipError_t e = hipLaunchCooperativeKernel((const void *)k_fused<EC>,
                                              dim3(F->W), dim3(threads), kargs,
                                              0, s);
    if (e != hipSuccess) return status_of_hip(e, "hipLaunchCooperativeKernel(k_fused)");
    return RB_OK;
}

rb_status solve_fused(rb_plan *p, const double *d_b, int ldb, double *d_x,
                      int ldx, int nrhs, hipStream_t s) {
    FusedPlan *F = p->fused;
    if (!F) {
        set_error(RB_ERR_INTERNAL, "fused schedule not built");
        return RB_ERR_INTERNAL;
    }
    const int m = p->m;
    if (m == 0) return RB_OK;
    const long nslots = (long)F->nsteps * 64;
    for (int k = 0; k < nrhs; k++) {
        const double *b = d_b + (size_t)k * ldb;
        double *x = d_x + (size_t)k * ldx;
        const long nt = std::max(nslots, (long)m);
        k_fused_prologue<<<(int)((nt + 255) / 256), 256, 0, s>>>(
            F->d_orig, nslots, b, F->d_brec, F->d_xq, m);
        RB_HIP(hipGetLastError());
        FusedArgs args{F->d_step0, F->d_gent0, F->d_rec, F->d_brec,
                       F->d_gsrc, F->d_gval, F->d_gmeta, F->d_xq, x,
                       p->d_status, F->d_trace};
        switch (F->EC) {
        case 2: RB_TRY(launch_fused<2>(F, args, s)); break;
        case 4: RB_TRY(launch_fused<4>(F, args, s)); break;
        default: RB_TRY(launch_fused<8>(F, args, s)); break;
        }
    }
    return RB_OK;
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
