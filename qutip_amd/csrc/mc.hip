// Batched Monte-Carlo wave-function trajectories (mcsolve) on gfx950.
//
// One workgroup runs one trajectory from the first to the last output time
// (grid-stride over trajectories), with no host round trip: the explicit
// Runge-Kutta integrator in its "mcstep" form, the norm-based collapse
// search and the quantum jumps all run on the device.  Trajectories advance
// independently with their own step sizes and jump counts; a workgroup that
// finishes early simply picks up the next trajectory.
//
// Restated algorithms (qutip/solver/...):
//   RK step, step control, dense output     integrator/explicit_rk.pyx:204-467
//   integrate(t, step=True)  (mcstep)        explicit_rk.pyx:278-331
//   MCIntegrator.set_state / integrate       mcsolve.py:245-302
//   _find_collapse_time (log-ratio secant)   mcsolve.py:321-369
//   _do_collapse (channel choice, jump)      mcsolve.py:371-406
//   H_eff = -iH - 1/2 sum c^dag c            mcsolve.py:491-496 (built on host)
// Random numbers: the uniform draws of numpy's Generator (PCG64 seeded by
// SeedSequence(seed), multitraj.py:354-393) are produced on the host in the
// order the reference consumes them (mcsolve.py:276-279, 388, 405) and
// streamed in; a trajectory that uses up its draws stops with a status the
// host acts on.
//
// Layout: each thread owns one or two rows of the state (d <= 1024).  The
// RK stage vectors stay on the chip for the whole trajectory: all sixteen in
// LDS when they take at most half of it (two workgroups per CU), else
// k_0..k_9 in registers and k_10..k_15 in LDS.  The vector an RHS product
// reads is staged in LDS.  H_eff is
// a column-major ELL in HBM (entry u of row r at u d + r), so the 64 lanes of
// a wave, which own 64 consecutive rows, load 64 consecutive columns and
// values per instruction; c_ops and e_ops (used at jumps and outputs only)
// are CSR.  All operators are shared by all trajectories and served from L2.  Norms, error norms and expectation values accumulate in FP64;
// with the complex64 variant the state, stages and RHS products are FP32.
#include <algorithm>
#include <math.h>
#include <string.h>

#include "qa_api_internal.h"
#include "qa_common.h"

using namespace qa;

namespace {

constexpr int kMcStages = 16;
constexpr int kMcOrder = 8;
constexpr int kMcMaxChan = 32;
constexpr int kMcMaxEops = 8;
constexpr int kMcThreads = 512;  // 2 waves per SIMD: 256 VGPRs per lane

// trajectory status codes (beside the RK codes of explicit_rk.pxd:5-12)
enum McStatus {
    MC_OK = 0,
    MC_TOO_MUCH_WORK = -1,
    MC_DT_UNDERFLOW = -2,
    MC_OUTSIDE_RANGE = -3,
    MC_COLLAPSE_NOT_FOUND = -10,  // mcsolve.py:363-367 RuntimeError
    MC_DRAWS_EXHAUSTED = -11,     // host re-runs with more pre-drawn numbers
    MC_NO_PROGRESS = -12
};
enum RkStatus { RK_AT_FRONT = 2, RK_INTERPOLATED = 1, RK_NORMAL = 0 };

struct McTableau {
    double a[kMcStages][kMcStages];
    double b[kMcStages];
    double c[kMcStages];
    double e[kMcStages];
    double bi[kMcStages][kMcOrder];
    int order, rk_step, rk_extra, dorder, interpolate, adaptive, fsal, pad;
};

struct McArgs {
    int d, ntraj, ntimes, nchan, neops, max_draws, max_jumps, traj_offset;
    int h_width;       // ELL width of H_eff
    const int *h_col;  // [h_width][d] column-major ELL (entry u of row r at u d + r)
    const void *h_vals;
    const int *c_rowptr, *c_col;  // nchan blocks of d + 1 row pointers (absolute)
    const void *c_vals;
    const int *e_rowptr, *e_col;
    const void *e_vals;
    const McTableau *tab;
    const double *times;
    const double2 *psi0;
    const double *draws;  // [ntraj][max_draws]
    double rtol, atol, first_step, min_step, max_step;
    int nsteps, norm_steps;
    double norm_t_tol, norm_tol, norm_min_step, mc_corr_eps;
    double2 *expect_out;  // [ntraj][ntimes][neops]
    double *col_times;    // [ntraj][max_jumps]
    int *col_which;
    int *info;  // [ntraj][4]: status, jumps, draws used, RK attempts
    // whole-solve mode (rk_small_kernel): per-system operator stride
    // (entries; 0 = one shared operator), start time, linear / quadratic
    // functionals of the state, outputs
    long long h_stride;
    double t0;
    int nfun;
    const int *f_kind, *f_ptr, *f_row, *f_col;
    const double2 *f_w;
    double2 *fun_out;  // [nsys][ntimes][nfun]
    double2 *y_final;  // [nsys][d] or null
    int fun_lds;       // LDS bytes reserved for the functionals and times (0: none)
    int x2;            // a second LDS copy of the product's input vector is reserved
};

template <typename R>
struct cplx_t;
template <>
struct cplx_t<double> {
    using type = double2;
};
template <>
struct cplx_t<float> {
    using type = float2;
};

template <typename R>
using C2 = typename cplx_t<R>::type;

template <typename R>
__device__ __forceinline__ C2<R> mk(R x, R y) {
    C2<R> v;
    v.x = x;
    v.y = y;
    return v;
}
template <typename R>
__device__ __forceinline__ void cfmaR(C2<R> &acc, C2<R> a, C2<R> b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}
template <typename R>
__device__ __forceinline__ void raxpyR(C2<R> &acc, R s, C2<R> v) {
    acc.x = fma(s, v.x, acc.x);
    acc.y = fma(s, v.y, acc.y);
}
template <typename R>
__device__ __forceinline__ double nrm2(C2<R> v) {
    return (double)v.x * (double)v.x + (double)v.y * (double)v.y;
}

// RS (a template parameter): stages k_0..k_{RS-1} live in registers, the
// rest in LDS.  RS = 10 keeps the step stages of vern7 in registers (512
// threads, up to 256 VGPRs each); RS = 0 keeps every stage in LDS, which
// frees the registers for 1024-thread workgroups (more waves in flight) when
// the LDS holds all 18 vectors.
constexpr int kRegStagesMax = 10;

// Per-workgroup LDS (dynamic): the vector an RHS product reads, the stage
// vectors k_10.. (the dense-output stages of vern7) and the FSAL slot, one
// entry per row, plus reduction scratch.
template <typename R>
struct McLds {
    C2<R> *x;      // [npad]
    C2<R> *xalt;   // [npad] the other input buffer, or null (one buffer)
    C2<R> *kd;     // [kMcStages - RS][npad]
    C2<R> *kfsal;  // [npad]
    double *red;   // [2 * kFunGroup * threads / 64]
    double *probs; // [kMcMaxChan]
    double *bfs;   // [2][kMcStages] dense-output weights (two buffers, alternating)
    int bsel;      // the buffer the next interpolation writes
    double *rsum;  // [2][2 * threads / 64] block-sum partials (two buffers, alternating)
    int rsel;      // the buffer the next block sum writes
    int npad;
};

// Functionals summed per block reduction (one pair of barriers for each
// group) by the whole-solve kernel.
constexpr int kFunGroup = 4;

template <typename R>
inline size_t mc_lds_bytes(int npad, int threads, int RS) {
    return sizeof(C2<R>) * (size_t)npad * (2 + kMcStages - RS) +
           sizeof(double) * ((2 * kFunGroup + 4) * (threads / 64) + kMcMaxChan + 2 * kMcStages) +
           64;
}

// A per-row value: each thread owns NR rows (r = threadIdx.x + q * blockDim.x).
template <typename R, int NR>
struct V {
    C2<R> v[NR];
};

template <typename R, int NR>
__device__ __forceinline__ V<R, NR> vzero() {
    V<R, NR> o;
#pragma unroll
    for (int q = 0; q < NR; ++q) o.v[q] = mk<R>(0, 0);
    return o;
}

// Per-trajectory integrator state.  Vectors are per thread; scalars are
// uniform (every thread holds the same value).
template <typename R, int NR, int RS>
struct Traj {
    V<R, NR> k[RS > 0 ? RS : 1];
    V<R, NR> yprev, yfront, y;
    double t, t_prev, t_front, dt_safe, dt_int;
    int status;    // trajectory status (McStatus)
    int rk_state;  // last RK status (INTERPOLATED marks a prepared dense output)
    long long attempts;
    int ndense;  // dense-output preparations
};

template <typename R, int NR, int RS>
struct Ctx {
    const McArgs &A;
    McLds<R> &S;
    const McTableau &T;
    int row[NR];
    bool act[NR];
    int rowc[NR];  // the row clamped into [0, d) for the ELL loads
    const int *hcol;    // this system's H_eff ELL
    const void *hval;
};

// Stage vector j; j is a compile-time constant after unrolling, so the
// register / LDS choice costs nothing.
template <typename R, int NR, int RS>
__device__ __forceinline__ V<R, NR> kget(Ctx<R, NR, RS> &c, Traj<R, NR, RS> &s, int j) {
    if (j < RS) return s.k[j < RS ? j : 0];
    V<R, NR> o;
#pragma unroll
    for (int q = 0; q < NR; ++q)
        o.v[q] = c.S.kd[(size_t)(j - RS) * c.S.npad + c.row[q]];
    return o;
}
template <typename R, int NR, int RS>
__device__ __forceinline__ void kset(Ctx<R, NR, RS> &c, Traj<R, NR, RS> &s, int j, const V<R, NR> &x) {
    if (j < RS) {
        s.k[j < RS ? j : 0] = x;
    } else {
#pragma unroll
        for (int q = 0; q < NR; ++q) c.S.kd[(size_t)(j - RS) * c.S.npad + c.row[q]] = x.v[q];
    }
}

// Block sums write their wave partials to one of two LDS buffers in turn,
// so one barrier per sum suffices: a buffer is written again only after the
// barrier of the sum in between, which every thread passes after its reads
// of it (the buffers serve nothing else).
template <typename R, int NR, int RS>
__device__ __forceinline__ double bsum(Ctx<R, NR, RS> &c, double v) {
    const int lane = threadIdx.x & (QA_WAVE - 1), wid = threadIdx.x / QA_WAVE;
    const int nw = blockDim.x / QA_WAVE;
    double *rb = c.S.rsum + 2 * nw * c.S.rsel;
    c.S.rsel ^= 1;
    v = wave_sum(v);
    if (lane == 0) rb[wid] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += rb[w];
    return t;
}

// Two block sums with one pair of barriers; each component is summed in
// the order bsum uses, so the results are bsum's bits.
template <typename R, int NR, int RS>
__device__ __forceinline__ double2 bsum2(Ctx<R, NR, RS> &c, double a, double b) {
    const int lane = threadIdx.x & (QA_WAVE - 1), wid = threadIdx.x / QA_WAVE;
    const int nw = blockDim.x / QA_WAVE;
    double *rb = c.S.rsum + 2 * nw * c.S.rsel;
    c.S.rsel ^= 1;
    a = wave_sum(a);
    b = wave_sum(b);
    if (lane == 0) {
        rb[wid] = a;
        rb[nw + wid] = b;
    }
    __syncthreads();
    double ta = 0.0, tb = 0.0;
    for (int w = 0; w < nw; ++w) {
        ta += rb[w];
        tb += rb[nw + w];
    }
    return make_double2(ta, tb);
}

// The input of a product (or of the functionals) into LDS, for every thread
// to gather from.  With two buffers the writes go to the one not read last,
// so only the barrier after them is needed: a thread still gathering from
// the other buffer has not reached that barrier, and every read of a
// buffer lies between the barrier after its writes and the barrier after
// the next input's, before the buffer is written again.
template <typename R, int NR, int RS>
__device__ __forceinline__ void stage_in(Ctx<R, NR, RS> &c, const V<R, NR> &x) {
    if (c.S.xalt) {  // uniform
        C2<R> *t = c.S.x;
        c.S.x = c.S.xalt;
        c.S.xalt = t;
    } else {
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) c.S.x[c.row[q]] = x.v[q];
    __syncthreads();
}

// M ELL entries (u0 .. u0 + M - 1) of this thread's rows: all column and
// value loads first, then the LDS gathers and multiply-adds.
template <int M, typename R, int NR, int RS>
__device__ __forceinline__ void ell_chunk(Ctx<R, NR, RS> &c, int u0, V<R, NR> &acc) {
    const C2<R> *hv = (const C2<R> *)c.hval;
    const size_t d = (size_t)c.A.d;
    int cc[NR][M];
    C2<R> vv[NR][M];
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
        for (int w = 0; w < M; ++w) cc[q][w] = c.hcol[(size_t)(u0 + w) * d + c.rowc[q]];
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
        for (int w = 0; w < M; ++w) vv[q][w] = hv[(size_t)(u0 + w) * d + c.rowc[q]];
#pragma unroll
    for (int q = 0; q < NR; ++q)
#pragma unroll
        for (int w = 0; w < M; ++w) cfmaR<R>(acc.v[q], vv[q][w], c.S.x[cc[q][w]]);
}

// k = H_eff x for this thread's rows; x is staged in LDS first.  The ELL
// width is a multiple of 4 (the host pads it), so entries go in chunks of 8
// (4 for two complex128 rows per thread, which leaves room for the stage
// vectors in registers) and at most one chunk of 4; no load is predicated.
template <typename R, int NR, int RS>
__device__ __forceinline__ V<R, NR> rhs(Ctx<R, NR, RS> &c, const V<R, NR> &x) {
    constexpr int K = (NR == 2 && sizeof(R) == 8) ? 4 : 8;
    stage_in(c, x);
    V<R, NR> acc = vzero<R, NR>();
    const int W = c.A.h_width;
    int u = 0;
    for (; u + K <= W; u += K) ell_chunk<K>(c, u, acc);
    if (K == 8 && u < W) ell_chunk<4>(c, u, acc);
#pragma unroll
    for (int q = 0; q < NR; ++q)
        if (!c.act[q]) acc.v[q] = mk<R>(0, 0);  // padding rows stay zero
    return acc;
}

// (O x)_r for a CSR operator (c_ops / e_ops block); x already in LDS.
template <typename R, int NR, int RS>
__device__ __forceinline__ V<R, NR> op_rows(Ctx<R, NR, RS> &c, const int *rowptr, const int *col,
                                            const void *vals) {
    V<R, NR> acc = vzero<R, NR>();
    const C2<R> *v = (const C2<R> *)vals;
#pragma unroll
    for (int q = 0; q < NR; ++q)
        if (c.act[q])
            for (int p = rowptr[c.row[q]]; p < rowptr[c.row[q] + 1]; ++p)
                cfmaR<R>(acc.v[q], v[p], c.S.x[col[p]]);
    return acc;
}

template <typename R, int NR, int RS>
__device__ __forceinline__ double norm2(Ctx<R, NR, RS> &c, const V<R, NR> &x) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < NR; ++q) s += nrm2<R>(x.v[q]);  // inactive rows hold zeros
    return bsum(c, s);
}

// base + dt * sum_{j < n} coef[j] k_j (zero terms skipped like iadd_data)
template <typename R, int NR, int RS>
__device__ __forceinline__ V<R, NR> combine(Ctx<R, NR, RS> &c, Traj<R, NR, RS> &s, const V<R, NR> &base,
                                            const double *coef, int n, double dt) {
    V<R, NR> o = base;
#pragma unroll
    for (int j = 0; j < kMcStages; ++j) {
        if (j < n) {
            const double f = dt * coef[j];
            if (f != 0.0) {
                const V<R, NR> kj = kget(c, s, j);
#pragma unroll
                for (int q = 0; q < NR; ++q) raxpyR<R>(o.v[q], (R)f, kj.v[q]);
            }
        }
    }
    return o;
}

// _compute_step (explicit_rk.pyx:358-389) + _error (:391-397)
template <typename R, int NR, int RS>
__device__ __forceinline__ double compute_step(Ctx<R, NR, RS> &c, Traj<R, NR, RS> &s, double dt) {
    const McTableau &T = c.T;
    s.attempts++;
    if (T.fsal) {
        V<R, NR> k0;
#pragma unroll
        for (int q = 0; q < NR; ++q) k0.v[q] = c.S.kfsal[c.row[q]];
        kset(c, s, 0, k0);
    } else {
        kset(c, s, 0, rhs(c, s.yprev));
    }
#pragma unroll
    for (int i = 1; i < kMcStages; ++i) {
        if (i < T.rk_step) {
            const V<R, NR> x = combine(c, s, s.yprev, T.a[i], i, dt);
            kset(c, s, i, rhs(c, x));
        }
    }
    s.yfront = combine(c, s, s.yprev, T.b, T.rk_step, dt);
    s.t_front = s.t_prev + dt;
    if (!T.adaptive) return 0.0;
    const V<R, NR> ev = combine(c, s, vzero<R, NR>(), T.e, T.rk_step, dt);
    double qsum = 0.0;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        if (c.act[q]) {
            const double ratio = sqrt(nrm2<R>(ev.v[q])) /
                                 (c.A.atol + c.A.rtol * sqrt(nrm2<R>(s.yfront.v[q])));
            qsum += ratio * ratio;
        }
    }
    return sqrt(bsum(c, qsum) / (double)c.A.d);
}

template <typename R, int NR, int RS>
__device__ __forceinline__ void prep_dense_out(Ctx<R, NR, RS> &c, Traj<R, NR, RS> &s) {
    const McTableau &T = c.T;
    s.ndense++;
#pragma unroll
    for (int i = 0; i < kMcStages; ++i) {
        if (i >= T.rk_step && i < T.rk_extra) {
            const V<R, NR> x = combine(c, s, s.yprev, T.a[i], i, s.dt_int);
            kset(c, s, i, rhs(c, x));
        }
    }
}

// The weights b_i(tau) (explicit_rk.pyx:412-430) are formed by threads
// i < kMcStages, each with its row of the table loaded at once (a loop to
// the run-time order would wait on every load before the next step), and
// shared through LDS.  Two weight buffers alternate, so one barrier per
// interpolation suffices: a buffer is written again only after the barrier
// of the interpolation in between, which every thread passes after its
// reads of it.
template <typename R, int NR, int RS>
__device__ __forceinline__ V<R, NR> interpolate(Ctx<R, NR, RS> &c, Traj<R, NR, RS> &s, double t) {
    const McTableau &T = c.T;
    double *bw = c.S.bfs + kMcStages * c.S.bsel;
    c.S.bsel ^= 1;
    const int i = threadIdx.x;
    if (i < kMcStages) {
        const double tau = (t - s.t_prev) / s.dt_int;
        double bv[kMcOrder];
#pragma unroll
        for (int j = 0; j < kMcOrder; ++j) bv[j] = T.bi[i][j];
        double v = 0.0;
        if (i < T.rk_extra) {
#pragma unroll
            for (int j = kMcOrder - 1; j >= 0; --j)
                if (j < T.dorder) {  // Horner from j = dorder - 1 down
                    v += bv[j];
                    v *= tau;
                }
        }
        bw[i] = v;
    }
    __syncthreads();
    double bf[kMcStages];
#pragma unroll
    for (int k = 0; k < kMcStages; ++k) bf[k] = bw[k];
    return combine(c, s, s.yprev, bf, T.rk_extra, s.dt_int);
}

template <typename R, int NR, int RS>
__device__ __forceinline__ double estimate_first_step(Ctx<R, NR, RS> &c, Traj<R, NR, RS> &s,
                                                      const V<R, NR> &y0) {
    const McTableau &T = c.T;
    if (!T.adaptive) return 0.0;
    double norm = sqrt(norm2(c, y0));
    const double tol = c.A.atol + norm * c.A.rtol;
    // f(y0) and f(y0 + dt1/100 f(y0)) are local temporaries (the reference
    // borrows k[0] and k[1] for them; the next step overwrites both anyway)
    V<R, NR> f0;
    if (T.fsal) {
#pragma unroll
        for (int q = 0; q < NR; ++q) f0.v[q] = c.S.kfsal[c.row[q]];
    } else {
        f0 = rhs(c, y0);
    }
    if (norm <= c.A.atol) norm = 1.0;
    double tmp = sqrt(norm2(c, f0));
    double factorial = 1.0;
    for (int i = 1; i <= T.order; ++i) factorial *= i;
    double dt1, dt2;
    if (tmp >= c.A.atol * 1e-6)
        dt1 = pow(tol * factorial * pow(norm, T.order), 1.0 / (T.order + 1)) / tmp;
    else
        dt1 = pow(tol * factorial, 1.0 / (T.order + 1)) * norm * 0.5;
    V<R, NR> yt = y0;
#pragma unroll
    for (int q = 0; q < NR; ++q) raxpyR<R>(yt.v[q], (R)(dt1 / 100), f0.v[q]);
    const V<R, NR> f1 = rhs(c, yt);
    tmp = sqrt(norm2(c, f1));
    if (tmp >= c.A.atol * 1e-6)
        dt2 = pow(tol * factorial * pow(norm, T.order), 1.0 / (T.order + 1)) / tmp;
    else
        dt2 = dt1;
    double dt = fmin(dt1, dt2);
    if (c.A.max_step != 0.0) dt = fmin(c.A.max_step, dt);
    if (c.A.min_step != 0.0) dt = fmax(c.A.min_step, dt);
    return dt;
}

// set_initial_value (explicit_rk.pyx:204-230)
template <typename R, int NR, int RS>
__device__ __forceinline__ void set_initial_value(Ctx<R, NR, RS> &c, Traj<R, NR, RS> &s,
                                                  const V<R, NR> &y0, double t0) {
    s.t = s.t_prev = s.t_front = t0;
    s.dt_int = 0.0;
    s.y = y0;
    s.yfront = y0;
    s.yprev = y0;
    s.rk_state = RK_NORMAL;
    if (c.T.fsal) {
        const V<R, NR> kf = rhs(c, y0);
#pragma unroll
        for (int q = 0; q < NR; ++q) c.S.kfsal[c.row[q]] = kf.v[q];
    }
    s.dt_safe = c.A.first_step != 0.0 ? c.A.first_step : estimate_first_step(c, s, y0);
}

template <typename R, int NR, int RS>
__device__ __forceinline__ void recompute_safe_step(Ctx<R, NR, RS> &c, Traj<R, NR, RS> &s, double err,
                                                    double dt) {
    if (!c.T.adaptive) return;
    double f;
    if (err == 0.0) {
        f = 10.0;
    } else {
        f = 0.9 * pow(err, -1.0 / (c.T.order + 1));
        f = fmin(10.0, f);
        f = fmax(0.2, f);
    }
    s.dt_safe = dt * f;
    if (c.A.max_step != 0.0) s.dt_safe = fmin(c.A.max_step, s.dt_safe);
    if (c.A.min_step != 0.0) s.dt_safe = fmax(c.A.min_step, s.dt_safe);
}

template <typename R, int NR, int RS>
__device__ __forceinline__ double get_timestep(Ctx<R, NR, RS> &c, Traj<R, NR, RS> &s, double t) {
    const double dt_needed = t - s.t_prev;
    if (!c.T.adaptive) return dt_needed;
    if (c.T.interpolate) return s.dt_safe;
    if (dt_needed <= s.dt_safe) return dt_needed;
    return dt_needed / (double)((long long)(dt_needed / s.dt_safe) + 1);
}

// _step_in_err (explicit_rk.pyx:333-356); returns attempts, sets status
template <typename R, int NR, int RS>
__device__ __forceinline__ int step_in_err(Ctx<R, NR, RS> &c, Traj<R, NR, RS> &s, double t, int max_step) {
    double error = 1.0;
    int n = 0;
    while (error >= 1.0) {
        const double dt = get_timestep(c, s, t);
        error = compute_step(c, s, dt);
        s.dt_int = dt;
        recompute_safe_step(c, s, error, dt);
        if (dt == c.A.min_step && error > 1.0) {
            s.status = MC_DT_UNDERFLOW;
            break;
        }
        n++;
        if (n > max_step) {
            s.status = MC_TOO_MUCH_WORK;
            break;
        }
    }
    if (c.T.fsal) {
#pragma unroll
        for (int j = 0; j < kMcStages; ++j) {
            if (j == c.T.rk_step - 1) {
                const V<R, NR> kl = kget(c, s, j);
#pragma unroll
                for (int q = 0; q < NR; ++q) c.S.kfsal[c.row[q]] = kl.v[q];
            }
        }
    }
    return n;
}

// integrate(t, step=True): explicit_rk.pyx:278-331 with step=True.
template <typename R, int NR, int RS>
__device__ __forceinline__ void mcstep(Ctx<R, NR, RS> &c, Traj<R, NR, RS> &s, double t) {
    int nsteps_left = c.A.nsteps;
    if (t == s.t) return;
    if (t < s.t_prev) {
        s.status = MC_OUTSIDE_RANGE;
        return;
    }
    if (c.T.interpolate && t < s.t_front) {
        if (s.rk_state != RK_INTERPOLATED) {
            prep_dense_out(c, s);
            s.rk_state = RK_INTERPOLATED;
        }
        s.y = interpolate(c, s, t);
        s.t = t;
        return;
    }
    s.rk_state = RK_NORMAL;
    if (s.t < s.t_front && t > s.t_front) t = s.t_front;
    if (s.t_front < t) {
        s.yprev = s.yfront;
        s.t_prev = s.t_front;
        nsteps_left -= step_in_err(c, s, t, nsteps_left);
    }
    if (s.status < 0) return;
    if (s.t_front > t) {
        prep_dense_out(c, s);
        s.rk_state = RK_INTERPOLATED;
        s.t = t;
        s.y = interpolate(c, s, t);
    } else {
        s.rk_state = RK_AT_FRONT;
        s.t = s.t_front;
        s.y = s.yfront;
    }
}

struct Rng {
    const double *d;
    int n, used;
};

__device__ __forceinline__ bool draw(Rng &g, double &u) {
    if (g.used >= g.n) return false;
    u = g.d[g.used++];
    return true;
}

template <typename R, int NR, int RS>
__device__ __forceinline__ void write_expect(Ctx<R, NR, RS> &c, const V<R, NR> &y, int traj, int k) {
    const double n2 = norm2(c, y);
    stage_in(c, y);
    for (int e = 0; e < c.A.neops; ++e) {
        const int *rp = c.A.e_rowptr + (size_t)e * (c.A.d + 1);
        const V<R, NR> oy = op_rows(c, rp, c.A.e_col, c.A.e_vals);
        double re = 0.0, im = 0.0;  // conj(y_r) (O y)_r in FP64
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            re += (double)y.v[q].x * oy.v[q].x + (double)y.v[q].y * oy.v[q].y;
            im += (double)y.v[q].x * oy.v[q].y - (double)y.v[q].y * oy.v[q].x;
        }
        const double2 r = bsum2(c, re, im);
        if (threadIdx.x == 0)
            c.A.expect_out[((size_t)traj * c.A.ntimes + k) * c.A.neops + e] =
                make_double2(r.x / n2, r.y / n2);
    }
}

// LDS carving, tableau copy and zeroing shared by the kernels of this file.
template <typename R, int NR, int RS>
__device__ __forceinline__ void kernel_prologue(const McArgs &A, unsigned char *lds,
                                                McTableau &T, McLds<R> &S) {
    const int nt = blockDim.x;
    const int npad = nt * NR;
    {
        unsigned char *p = lds;
        S.npad = npad;
        S.x = (C2<R> *)p;
        p += sizeof(C2<R>) * npad;
        S.kd = (C2<R> *)p;
        p += sizeof(C2<R>) * npad * (kMcStages - RS);
        S.kfsal = (C2<R> *)p;
        p += sizeof(C2<R>) * npad;
        S.red = (double *)p;
        p += sizeof(double) * 2 * kFunGroup * (nt / QA_WAVE);
        S.probs = (double *)p;
        p += sizeof(double) * kMcMaxChan;
        S.bfs = (double *)p;
        S.bsel = 0;
        p += sizeof(double) * 2 * kMcStages;
        S.rsum = (double *)p;
        S.rsel = 0;
        S.xalt = nullptr;
    }
    {
        const double *src = (const double *)A.tab;
        double *dst = (double *)&T;
        const int nd = (int)(sizeof(McTableau) / sizeof(double));
        for (int i = threadIdx.x; i < nd; i += nt) dst[i] = src[i];
    }
    for (int i = threadIdx.x; i < (kMcStages - RS + 2) * npad; i += nt)
        S.x[i] = mk<R>(0, 0);  // x, the LDS stages and the FSAL slot start at zero
    __syncthreads();
}

template <typename R, int NR, int RS>
__device__ __forceinline__ void ctx_rows(Ctx<R, NR, RS> &c, const int *hcol, const void *hval) {
    c.hcol = hcol;
    c.hval = hval;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        c.row[q] = threadIdx.x + q * blockDim.x;
        c.act[q] = c.row[q] < c.A.d;
        c.rowc[q] = c.act[q] ? c.row[q] : 0;
    }
}

// integrate(t, step=False) (explicit_rk.pyx:278-331): steps until t_front
// reaches t within one budget of nsteps attempts, then interpolates (or
// takes y_front).
template <typename R, int NR, int RS>
__device__ __forceinline__ void integrate_to(Ctx<R, NR, RS> &c, Traj<R, NR, RS> &s, double t) {
    int nsteps_left = c.A.nsteps;
    if (t == s.t) return;
    if (t < s.t_prev) {
        s.status = MC_OUTSIDE_RANGE;
        return;
    }
    if (c.T.interpolate && t < s.t_front) {
        if (s.rk_state != RK_INTERPOLATED) {
            prep_dense_out(c, s);
            s.rk_state = RK_INTERPOLATED;
        }
        s.y = interpolate(c, s, t);
        s.t = t;
        return;
    }
    s.rk_state = RK_NORMAL;
    while (s.t_front < t && s.status >= 0) {
        s.yprev = s.yfront;
        s.t_prev = s.t_front;
        nsteps_left -= step_in_err(c, s, t, nsteps_left);
    }
    if (s.status < 0) return;
    if (s.t_front > t) {
        prep_dense_out(c, s);
        s.rk_state = RK_INTERPOLATED;
        s.t = t;
        s.y = interpolate(c, s, t);
    } else {
        s.rk_state = RK_AT_FRONT;
        s.t = s.t_front;
        s.y = s.yfront;
    }
}

// p rounded up to 16 bytes as an offset from p itself, so the result keeps
// p's provenance: a pointer into the LDS stays one the compiler knows to be
// in the LDS (DS instructions, not FLAT ones, which wait on both the memory
// and the LDS counters).
template <typename T>
__device__ __forceinline__ unsigned char *align16(T *p) {
    unsigned char *b = (unsigned char *)p;
    return b + ((16 - ((uintptr_t)b & 15)) & 15);
}

// The functionals as the whole-solve kernel reads them: kind, CSR-style
// entry pointer, rows, columns and weights, in LDS when they were staged
// there, else in global memory.
struct FunView {
    const int *kind, *ptr, *row, *col;
    const double2 *w;
};

// Functionals of the state y (linear: sum w y[col]; quadratic:
// sum w conj(y[row]) y[col]), accumulated in FP64 in a fixed order, into
// out[0 .. nfun-1] (thread 0 writes).
template <typename R, int NR, int RS>
__device__ __forceinline__ void write_functionals(Ctx<R, NR, RS> &c, const V<R, NR> &y,
                                                  const FunView &fv, double2 *out) {
    stage_in(c, y);
    const int lane = threadIdx.x & (QA_WAVE - 1), wid = threadIdx.x / QA_WAVE;
    const int nw = blockDim.x / QA_WAVE;
    // kFunGroup functionals per block reduction: each wave's sums go to LDS
    // as they are formed (no array of them in registers), then one barrier,
    // and thread t < kFunGroup adds functional f0 + t over the waves in
    // order, which is bsum2's order, so the bits are bsum2's.  The barrier
    // after stage_in's writes (first group) or at the top of the loop (the
    // previous group's sums read) comes before a wave writes LDS sums.
    for (int f0 = 0; f0 < c.A.nfun; f0 += kFunGroup) {
        if (f0 > 0) __syncthreads();
#pragma unroll 1
        for (int g = 0; g < kFunGroup; ++g) {
            const int f = f0 + g;
            if (f >= c.A.nfun) break;  // uniform
            const int quad = fv.kind[f];
            const int e1 = fv.ptr[f + 1];
            double re = 0.0, im = 0.0;
            for (int e = fv.ptr[f] + threadIdx.x; e < e1; e += blockDim.x) {
                const C2<R> xc = c.S.x[fv.col[e]];
                const double2 w = fv.w[e];
                double vr = w.x * (double)xc.x - w.y * (double)xc.y;
                double vi = w.x * (double)xc.y + w.y * (double)xc.x;
                if (quad) {
                    const C2<R> xr = c.S.x[fv.row[e]];
                    const double lr = (double)xr.x, li = -(double)xr.y;
                    const double tr = lr * vr - li * vi;
                    vi = lr * vi + li * vr;
                    vr = tr;
                }
                re += vr;
                im += vi;
            }
            re = wave_sum(re);
            im = wave_sum(im);
            if (lane == 0) {
                c.S.red[(2 * g) * nw + wid] = re;
                c.S.red[(2 * g + 1) * nw + wid] = im;
            }
        }
        __syncthreads();
        if (threadIdx.x < kFunGroup && f0 + (int)threadIdx.x < c.A.nfun) {
            double ta = 0.0, tb = 0.0;
            for (int w = 0; w < nw; ++w) {
                ta += c.S.red[2 * threadIdx.x * nw + w];
                tb += c.S.red[(2 * threadIdx.x + 1) * nw + w];
            }
            out[f0 + threadIdx.x] = make_double2(ta, tb);
        }
    }
}

template <typename R, int NR, int RS>
__global__ __launch_bounds__(RS == 0 ? 2 * kMcThreads : kMcThreads) void mc_traj_kernel(McArgs A) {
    extern __shared__ __align__(16) unsigned char mc_dyn_lds[];
    __shared__ McTableau T;
    McLds<R> S;
    const int nt = blockDim.x;
    const int npad = nt * NR;
    {
        unsigned char *p = mc_dyn_lds;
        S.npad = npad;
        S.x = (C2<R> *)p;
        p += sizeof(C2<R>) * npad;
        S.kd = (C2<R> *)p;
        p += sizeof(C2<R>) * npad * (kMcStages - RS);
        S.kfsal = (C2<R> *)p;
        p += sizeof(C2<R>) * npad;
        S.red = (double *)p;
        p += sizeof(double) * 2 * kFunGroup * (nt / QA_WAVE);
        S.probs = (double *)p;
        p += sizeof(double) * kMcMaxChan;
        S.bfs = (double *)p;
        S.bsel = 0;
        p += sizeof(double) * 2 * kMcStages;
        S.rsum = (double *)p;
        S.rsel = 0;
        // the second input buffer, when the launcher reserved it
        S.xalt = A.x2 ? (C2<R> *)align16(S.rsum + 4 * (nt / QA_WAVE)) : nullptr;
    }
    {
        const double *src = (const double *)A.tab;
        double *dst = (double *)&T;
        const int nd = (int)(sizeof(McTableau) / sizeof(double));
        for (int i = threadIdx.x; i < nd; i += nt) dst[i] = src[i];
    }
    for (int i = threadIdx.x; i < (kMcStages - RS + 2) * npad; i += nt)
        S.x[i] = mk<R>(0, 0);  // x, the LDS stages and the FSAL slot start at zero
    __syncthreads();
    // (written out here rather than through kernel_prologue / ctx_rows: the
    // shared helpers cost this kernel 100 B/lane more register spill)
    Ctx<R, NR, RS> c{A, S, T};
    c.hcol = A.h_col;
    c.hval = A.h_vals;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        c.row[q] = threadIdx.x + q * nt;
        c.act[q] = c.row[q] < A.d;
        c.rowc[q] = c.act[q] ? c.row[q] : 0;
    }

    for (int traj = blockIdx.x; traj < A.ntraj; traj += gridDim.x) {
        Traj<R, NR, RS> s;
        s.status = MC_OK;
        s.attempts = 0;
        s.ndense = 0;
        Rng g{A.draws + (size_t)traj * A.max_draws, A.max_draws, 0};
        int jumps = 0;
        double target = 0.0;
        V<R, NR> y0;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const double2 p0 = c.act[q] ? A.psi0[c.row[q]] : make_double2(0.0, 0.0);
            y0.v[q] = mk<R>((R)p0.x, (R)p0.y);
        }
        if (!draw(g, target)) s.status = MC_DRAWS_EXHAUSTED;
        if (s.status == MC_OK) set_initial_value(c, s, y0, A.times[0]);
        if (s.status == MC_OK) write_expect(c, y0, traj, 0);
        for (int k = 1; k < A.ntimes && s.status == MC_OK; ++k) {
            const double tk = A.times[k];
            // MCIntegrator.integrate (mcsolve.py:286-302)
            // (the state at t_old is always s.y: every path below that moves
            // t_old also moves s.y there)
            double t_old = s.t;
            double norm_old = norm2(c, s.y);
            long long guard = 0;
            while (t_old < tk && s.status == MC_OK) {
                if (++guard > 100000) {
                    s.status = MC_NO_PROGRESS;
                    break;
                }
                mcstep(c, s, tk);
                if (s.status < 0) break;
                const double t_step = s.t;
                double nrm = norm2(c, s.y);
                if (nrm <= target) {
                    // _find_collapse_time (mcsolve.py:321-369)
                    int tries = 0;
                    double tp = t_old, tf = t_step, nold = norm_old, t_guess = tf;
                    while (tries < A.norm_steps) {
                        tries++;
                        if (tf - tp < A.norm_t_tol) {
                            t_guess = tf;
                            break;
                        }
                        const double dtc = tf - tp;
                        double ratio = log(nold / target) / log(nold / nrm);
                        if (ratio < A.norm_min_step) ratio = A.norm_min_step;
                        if (ratio > 1.0 - A.norm_min_step) ratio = 1.0 - A.norm_min_step;
                        t_guess = tp + dtc * ratio;
                        if (t_guess - tp < A.norm_t_tol) t_guess = tp + A.norm_t_tol;
                        mcstep(c, s, t_guess);
                        if (s.status < 0) break;
                        const double n2g = norm2(c, s.y);
                        if (fabs(target - n2g) < A.norm_tol * target) {
                            break;
                        } else if (n2g < target) {
                            tf = t_guess;
                            nrm = n2g;
                        } else {
                            tp = t_guess;
                            nold = n2g;
                        }
                    }
                    if (s.status < 0) break;
                    if (tries >= A.norm_steps) {
                        s.status = MC_COLLAPSE_NOT_FOUND;
                        break;
                    }
                    // _do_collapse (mcsolve.py:371-406); the state is s.y
                    const V<R, NR> psi = s.y;
                    stage_in(c, psi);
                    int which = 0;
                    if (A.nchan > 1) {
                        double tot = 0.0;
                        for (int ch = 0; ch < A.nchan; ++ch) {
                            const int *rp = A.c_rowptr + (size_t)ch * (A.d + 1);
                            const double pr = norm2(c, op_rows(c, rp, A.c_col, A.c_vals));
                            if (threadIdx.x == 0) S.probs[ch] = pr;
                            tot += pr;
                        }
                        __syncthreads();
                        double u;
                        if (!draw(g, u)) {
                            s.status = MC_DRAWS_EXHAUSTED;
                            break;
                        }
                        double tg = tot * u - S.probs[0];
                        while (tg > 0.0 && which < A.nchan - 1) {
                            which++;
                            tg -= S.probs[which];
                        }
                    }
                    const int *rp = A.c_rowptr + (size_t)which * (A.d + 1);
                    V<R, NR> ynew = op_rows(c, rp, A.c_col, A.c_vals);
                    const double nn = sqrt(norm2(c, ynew));
                    if (nn < A.mc_corr_eps) {
                        const double np = sqrt(norm2(c, psi));
#pragma unroll
                        for (int q = 0; q < NR; ++q) {
                            ynew.v[q].x = (R)(psi.v[q].x / np);
                            ynew.v[q].y = (R)(psi.v[q].y / np);
                        }
                    } else {
#pragma unroll
                        for (int q = 0; q < NR; ++q) {
                            ynew.v[q].x = (R)(ynew.v[q].x / nn);
                            ynew.v[q].y = (R)(ynew.v[q].y / nn);
                        }
                        if (threadIdx.x == 0 && jumps < A.max_jumps) {
                            A.col_times[(size_t)traj * A.max_jumps + jumps] = t_guess;
                            A.col_which[(size_t)traj * A.max_jumps + jumps] = which;
                        }
                        jumps++;
                        if (!draw(g, target)) {
                            s.status = MC_DRAWS_EXHAUSTED;
                            break;
                        }
                    }
                    set_initial_value(c, s, ynew, t_guess);
                    t_old = s.t;
                    norm_old = 1.0;
                } else {
                    t_old = t_step;
                    norm_old = nrm;
                }
            }
            if (s.status != MC_OK) break;
            write_expect(c, s.y, traj, k);
        }
        if (threadIdx.x == 0) {
            A.info[(size_t)traj * 4 + 0] = s.status;
            A.info[(size_t)traj * 4 + 1] = jumps;
            A.info[(size_t)traj * 4 + 2] = g.used;
            A.info[(size_t)traj * 4 + 3] = (int)s.attempts;
        }
        __syncthreads();
    }
}

// Whole solves of small systems (sesolve / mesolve with n <= 1024): one
// workgroup integrates one system from t0 through every output time (the
// integrate(t) loop of the reference with interpolation), evaluating the
// functionals at each time, with no host round trip.  Systems are
// independent initial states (and optionally operators, h_stride apart),
// picked up grid-stride.  HL: the one operator of every system (h_stride
// 0) is copied into LDS after the other arrays once per workgroup, so the
// RHS products read it there instead of from L2 (the launcher picks this
// when it fits the configuration's LDS budget).
template <typename R, int NR, int RS, bool HL>
__global__ __launch_bounds__(RS == 0 ? 2 * kMcThreads : kMcThreads) void rk_small_kernel(McArgs A) {
    extern __shared__ __align__(16) unsigned char mc_dyn_lds[];
    __shared__ McTableau T;
    McLds<R> S;
    kernel_prologue<R, NR, RS>(A, mc_dyn_lds, T, S);
    Ctx<R, NR, RS> c{A, S, T};
    const int *lcol = nullptr;
    const C2<R> *lval = nullptr;
    unsigned char *tail = align16(S.rsum + 4 * (blockDim.x / QA_WAVE));
    if (A.x2) {  // the launcher reserved the second input buffer here
        S.xalt = (C2<R> *)tail;
        tail += sizeof(C2<R>) * (size_t)S.npad;
    }
    if constexpr (HL) {
        const int nent = A.h_width * A.d;
        C2<R> *v = (C2<R> *)tail;
        int *col = (int *)(tail + sizeof(C2<R>) * (size_t)nent);
        const C2<R> *gv = (const C2<R> *)A.h_vals;
        for (int i = threadIdx.x; i < nent; i += blockDim.x) {
            v[i] = gv[i];
            col[i] = A.h_col[i];
        }
        lcol = col;
        lval = v;
        tail = align16(col + nent);
    }
    // The functionals, then the output times, into the fun_lds bytes the
    // launcher reserved after the arrays above, each when it fits there
    // (the sizes are checked here, where the entry count is known).
    FunView fv{A.f_kind, A.f_ptr, A.f_row, A.f_col, A.f_w};
    const double *times = A.times;
    if (A.fun_lds > 0) {
        size_t used = 0;
        if (A.nfun > 0) {
            const int nfun = A.nfun, nnz = A.f_ptr[nfun];
            const size_t need = sizeof(double2) * (size_t)nnz +
                                sizeof(int) * (2 * (size_t)nnz + 2 * (size_t)nfun + 1);
            if (nnz >= 0 && need <= (size_t)A.fun_lds) {
                double2 *w = (double2 *)tail;
                int *ptr = (int *)(w + nnz);
                int *kind = ptr + nfun + 1;
                int *row = kind + nfun;
                int *col = row + nnz;
                for (int i = threadIdx.x; i < nnz; i += blockDim.x) {
                    w[i] = A.f_w[i];
                    row[i] = A.f_row[i];
                    col[i] = A.f_col[i];
                }
                for (int i = threadIdx.x; i <= nfun; i += blockDim.x) {
                    ptr[i] = A.f_ptr[i];
                    if (i < nfun) kind[i] = A.f_kind[i];
                }
                fv = FunView{kind, ptr, row, col, w};
                used = (size_t)(align16(col + nnz) - tail);
            }
        }
        if (used + sizeof(double) * (size_t)A.ntimes <= (size_t)A.fun_lds) {
            double *tl = (double *)(tail + used);
            for (int i = threadIdx.x; i < A.ntimes; i += blockDim.x) tl[i] = A.times[i];
            times = tl;
        }
    }
    __syncthreads();
    for (int sys = blockIdx.x; sys < A.ntraj; sys += gridDim.x) {
        const size_t hoff = (size_t)sys * (size_t)A.h_stride;
        if constexpr (HL)
            ctx_rows(c, lcol, (const void *)lval);
        else
            ctx_rows(c, A.h_col + hoff, (const void *)((const C2<R> *)A.h_vals + hoff));
        Traj<R, NR, RS> s;
        s.status = MC_OK;
        s.attempts = 0;
        s.ndense = 0;
        V<R, NR> y0;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const double2 p0 =
                c.act[q] ? A.psi0[(size_t)sys * A.d + c.row[q]] : make_double2(0.0, 0.0);
            y0.v[q] = mk<R>((R)p0.x, (R)p0.y);
        }
        set_initial_value(c, s, y0, A.t0);
        int k = 0;
        for (; k < A.ntimes && s.status >= 0; ++k) {
            integrate_to(c, s, times[k]);
            if (s.status < 0) break;
            write_functionals(c, s.y, fv, A.fun_out + ((size_t)sys * A.ntimes + k) * A.nfun);
        }
        if (A.y_final) {
#pragma unroll
            for (int q = 0; q < NR; ++q)
                if (c.act[q])
                    A.y_final[(size_t)sys * A.d + c.row[q]] =
                        make_double2((double)s.y.v[q].x, (double)s.y.v[q].y);
        }
        if (threadIdx.x == 0) {
            A.info[(size_t)sys * 4 + 0] = s.status;
            A.info[(size_t)sys * 4 + 1] = k;  // outputs written
            A.info[(size_t)sys * 4 + 2] = s.ndense;
            A.info[(size_t)sys * 4 + 3] = (int)s.attempts;
        }
        __syncthreads();
    }
}

constexpr size_t kLdsBudget = 160 * 1024;  // LDS of one gfx950 CU

constexpr size_t kFunLdsMax = 16 * 1024;  // LDS for the functionals and times

// Static LDS of both kernels (the `__shared__ McTableau T`), rounded up to
// the 16-byte granule the dynamic part starts on.  Every budget below counts
// it: a workgroup's LDS is the static part plus the dynamic bytes.
constexpr size_t kLdsStatic = (sizeof(McTableau) + 15) & ~(size_t)15;

// The LDS plan of one launch: dynamic bytes, the operator in LDS (HL), and
// the bytes reserved for functionals and times.
struct McPlan {
    size_t lds;   // dynamic LDS bytes per workgroup
    bool hl;      // the shared ELL operator is staged in LDS
    int fun_lds;  // bytes for the functionals and times (0: none)
    bool x2;      // a second input buffer (one barrier per product instead of two)
};

// RS == 0 runs two workgroups per CU, so each gets half of the CU's LDS,
// its static part included.
inline size_t mc_lds_cap(int RS) {
    return (RS == 0 ? kLdsBudget / 2 : kLdsBudget) - kLdsStatic;
}

template <typename R>
McPlan plan_mc(const McArgs &A, int threads, int NR, int RS, bool small) {
    McPlan P;
    P.lds = mc_lds_bytes<R>(threads * NR, threads, RS);
    P.hl = false;
    P.fun_lds = 0;
    P.x2 = false;
    const size_t cap = mc_lds_cap(RS);
    if (small && A.h_stride == 0 && A.h_width > 0) {
        // the shared operator in LDS when it fits the budget of this shape
        const size_t ell = (size_t)A.h_width * A.d * (sizeof(C2<R>) + sizeof(int));
        if (P.lds + ell <= cap) {
            P.lds += ell;
            P.hl = true;
        }
    }
    {
        // the second input buffer (with 16 bytes for its alignment) when it
        // fits and keeps the number of workgroups an LDS-limited CU holds
        const size_t xb = 16 + sizeof(C2<R>) * (size_t)threads * NR;
        if (P.lds + xb <= cap &&
            kLdsBudget / (P.lds + xb + kLdsStatic) == kLdsBudget / (P.lds + kLdsStatic)) {
            P.lds += xb;
            P.x2 = true;
        }
    }
    if (small) {
        if (A.nfun > 0 || A.ntimes > 0) {
            // room for the functionals and times, up to kFunLdsMax, that
            // keeps the number of workgroups an LDS-limited CU holds
            const size_t per_wg = P.lds + kLdsStatic;
            const size_t nwg = kLdsBudget / per_wg;  // >= 1: P.lds <= cap
            const size_t lim = std::min(kLdsBudget / nwg - kLdsStatic, cap);
            const size_t room = lim > P.lds ? std::min(lim - P.lds, kFunLdsMax + 16) : 0;
            if (room >= 16 + 256) {
                P.fun_lds = (int)((room - 16) & ~(size_t)15);
                P.lds += 16 + (size_t)P.fun_lds;
            }
        }
    }
    return P;
}

template <typename R, int NR, int RS, bool SMALL>
int launch_mc_cfg(const McArgs &A0, int threads, int blocks, hipStream_t s) {
    McArgs A = A0;
    const McPlan P = plan_mc<R>(A, threads, NR, RS, SMALL);
    if (P.lds > mc_lds_cap(RS)) {
        qa::set_error("trajectory kernel: %zu B of dynamic LDS exceed the %zu B budget", P.lds,
                      mc_lds_cap(RS));
        return QA_ERR_ARG;
    }
    A.fun_lds = P.fun_lds;
    A.x2 = P.x2 ? 1 : 0;
    const size_t lds = P.lds;
    const bool hl = P.hl;
    const void *fn = !SMALL ? (const void *)mc_traj_kernel<R, NR, RS>
                     : hl   ? (const void *)rk_small_kernel<R, NR, RS, true>
                            : (const void *)rk_small_kernel<R, NR, RS, false>;
    QA_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (!SMALL)
        hipLaunchKernelGGL((mc_traj_kernel<R, NR, RS>), dim3(blocks), dim3(threads), lds, s, A);
    else if (hl)
        hipLaunchKernelGGL((rk_small_kernel<R, NR, RS, true>), dim3(blocks), dim3(threads), lds,
                           s, A);
    else
        hipLaunchKernelGGL((rk_small_kernel<R, NR, RS, false>), dim3(blocks), dim3(threads), lds,
                           s, A);
    QA_CHECK_LAUNCH();
    return QA_OK;
}

// Launch shape (measured on the TFIM trajectories at d = 512 and 1024):
// (threads, rows per thread NR, register stages RS) for d rows, false when
// d does not fit.  Every stage in LDS, one row per thread, when the 18
// vectors take at most half of the LDS: two workgroups then share a CU,
// which beats keeping the step stages in registers.  Otherwise the step
// stages stay in registers, one row per thread up to 512 rows and two rows
// per thread above.
template <typename R>
bool shape_mc(int d, int *threads, int *NR, int *RS) {
    const int t1 = ((d + QA_WAVE - 1) / QA_WAVE) * QA_WAVE;
    if (t1 <= 2 * kMcThreads && mc_lds_bytes<R>(t1, t1, 0) <= mc_lds_cap(0)) {
        *threads = t1, *NR = 1, *RS = 0;
        return true;
    }
    if (t1 <= kMcThreads && mc_lds_bytes<R>(t1, t1, kRegStagesMax) <= mc_lds_cap(kRegStagesMax)) {
        *threads = t1, *NR = 1, *RS = kRegStagesMax;
        return true;
    }
    const int rows = (d + 1) / 2;
    const int t2 = ((rows + QA_WAVE - 1) / QA_WAVE) * QA_WAVE;
    if (t2 > kMcThreads ||
        mc_lds_bytes<R>(2 * t2, t2, kRegStagesMax) > mc_lds_cap(kRegStagesMax))
        return false;
    *threads = t2, *NR = 2, *RS = kRegStagesMax;
    return true;
}

template <typename R, bool SMALL = false>
int launch_mc(const McArgs &A, int blocks, hipStream_t s) {
    int threads, NR, RS;
    if (!shape_mc<R>(A.d, &threads, &NR, &RS)) {
        qa::set_error("qa_mc_run: d = %d does not fit the trajectory kernel", A.d);
        return QA_ERR_ARG;
    }
    if (NR == 1 && RS == 0) return launch_mc_cfg<R, 1, 0, SMALL>(A, threads, blocks, s);
    if (NR == 1) return launch_mc_cfg<R, 1, kRegStagesMax, SMALL>(A, threads, blocks, s);
    return launch_mc_cfg<R, 2, kRegStagesMax, SMALL>(A, threads, blocks, s);
}

}  // namespace

extern "C" {

int qa_mc_tableau_size(void) { return (int)sizeof(McTableau); }

int qa_mc_lds_plan(int d, int h_width, int nfun, int ntimes, int fp32, int small,
                   long long *out) {
    QA_REQUIRE(out != nullptr, "qa_mc_lds_plan: null output");
    QA_REQUIRE(d >= 1 && h_width >= 0 && nfun >= 0 && ntimes >= 0, "qa_mc_lds_plan: bad size");
    McArgs A;
    memset(&A, 0, sizeof(A));
    A.d = d;
    A.h_width = h_width;
    A.nfun = nfun;
    A.ntimes = ntimes;
    int threads, NR, RS;
    const bool ok = fp32 ? shape_mc<float>(d, &threads, &NR, &RS)
                         : shape_mc<double>(d, &threads, &NR, &RS);
    if (!ok) {
        qa::set_error("qa_mc_lds_plan: d = %d does not fit the trajectory kernel", d);
        return QA_ERR_ARG;
    }
    const McPlan P = fp32 ? plan_mc<float>(A, threads, NR, RS, small != 0)
                          : plan_mc<double>(A, threads, NR, RS, small != 0);
    out[0] = threads;
    out[1] = NR;
    out[2] = RS;
    out[3] = (long long)P.lds;
    out[4] = (long long)kLdsStatic;
    out[5] = P.hl ? 1 : 0;
    out[6] = P.fun_lds;
    out[7] = (long long)mc_lds_cap(RS);
    return QA_OK;
}

// Pack a tableau (host arrays as for qa_rk_create) into the device layout
// the trajectory kernel reads; `out` is a host buffer of
// qa_mc_tableau_size() bytes to be copied to the device by the caller.
int qa_mc_pack_tableau(void *out, int order, int rk_step, int rk_extra, int dorder,
                       const double *a, const double *b, const double *c, const double *e,
                       const double *bi, int interpolate) {
    QA_REQUIRE(out && a && b && c, "qa_mc_pack_tableau: null argument");
    QA_REQUIRE(rk_step >= 1 && rk_step <= rk_extra && rk_extra <= kMcStages,
               "qa_mc_pack_tableau: 1 <= rk_step <= rk_extra <= %d", kMcStages);
    QA_REQUIRE(!bi || (dorder >= 1 && dorder <= kMcOrder), "qa_mc_pack_tableau: dense order");
    McTableau *T = (McTableau *)out;
    memset(T, 0, sizeof(McTableau));
    for (int i = 0; i < rk_extra; ++i) {
        for (int j = 0; j < rk_extra; ++j) T->a[i][j] = a[(size_t)i * rk_extra + j];
        T->c[i] = c[i];
        if (bi)
            for (int j = 0; j < dorder; ++j) T->bi[i][j] = bi[(size_t)i * dorder + j];
    }
    for (int i = 0; i < rk_step; ++i) {
        T->b[i] = b[i];
        T->e[i] = e ? e[i] : 0.0;
    }
    T->order = order;
    T->rk_step = rk_step;
    T->rk_extra = (bi && interpolate) ? rk_extra : rk_step;
    T->dorder = bi ? dorder : 0;
    T->interpolate = (bi && interpolate) ? 1 : 0;
    T->adaptive = e != nullptr;
    int fsal = 1;
    for (int j = 0; j < rk_step; ++j) {
        const double x = a[(size_t)(rk_step - 1) * rk_extra + j], y = b[j];
        if (fabs(x - y) > 1e-14 + 1e-5 * fabs(y)) fsal = 0;
    }
    T->fsal = fsal;
    return QA_OK;
}

int qa_mc_run(int d, int ntraj, int ntimes, const double *times_dev, const void *psi0_dev,
              int h_width, const int *h_col, const void *h_vals, int nchan,
              const int *c_rowptr, const int *c_col, const void *c_vals, int neops,
              const int *e_rowptr, const int *e_col, const void *e_vals,
              const void *tableau_dev, const double *draws_dev, int max_draws,
              double rtol, double atol, int nsteps, double first_step, double min_step,
              double max_step, int norm_steps, double norm_t_tol, double norm_tol,
              double norm_min_step, double mc_corr_eps, int fp32, void *expect_out_dev,
              double *col_times_dev, int *col_which_dev, int max_jumps, int *info_dev,
              int grid, void *stream) {
    QA_REQUIRE(d >= 1 && d <= 2 * kMcThreads, "qa_mc_run: 1 <= d <= %d (rows held per thread)",
               2 * kMcThreads);
    QA_REQUIRE(ntraj >= 0 && ntimes >= 1, "qa_mc_run: need ntimes >= 1");
    QA_REQUIRE(nchan >= 1 && nchan <= kMcMaxChan, "qa_mc_run: 1 <= nchan <= %d", kMcMaxChan);
    QA_REQUIRE(neops >= 0 && neops <= kMcMaxEops, "qa_mc_run: neops <= %d", kMcMaxEops);
    QA_REQUIRE(max_draws >= 1 && max_jumps >= 1, "qa_mc_run: need max_draws, max_jumps >= 1");
    QA_REQUIRE(norm_steps >= 1 && nsteps >= 0, "qa_mc_run: bad step limits");
    QA_REQUIRE(h_width >= 0 && h_width % 4 == 0 && h_width <= d + 3,
               "qa_mc_run: h_width must be a multiple of 4, at most d + 3");
    QA_REQUIRE(times_dev && psi0_dev && tableau_dev && draws_dev && info_dev &&
                   expect_out_dev && col_times_dev && col_which_dev && c_rowptr &&
                   (h_width == 0 || (h_col && h_vals)),
               "qa_mc_run: null pointer");
    if (ntraj == 0) return QA_OK;
    McArgs A;
    A.d = d;
    A.ntraj = ntraj;
    A.ntimes = ntimes;
    A.nchan = nchan;
    A.neops = neops;
    A.max_draws = max_draws;
    A.max_jumps = max_jumps;
    A.traj_offset = 0;
    A.h_width = h_width;
    A.h_col = h_col;
    A.h_vals = h_vals;
    A.c_rowptr = c_rowptr;
    A.c_col = c_col;
    A.c_vals = c_vals;
    A.e_rowptr = e_rowptr;
    A.e_col = e_col;
    A.e_vals = e_vals;
    A.tab = (const McTableau *)tableau_dev;
    A.times = times_dev;
    A.psi0 = (const double2 *)psi0_dev;
    A.draws = draws_dev;
    A.rtol = rtol;
    A.atol = atol;
    A.first_step = first_step;
    A.min_step = min_step != 0.0 ? min_step : 1e-15;
    A.max_step = max_step;
    A.nsteps = nsteps;
    A.norm_steps = norm_steps;
    A.norm_t_tol = norm_t_tol;
    A.norm_tol = norm_tol;
    A.norm_min_step = norm_min_step;
    A.mc_corr_eps = mc_corr_eps;
    A.expect_out = (double2 *)expect_out_dev;
    A.col_times = col_times_dev;
    A.col_which = col_which_dev;
    A.info = info_dev;
    if (grid <= 0) grid = ntraj;
    if (grid > ntraj) grid = ntraj;
    hipStream_t s = (hipStream_t)stream;
    return fp32 ? launch_mc<float>(A, grid, s) : launch_mc<double>(A, grid, s);
}


int qa_rk_small_run(int n, int nsys, double t0, int ntimes, const double *times_dev,
                    const void *y0_dev, int h_width, const int *h_col, const void *h_vals,
                    long long h_stride, int nfun, const int *fkind, const int *fptr,
                    const int *frows, const int *fcols, const void *fw, const void *tableau_dev,
                    double rtol, double atol, int nsteps, double first_step, double min_step,
                    double max_step, int fp32, void *fun_out_dev, void *y_final_dev,
                    int *info_dev, int grid, void *stream) {
    QA_REQUIRE(n >= 1 && n <= 2 * kMcThreads, "qa_rk_small_run: 1 <= n <= %d", 2 * kMcThreads);
    QA_REQUIRE(nsys >= 0 && ntimes >= 0 && nfun >= 0, "qa_rk_small_run: negative size");
    QA_REQUIRE(h_width >= 0 && h_width % 4 == 0 && h_width <= n + 3 && h_stride >= 0,
               "qa_rk_small_run: h_width must be a multiple of 4, at most n + 3");
    QA_REQUIRE(y0_dev && tableau_dev && info_dev && (ntimes == 0 || times_dev) &&
                   (h_width == 0 || (h_col && h_vals)) &&
                   (nfun == 0 || (fkind && fptr && frows && fcols && fw && fun_out_dev)),
               "qa_rk_small_run: null pointer");
    QA_REQUIRE(nsteps >= 0, "qa_rk_small_run: nsteps >= 0");
    if (nsys == 0) return QA_OK;
    McArgs A;
    memset(&A, 0, sizeof(A));
    A.d = n;
    A.ntraj = nsys;
    A.ntimes = ntimes;
    A.h_width = h_width;
    A.h_col = h_col;
    A.h_vals = h_vals;
    A.h_stride = h_stride;
    A.tab = (const McTableau *)tableau_dev;
    A.times = times_dev;
    A.t0 = t0;
    A.psi0 = (const double2 *)y0_dev;
    A.rtol = rtol;
    A.atol = atol;
    A.first_step = first_step;
    A.min_step = min_step;
    A.max_step = max_step;
    A.nsteps = nsteps;
    A.nfun = nfun;
    A.f_kind = fkind;
    A.f_ptr = fptr;
    A.f_row = frows;
    A.f_col = fcols;
    A.f_w = (const double2 *)fw;
    A.fun_out = (double2 *)fun_out_dev;
    A.y_final = (double2 *)y_final_dev;
    A.info = info_dev;
    if (grid <= 0 || grid > nsys) grid = nsys;
    hipStream_t s = (hipStream_t)stream;
    return fp32 ? launch_mc<float, true>(A, grid, s) : launch_mc<double, true>(A, grid, s);
}

}  // extern "C"
