// Batched stochastic trajectories (ssesolve / smesolve) on gfx950.
//
// One workgroup runs one trajectory over the whole tlist (grid-stride over
// trajectories); all trajectories advance in one launch with no host round
// trip.  Restated algorithms (qutip/solver/...):
//   ssesolve system: drift = (-iH - 1/2 sum c^dag c) psi
//       + sum_i (e_i / 2 c_i psi - e_i^2 / 8 psi), e_i = <psi|c_i + c_i^dag|psi>,
//       diffusion_i = c_i psi - e_i / 2 psi        sode/ssystem.pyx:119-176
//   smesolve system: drift = L vec(rho), diffusion_i = C_i vec(rho)
//       - tr(C_i rho) vec(rho), C_i = spre(c_i) + spost(c_i^dag)
//                                                  sode/ssystem.pyx:195-250
//   Euler and Platen steps                         sode/_sode.pyx:17-120
//   steps per output interval                      sode/sode.py:155-177
// The Wiener increments are the reference's (normal(0, sqrt dt) from each
// trajectory's generator, sode/_noise.py:8-30), drawn on the host in that
// order and streamed in.
//
// Layout: thread t owns rows t, t + 256, ... (NR rows, n <= 1024); every
// vector of the step (state, d1, the diffusions, the products) stays in
// registers and only the vector a product reads is staged in LDS.  The
// drift operator L and the operators M_i (c_i for ssesolve, C_i for
// smesolve) are column-major ELL matrices (entry u of row r at u n + r) in
// HBM, shared by all trajectories and served from L2.
#include <math.h>

#include "qa_api_internal.h"
#include "qa_common.h"

using namespace qa;

namespace {

constexpr int kSdeThreads = 256;
constexpr int kSdeWaves = kSdeThreads / QA_WAVE;
constexpr int kSdeMaxSc = 4;
constexpr int kSdeMaxE = 16;
constexpr int kSdeMaxN = 1024;

enum SdeMethod { SDE_EULER = 0, SDE_PLATEN = 1 };

struct SdeArgs {
    int n, d, nsc, method, nout, ne, ntraj;
    double dt;
    long long nsteps;        // steps of one trajectory (sum of counts)
    const int *counts;       // steps of every output interval (nout - 1)
    int lw;                  // ELL width of L
    const int *lcol;
    const double2 *lval;
    int mw;                  // ELL width of every M_i (block i at i mw n)
    const int *mcol;
    const double2 *mval;
    const double *dW;        // [ntraj][nsteps][nsc]
    const double2 *x0;       // the initial state (n)
    const int *erow;         // e_ops CSR blocks: absolute row pointers, one
    const int *ecol;         //   block of rows + 1 per operator (rows = n
    const double2 *ew;       //   for ssesolve, d for smesolve)
    double2 *out;            // [ntraj][nout][ne]
    int *info;               // [ntraj]: 0, or -1 when the state stopped being finite
};

struct SdeShared {
    double2 x[kSdeMaxN];
    double red[2 * kSdeMaxE * kSdeWaves];
};

template <int NR>
struct SV {
    double2 v[NR];
};

template <int NR>
__device__ __forceinline__ SV<NR> szero() {
    SV<NR> o;
#pragma unroll
    for (int q = 0; q < NR; ++q) o.v[q] = make_double2(0.0, 0.0);
    return o;
}

__device__ __forceinline__ double2 cscale(double2 s, double2 v) { return cmul(s, v); }

// o = a + s b (complex s)
template <int NR>
__device__ __forceinline__ SV<NR> axpy(const SV<NR> &a, double2 s, const SV<NR> &b) {
    SV<NR> o;
#pragma unroll
    for (int q = 0; q < NR; ++q) o.v[q] = cadd(a.v[q], cmul(s, b.v[q]));
    return o;
}

template <int NR>
__device__ __forceinline__ SV<NR> raxpyv(const SV<NR> &a, double s, const SV<NR> &b) {
    SV<NR> o;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        o.v[q] = a.v[q];
        raxpy(o.v[q], s, b.v[q]);
    }
    return o;
}

// K block sums (deterministic: wave sums, then the waves in order).
template <int K>
__device__ __forceinline__ void block_sums(SdeShared &S, double (&v)[K]) {
    const int lane = threadIdx.x & (QA_WAVE - 1), wid = threadIdx.x / QA_WAVE;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) S.red[k * kSdeWaves + wid] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double t = 0.0;
        for (int w = 0; w < kSdeWaves; ++w) t += S.red[k * kSdeWaves + w];
        v[k] = t;
    }
}

template <int NR>
__device__ __forceinline__ void stage(SdeShared &S, const SV<NR> &x, int n) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        const int r = threadIdx.x + q * kSdeThreads;
        if (r < n) S.x[r] = x.v[q];
    }
    __syncthreads();
}

// (A x) for this thread's rows, A a column-major ELL of width w (a multiple
// of 4; padding entries have value 0 and a column inside [0, n)); x staged.
template <int NR>
__device__ __forceinline__ SV<NR> ell(const SdeShared &S, int n, int w, const int *__restrict__ col,
                                      const double2 *__restrict__ val) {
    SV<NR> acc = szero<NR>();
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        const int r = threadIdx.x + q * kSdeThreads;
        if (r >= n) continue;
        for (int u = 0; u < w; u += 4) {
            int c[4];
            double2 a[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                c[k] = col[(size_t)(u + k) * n + r];
                a[k] = val[(size_t)(u + k) * n + r];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) cfma(acc.v[q], a[k], S.x[c[k]]);
        }
    }
    return acc;
}

// The operator functional of every y_i = M_i x: ssesolve <x|c + c^dag|x> =
// 2 Re <x|y> (real), smesolve tr(C_i rho) = the sum of y's diagonal
// entries (rows k (d + 1)).
template <int NR, int NSC, bool OPEN>
__device__ __forceinline__ void functionals(SdeShared &S, int n, int d, const SV<NR> &x,
                                            const SV<NR> (&y)[NSC], double2 (&s)[NSC]) {
    double v[2 * NSC];
#pragma unroll
    for (int i = 0; i < NSC; ++i) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const int r = threadIdx.x + q * kSdeThreads;
            if (r >= n) continue;
            if (OPEN) {
                if (r % (d + 1) == 0) {
                    re += y[i].v[q].x;
                    im += y[i].v[q].y;
                }
            } else {
                re += x.v[q].x * y[i].v[q].x + x.v[q].y * y[i].v[q].y;
            }
        }
        v[2 * i] = re;
        v[2 * i + 1] = im;
    }
    block_sums<2 * NSC>(S, v);
#pragma unroll
    for (int i = 0; i < NSC; ++i)
        s[i] = OPEN ? make_double2(v[2 * i], v[2 * i + 1]) : make_double2(2.0 * v[2 * i], 0.0);
}

// y_i = M_i x for every i (x staged).
template <int NR, int NSC>
__device__ __forceinline__ void apply_m(const SdeShared &S, const SdeArgs &A, SV<NR> (&y)[NSC]) {
#pragma unroll
    for (int i = 0; i < NSC; ++i)
        y[i] = ell<NR>(S, A.n, A.mw, A.mcol + (size_t)i * A.mw * A.n,
                       A.mval + (size_t)i * A.mw * A.n);
}

// drift(x) with x staged; y_i = M_i x and their functionals are left in y
// and s (ssesolve computes them for the drift; smesolve only when asked).
template <int NR, int NSC, bool OPEN>
__device__ __forceinline__ SV<NR> drift(SdeShared &S, const SdeArgs &A, const SV<NR> &x,
                                        SV<NR> (&y)[NSC], double2 (&s)[NSC], bool with_m) {
    SV<NR> out = ell<NR>(S, A.n, A.lw, A.lcol, A.lval);
    if (!OPEN || with_m) {
        apply_m<NR, NSC>(S, A, y);
        functionals<NR, NSC, OPEN>(S, A.n, A.d, x, y, s);
    }
    if (!OPEN) {
#pragma unroll
        for (int i = 0; i < NSC; ++i) {
            const double e = s[i].x;
            out = raxpyv<NR>(out, 0.5 * e, y[i]);
            out = raxpyv<NR>(out, -0.125 * e * e, x);
        }
    }
    return out;
}

// diffusion_i = y_i - g_i x with g = e / 2 (ssesolve) or tr(C_i rho)
// (smesolve).
template <int NR, bool OPEN>
__device__ __forceinline__ SV<NR> diffusion(const SV<NR> &y, double2 s, const SV<NR> &x) {
    const double2 g = OPEN ? s : make_double2(0.5 * s.x, 0.5 * s.y);
    return axpy<NR>(y, make_double2(-g.x, -g.y), x);
}

// <O>(x) of every e_op into out[0 .. ne): <x|O|x> (ssesolve) or tr(O rho)
// (smesolve); x staged.
template <int NR, bool OPEN>
__device__ __forceinline__ void expectations(SdeShared &S, const SdeArgs &A, const SV<NR> &x,
                                             double2 *out) {
    for (int e0 = 0; e0 < A.ne; e0 += 4) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double re = 0.0, im = 0.0;
            const int e = e0 + k;
            if (e < A.ne) {
                const int rows = OPEN ? A.d : A.n;
                const int *rp = A.erow + (size_t)e * (rows + 1);
#pragma unroll
                for (int q = 0; q < NR; ++q) {
                    const int r = threadIdx.x + q * kSdeThreads;
                    if (r >= rows) continue;
                    double2 acc = make_double2(0.0, 0.0);
                    for (int p = rp[r]; p < rp[r + 1]; ++p) {
                        const int c = A.ecol[p];
                        cfma(acc, A.ew[p], S.x[OPEN ? r * A.d + c : c]);
                    }
                    if (OPEN) {
                        re += acc.x;
                        im += acc.y;
                    } else {
                        const double2 xr = x.v[q];  // conj(x_r) acc
                        re += xr.x * acc.x + xr.y * acc.y;
                        im += xr.x * acc.y - xr.y * acc.x;
                    }
                }
            }
            v[2 * k] = re;
            v[2 * k + 1] = im;
        }
        block_sums<8>(S, v);
        if (threadIdx.x == 0)
            for (int k = 0; k < 4 && e0 + k < A.ne; ++k) out[e0 + k] = make_double2(v[2 * k], v[2 * k + 1]);
    }
}

template <int NR, int NSC, bool OPEN>
__device__ __forceinline__ SV<NR> platen_step(SdeShared &S, const SdeArgs &A, const SV<NR> &x,
                                              const double (&dW)[NSC]) {
    const double dt = A.dt, sq = sqrt(dt), sqi = 0.25 / sq;
    SV<NR> y[NSC], d2[NSC];
    double2 s[NSC];
    stage<NR>(S, x, A.n);
    SV<NR> dr = drift<NR, NSC, OPEN>(S, A, x, y, s, true);
    const SV<NR> d1 = raxpyv<NR>(x, dt, dr);
#pragma unroll
    for (int i = 0; i < NSC; ++i) d2[i] = diffusion<NR, OPEN>(y[i], s[i], x);
    // out = d1 / 2 + x / 2 + sum_i dW_i / 2 d2_i ; Vt = d1 + sum_i dW_i d2_i
    SV<NR> out, Vt = d1;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
        out.v[q] = make_double2(0.5 * d1.v[q].x + 0.5 * x.v[q].x, 0.5 * d1.v[q].y + 0.5 * x.v[q].y);
    }
#pragma unroll
    for (int i = 0; i < NSC; ++i) {
        out = raxpyv<NR>(out, 0.5 * dW[i], d2[i]);
        Vt = raxpyv<NR>(Vt, dW[i], d2[i]);
    }
    stage<NR>(S, Vt, A.n);
    dr = drift<NR, NSC, OPEN>(S, A, Vt, y, s, false);
    out = raxpyv<NR>(out, 0.5 * dt, dr);
#pragma unroll
    for (int i = 0; i < NSC; ++i) {
#pragma unroll
        for (int sg = 0; sg < 2; ++sg) {
            const double sign = sg == 0 ? 1.0 : -1.0;
            const SV<NR> V = raxpyv<NR>(d1, sign * sq, d2[i]);
            stage<NR>(S, V, A.n);
            apply_m<NR, NSC>(S, A, y);
            functionals<NR, NSC, OPEN>(S, A.n, A.d, V, y, s);
#pragma unroll
            for (int j = 0; j < NSC; ++j) {
                double c;
                if (i == j) {
                    const double dw2 = sqi * (dW[i] * dW[j] - dt), dw = 0.25 * dW[i];
                    c = sg == 0 ? dw2 + dw : -dw2 + dw;
                } else {
                    c = sign * sqi * dW[i] * dW[j];
                }
                out = raxpyv<NR>(out, c, diffusion<NR, OPEN>(y[j], s[j], V));
            }
        }
    }
    return out;
}

template <int NR, int NSC, bool OPEN>
__device__ __forceinline__ SV<NR> euler_step(SdeShared &S, const SdeArgs &A, const SV<NR> &x,
                                             const double (&dW)[NSC]) {
    SV<NR> y[NSC];
    double2 s[NSC];
    stage<NR>(S, x, A.n);
    const SV<NR> dr = drift<NR, NSC, OPEN>(S, A, x, y, s, true);
    SV<NR> out = raxpyv<NR>(x, A.dt, dr);
#pragma unroll
    for (int i = 0; i < NSC; ++i) out = raxpyv<NR>(out, dW[i], diffusion<NR, OPEN>(y[i], s[i], x));
    return out;
}

template <int NR, int NSC, bool OPEN>
__global__ __launch_bounds__(kSdeThreads) void sde_traj_kernel(SdeArgs A) {
    __shared__ SdeShared S;
    for (int traj = blockIdx.x; traj < A.ntraj; traj += gridDim.x) {
        SV<NR> x;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
            const int r = threadIdx.x + q * kSdeThreads;
            x.v[q] = r < A.n ? A.x0[r] : make_double2(0.0, 0.0);
        }
        double2 *out = A.out + (size_t)traj * A.nout * A.ne;
        const double *dWt = A.dW + (size_t)traj * A.nsteps * NSC;
        stage<NR>(S, x, A.n);
        expectations<NR, OPEN>(S, A, x, out);
        long long k = 0;
        bool finite = true;
        for (int o = 1; o < A.nout; ++o) {
            const int N = A.counts[o - 1];
            for (int st = 0; st < N; ++st, ++k) {
                double dW[NSC];
#pragma unroll
                for (int i = 0; i < NSC; ++i) dW[i] = dWt[k * NSC + i];
                x = A.method == SDE_PLATEN ? platen_step<NR, NSC, OPEN>(S, A, x, dW)
                                           : euler_step<NR, NSC, OPEN>(S, A, x, dW);
            }
            stage<NR>(S, x, A.n);
            expectations<NR, OPEN>(S, A, x, out + (size_t)o * A.ne);
#pragma unroll
            for (int q = 0; q < NR; ++q)
                finite = finite && isfinite(x.v[q].x) && isfinite(x.v[q].y);
        }
        // a workgroup-wide verdict (every thread reaches this barrier)
        double bad[1] = {finite ? 0.0 : 1.0};
        block_sums<1>(S, bad);
        if (threadIdx.x == 0) A.info[traj] = bad[0] > 0.0 ? -1 : 0;
    }
}

template <int NR, int NSC>
int launch_sde(const SdeArgs &A, bool open, int grid, hipStream_t s) {
    if (open)
        hipLaunchKernelGGL((sde_traj_kernel<NR, NSC, true>), dim3(grid), dim3(kSdeThreads), 0, s, A);
    else
        hipLaunchKernelGGL((sde_traj_kernel<NR, NSC, false>), dim3(grid), dim3(kSdeThreads), 0, s, A);
    QA_CHECK_LAUNCH();
    return QA_OK;
}

template <int NR>
int launch_sde_nsc(const SdeArgs &A, bool open, int grid, hipStream_t s) {
    switch (A.nsc) {
        case 1: return launch_sde<NR, 1>(A, open, grid, s);
        case 2: return launch_sde<NR, 2>(A, open, grid, s);
        case 3: return launch_sde<NR, 3>(A, open, grid, s);
        default: return launch_sde<NR, 4>(A, open, grid, s);
    }
}

}  // namespace

extern "C" {

int qa_sde_run(int n, int d, int open, int nsc, int method, int ntraj, int nout,
               const int *counts_dev, double dt, long long nsteps, int lw, const int *lcol,
               const void *lval, int mw, const int *mcol, const void *mval, const double *dW_dev,
               const void *x0_dev, int ne, const int *erow, const int *ecol, const void *ew,
               void *out_dev, int *info_dev, int grid, void *stream) {
    QA_REQUIRE(n >= 1 && n <= kSdeMaxN, "qa_sde_run: 1 <= n <= %d", kSdeMaxN);
    QA_REQUIRE(!open || (d >= 1 && d * d == n), "qa_sde_run: smesolve needs n = d^2");
    QA_REQUIRE(open || d == n, "qa_sde_run: ssesolve needs d = n");
    QA_REQUIRE(nsc >= 1 && nsc <= kSdeMaxSc, "qa_sde_run: 1 <= nsc <= %d", kSdeMaxSc);
    QA_REQUIRE(method == SDE_EULER || method == SDE_PLATEN, "qa_sde_run: method 0 or 1");
    QA_REQUIRE(ntraj >= 0 && nout >= 1 && ne >= 0 && ne <= kSdeMaxE,
               "qa_sde_run: ntraj >= 0, nout >= 1, 0 <= ne <= %d", kSdeMaxE);
    QA_REQUIRE(dt > 0.0 && nsteps >= 0, "qa_sde_run: dt > 0");
    QA_REQUIRE(lw >= 4 && lw % 4 == 0 && mw >= 4 && mw % 4 == 0,
               "qa_sde_run: ELL widths must be positive multiples of 4");
    QA_REQUIRE(lcol && lval && mcol && mval && x0_dev && (nout < 2 || counts_dev) &&
                   (nsteps == 0 || dW_dev) && (ne == 0 || (erow && ecol && ew)) && out_dev &&
                   info_dev,
               "qa_sde_run: null buffer");
    if (ntraj == 0) return QA_OK;
    SdeArgs A;
    A.n = n;
    A.d = d;
    A.nsc = nsc;
    A.method = method;
    A.nout = nout;
    A.ne = ne;
    A.ntraj = ntraj;
    A.dt = dt;
    A.nsteps = nsteps;
    A.counts = counts_dev;
    A.lw = lw;
    A.lcol = lcol;
    A.lval = (const double2 *)lval;
    A.mw = mw;
    A.mcol = mcol;
    A.mval = (const double2 *)mval;
    A.dW = dW_dev;
    A.x0 = (const double2 *)x0_dev;
    A.erow = erow;
    A.ecol = ecol;
    A.ew = (const double2 *)ew;
    A.out = (double2 *)out_dev;
    A.info = info_dev;
    if (grid <= 0 || grid > ntraj) grid = ntraj;
    hipStream_t s = (hipStream_t)stream;
    const int nr = (n + kSdeThreads - 1) / kSdeThreads;
    if (nr <= 1) return launch_sde_nsc<1>(A, open != 0, grid, s);
    if (nr <= 2) return launch_sde_nsc<2>(A, open != 0, grid, s);
    return launch_sde_nsc<4>(A, open != 0, grid, s);
}

}  // extern "C"
