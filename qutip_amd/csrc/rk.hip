// Fused explicit Runge-Kutta engine for dy/dt = A y with a constant sparse
// generator A (the superoperator of mesolve, -iH of sesolve).
//
// Restates the algorithm of qutip/solver/integrator/explicit_rk.pyx (step
// control :333-356 and :439-467, stages :358-389, error :391-397, dense output
// :399-430) with the work laid out for gfx950:
//
// * one kernel per RK stage: the SELL-C spmv k_i = A x_i with an epilogue
//   that forms the next stage input x_{i+1} = y_prev + dt sum_j a_{i+1,j} k_j
//   row by row (the reference does a copy plus one zaxpy per coefficient per
//   stage, explicit_rk.pyx:432-437);
// * the last stage's epilogue forms y_front, the embedded error vector and
//   the weighted RMS norm partials (ode.pyx:27-52); a one-workgroup control
//   kernel reduces them in a fixed order and runs the step controller on the
//   device, so the host never waits on a step;
// * dense-output stages, the output-time interpolation and the expectation
//   values ("functionals") are further kernels gated by the device phase;
//   with one chunk per functional the last output workgroup to finish also
//   advances the output counters, so no completion launch follows;
// * one loop iteration (all of the above) is captured once into a hipGraph
//   and replayed; kernels whose phase does not match return immediately.
//
// Device state lives in an RKCtl block in HBM.  Vectors are complex128,
// padded to a multiple of 64 rows.  Operator values are complex128 or
// complex64 (widened to FP64 in registers; accumulation is always FP64).
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "qa_api_internal.h"
#include "qa_coef.h"
#include "qa_common.h"

using namespace qa;

namespace {

enum Phase { PH_IDLE = 0, PH_STEP = 1, PH_DENSE = 2, PH_OUTPUT = 3, PH_DONE = 4, PH_ERROR = 5 };
// integrator status codes: qutip/solver/integrator/explicit_rk.pxd:5-12
enum Status {
    ST_AT_FRONT = 2,
    ST_INTERPOLATED = 1,
    ST_NORMAL = 0,
    ST_TOO_MUCH_WORK = -1,
    ST_DT_UNDERFLOW = -2,
    ST_OUTSIDE_RANGE = -3,
    ST_NOT_INITIATED = -4
};

constexpr int kMaxStages = 16;
constexpr int kMaxPending = 4;     // output times handled per output launch
constexpr int kOutChunk = 16;      // functional entries per output workgroup
constexpr int kOutThreads = 64;    // one wavefront per output workgroup
constexpr int kStepBlocks = 256;   // threads per stage workgroup
constexpr int kMaxChunk = 64;      // loop iterations per host chunk
constexpr int kGraphBatch = 8;     // loop iterations per batch graph
constexpr int kBatchAfter = 128;   // iterations an engine runs before it batches

struct RKCtl {
    double t_prev, t_front, dt_safe, dt_int, dt_try, err, t_last;
    double rtol, atol, min_step, max_step;
    double *times;  // mutable copy of the requested output times
    double2 *fun_out;
    double2 *states_out;
    int ntimes, phase, cur, out_idx, nsteps_left, nsteps_call, dense_ready, status;
    int order, max_numsteps, interpolate, adaptive, single_step, nfun;
    long long n_attempts, n_accepted, n_dense, n_outputs;
};

// ---- step control (executed by one thread) ---------------------------------
__device__ double get_timestep(const RKCtl *c, double t) {
    // explicit_rk.pyx:439-449
    const double dt_needed = t - c->t_prev;
    if (!c->adaptive) return dt_needed;
    if (c->interpolate) return c->dt_safe;
    if (dt_needed <= c->dt_safe) return dt_needed;
    return dt_needed / (double)((long long)(dt_needed / c->dt_safe) + 1);
}

__device__ void recompute_safe_step(RKCtl *c, double err, double dt) {
    // explicit_rk.pyx:451-467
    if (!c->adaptive) return;
    double factor;
    if (err == 0.0) {
        factor = 10.0;
    } else {
        factor = 0.9 * pow(err, -1.0 / (c->order + 1));
        factor = fmin(10.0, factor);
        factor = fmax(0.2, factor);
    }
    c->dt_safe = dt * factor;
    if (c->max_step != 0.0) c->dt_safe = fmin(c->max_step, c->dt_safe);
    if (c->min_step != 0.0) c->dt_safe = fmax(c->min_step, c->dt_safe);
}

// y_prev <- y_front, t_prev <- t_front, then a fresh _step_in_err
// (explicit_rk.pyx:312-316).
__device__ void start_step(RKCtl *c, double target) {
    c->cur ^= 1;
    c->t_prev = c->t_front;
    c->dense_ready = 0;
    c->nsteps_call = 0;
    c->dt_try = get_timestep(c, target);
    c->phase = PH_STEP;
}

// Decide what the next output time needs (explicit_rk.pyx:288-331).
__device__ void route(RKCtl *c) {
    if (c->out_idx >= c->ntimes) {
        c->phase = PH_DONE;
        return;
    }
    double target = c->times[c->out_idx];
    if (target < c->t_prev) {
        c->status = ST_OUTSIDE_RANGE;
        c->phase = PH_ERROR;
        return;
    }
    if (target < c->t_front) {
        if (!c->interpolate) {
            c->status = ST_OUTSIDE_RANGE;
            c->phase = PH_ERROR;
            return;
        }
        c->phase = c->dense_ready ? PH_OUTPUT : PH_DENSE;
        return;
    }
    if (c->single_step && c->t_last < c->t_front && target > c->t_front) {
        // integrate(t, step=True) clamps to the front (explicit_rk.pyx:309-311)
        c->times[c->out_idx] = c->t_front;
        target = c->t_front;
    }
    if (target == c->t_front) {
        c->phase = PH_OUTPUT;
        return;
    }
    c->nsteps_left = c->max_numsteps;
    c->status = ST_NORMAL;
    start_step(c, target);
}

// After one attempt of _compute_step (explicit_rk.pyx:333-356).
__device__ void after_attempt(RKCtl *c, double err) {
    const double dt = c->dt_try;
    c->err = err;
    c->n_attempts++;
    c->dt_int = dt;
    recompute_safe_step(c, err, dt);
    if (dt == c->min_step && err > 1.0) {
        c->status = ST_DT_UNDERFLOW;
        c->phase = PH_ERROR;
        return;
    }
    c->nsteps_call++;
    if (c->nsteps_call > c->nsteps_left) {
        c->status = ST_TOO_MUCH_WORK;
        c->phase = PH_ERROR;
        return;
    }
    double target = c->times[c->out_idx];
    if (err >= 1.0) {
        c->dt_try = get_timestep(c, target);
        c->phase = PH_STEP;
        return;
    }
    c->t_front = c->t_prev + dt;
    c->nsteps_left -= c->nsteps_call;
    c->nsteps_call = 0;
    c->n_accepted++;
    c->dense_ready = 0;
    if (c->single_step && c->t_front < target) {
        // integrate(t, step=True) stops after one step at the front
        c->times[c->out_idx] = c->t_front;
        target = c->t_front;
    }
    if (c->t_front < target) {
        start_step(c, target);
    } else if (c->t_front > target) {
        if (c->interpolate) {
            c->phase = PH_DENSE;
        } else {
            c->status = ST_OUTSIDE_RANGE;
            c->phase = PH_ERROR;
        }
    } else {
        c->phase = PH_OUTPUT;
    }
}

// ---- kernels -----------------------------------------------------------------
enum Input { IN_YPREV = 0, IN_YFRONT = 1, IN_X = 2 };
enum Epi { EPI_NONE = 0, EPI_NEXT = 1, EPI_FINAL = 2 };

// Fields of a stage launch shared by the sparse and the dense stage kernels:
// where the stage vectors live and the stage's combine coefficients.
struct EpiArgs {
    int nrows;
    int64_t npad;
    double2 *y0, *y1, *x0, *x1, *K;
    double *partials;
    RKCtl *ctl;
    int stage, run_phase, input, epi, nterms;
    double cstage;                 // the stage's time is t_prev + cstage dt
    double coef_self, ecoef_self;  // coefficients of this stage's own k
    // the earlier stages with a non-zero coefficient, compacted: k_{tj[t]}
    // enters the next input / y_front with tc[t] and the error with te[t]
    int nt;
    int tj[kMaxStages];
    double tc[kMaxStages], te[kMaxStages];
};

template <typename V>
struct StageArgs : EpiArgs {
    const int *slice_ptr;
    const int *col;
    const V *vals;
    int nslices;
    // coded operator (FORM_CODED, FORM_CODED_UNIFORM): 16-bit codes packed
    // four to a word, the value and column-offset tables
    const uint2 *code;
    const double2 *vtab;
    const int *otab;
    int vbits;
    int uwords;  // > 0: every slice uwords words per lane (FORM_CODED_UNIFORM)
    // time-dependent generator (rk_stage_kernel<.., .., .., true>): the
    // coefficient programs and their value-table ranges
    CoefSet td;
};

// Operator forms of rk_stage_kernel.
enum StageForm { FORM_PLAIN = 0, FORM_CODED = 1, FORM_CODED_UNIFORM = 2 };

// The vectors a stage reads and writes, resolved from the control block.
struct StageVecs {
    double2 *yprev, *yfront;
    const double2 *xin;
    double dt;
};

__device__ __forceinline__ StageVecs stage_vecs(const EpiArgs &P, const RKCtl *ctl) {
    StageVecs v;
    const int cur = ctl->cur;
    v.yprev = cur ? P.y1 : P.y0;
    v.yfront = cur ? P.y0 : P.y1;
    v.dt = (P.run_phase == PH_STEP) ? ctl->dt_try : ctl->dt_int;
    v.xin = P.input == IN_YPREV ? v.yprev : P.input == IN_YFRONT ? v.yfront
                                                                   : ((P.stage & 1) ? P.x1 : P.x0);
    return v;
}

// Epilogue, part 1 (before the product is known): o = y_prev + dt sum_{j <
// stage} coef_j k_j and er = dt sum_{j < stage} ecoef_j k_j for element r,
// over the compacted non-zero terms.  The term count is uniform; it selects
// a body of exactly NT unconditional loads, so all of them are in flight at
// once (a load behind a branch would wait for its own data before the next).
template <int NT>
__device__ __forceinline__ void epi_terms(const EpiArgs &P, double dt, int64_t r, double2 &o,
                                          double2 &er) {
    double2 kv[NT > 0 ? NT : 1];
#pragma unroll
    for (int t = 0; t < NT; ++t) kv[t] = P.K[(int64_t)P.tj[t] * P.npad + r];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        if (P.tc[t] != 0.0) raxpy(o, dt * P.tc[t], kv[t]);
        if (P.te[t] != 0.0) raxpy(er, dt * P.te[t], kv[t]);
    }
}

__device__ __forceinline__ void epi_pre(const EpiArgs &P, const StageVecs &v, int64_t r,
                                        double2 &o, double2 &er) {
    o = make_double2(0.0, 0.0);
    er = make_double2(0.0, 0.0);
    if (P.epi == EPI_NONE) return;
    o = v.yprev[r];
    switch (P.nt) {
        case 0: break;
        case 1: epi_terms<1>(P, v.dt, r, o, er); break;
        case 2: epi_terms<2>(P, v.dt, r, o, er); break;
        case 3: epi_terms<3>(P, v.dt, r, o, er); break;
        case 4: epi_terms<4>(P, v.dt, r, o, er); break;
        case 5: epi_terms<5>(P, v.dt, r, o, er); break;
        case 6: epi_terms<6>(P, v.dt, r, o, er); break;
        case 7: epi_terms<7>(P, v.dt, r, o, er); break;
        case 8: epi_terms<8>(P, v.dt, r, o, er); break;
        case 9: epi_terms<9>(P, v.dt, r, o, er); break;
        case 10: epi_terms<10>(P, v.dt, r, o, er); break;
        case 11: epi_terms<11>(P, v.dt, r, o, er); break;
        case 12: epi_terms<12>(P, v.dt, r, o, er); break;
        case 13: epi_terms<13>(P, v.dt, r, o, er); break;
        case 14: epi_terms<14>(P, v.dt, r, o, er); break;
        default: epi_terms<15>(P, v.dt, r, o, er); break;
    }
}

// Epilogue, part 2: store k_stage, then the next stage input or y_front and
// the squared weighted error of element r (returned).
__device__ __forceinline__ double epi_post(const EpiArgs &P, const RKCtl *ctl, const StageVecs &v,
                                           int64_t r, double2 acc, double2 o, double2 er) {
    P.K[(int64_t)P.stage * P.npad + r] = acc;
    const double fb = v.dt * P.coef_self;
    const double fe = v.dt * P.ecoef_self;
    if (P.epi == EPI_NEXT) {
        if (fb != 0.0) raxpy(o, fb, acc);
        double2 *xout = ((P.stage + 1) & 1) ? P.x1 : P.x0;
        xout[r] = o;
    } else if (P.epi == EPI_FINAL) {
        if (fb != 0.0) raxpy(o, fb, acc);
        if (fe != 0.0) raxpy(er, fe, acc);
        v.yfront[r] = o;
        if (r < P.nrows) {
            const double q = sqrt(cabs2(er)) / (ctl->atol + ctl->rtol * sqrt(cabs2(o)));
            return q * q;
        }
    }
    return 0.0;
}

template <typename V, int L, int FORM, bool TD>
__global__ __launch_bounds__(kStepBlocks) void rk_stage_kernel(StageArgs<V> P) {
    constexpr bool CODED = FORM != FORM_PLAIN;
    static_assert(CODED || !TD, "a time-dependent generator needs the coded form");
    RKCtl *ctl = P.ctl;
    __shared__ double2 vt[CODED ? kCodeTabMax : 1];
    __shared__ int ot[CODED ? kCodeTabMax : 1];
    static_assert(kStepBlocks == 256, "code_tables_to_lds assumes 256 threads");
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int slice = (int)(t / QA_WAVE), lane = (int)(t % QA_WAVE);
    const int64_t r = (int64_t)slice * (QA_WAVE / L) + lane / L;  // the row of this lane
    // The tables of a coded operator (straight into LDS), the slice bounds
    // (or, for uniform slices, the code words themselves) and the control
    // block are all requested before anything waits, so their round trips
    // overlap.  The barrier comes before the phase branch: no wave exits
    // with a load into LDS outstanding.
    if constexpr (CODED) code_tables_to_lds(P.vtab, P.otab, P.vbits, vt, ot);
    const int sl = slice < P.nslices ? slice : P.nslices - 1;
    int sp0 = 0, sp1 = 0;
    uint2 cw[kUWordsMax];
    if constexpr (FORM == FORM_CODED_UNIFORM) {
        const int wb = sl * P.uwords * QA_WAVE + lane;
#pragma unroll
        for (int u = 0; u < kUWordsMax; ++u) cw[u] = P.code[wb + min(u, P.uwords - 1) * QA_WAVE];
    } else {
        sp0 = P.slice_ptr[sl];
        sp1 = P.slice_ptr[sl + 1];
    }
    const int phase = ctl->phase;
    const StageVecs v = stage_vecs(P, ctl);
    double t_stage = 0.0;
    if constexpr (TD) t_stage = ctl->t_prev + P.cstage * v.dt;
    if constexpr (CODED) __syncthreads();
    if (phase != P.run_phase) return;  // uniform over the whole grid
    if constexpr (TD) {
        // A(t) = A_0 + sum_e c_e(t) A_e: scale each element's table entries
        // by its coefficient at this stage's time
        __shared__ double2 cf[kMaxCoef];
        coef_scale_table(P.td, t_stage, vt, cf);
    }
    const int base = sp0, steps = (sp1 - sp0) / QA_WAVE;
    double part = 0.0;
    if (slice < P.nslices) {
        // Epilogue terms that do not depend on this stage's product are formed
        // first, so their loads are in flight with the product's.
        const bool owner = (lane % L == 0);
        double2 o = make_double2(0.0, 0.0), er = make_double2(0.0, 0.0);
        if (owner) epi_pre(P, v, r, o, er);
        double2 acc;
        const CodeSplit cs = {(1u << P.vbits) - 1u, P.vbits};
        if constexpr (FORM == FORM_CODED_UNIFORM)
            acc = sell_coded_row_dot_uniform<L>(cw, P.uwords, vt, ot, cs, v.xin, (int)r,
                                                P.nrows - 1);
        else if constexpr (FORM == FORM_CODED)
            acc = sell_coded_row_dot<L>(P.code, vt, ot, cs, v.xin, base, steps, lane, (int)r,
                                        P.nrows - 1);
        else
            acc = sell_row_dot<L>(P.col, P.vals, v.xin, base, steps, lane);
        if (owner) part = epi_post(P, ctl, v, r, acc, o, er);
    }
    if (P.epi == EPI_FINAL) {
        __shared__ double smem[kStepBlocks / QA_WAVE];
        const double s = block_sum(part, smem);
        if (threadIdx.x == 0) P.partials[blockIdx.x] = s;
    }
}

// The value table of a time-dependent coded generator at time t: vt_out =
// vtab with each element's entries scaled by its coefficient (one
// workgroup; for the ungated products at a given time -- the first-step
// estimate and apply).
__global__ __launch_bounds__(256) void coef_table_kernel(CoefSet S, double t,
                                                         const double2 *__restrict__ vtab,
                                                         double2 *__restrict__ vt_out, int nv) {
    __shared__ double2 cf[kMaxCoef];
    if ((int)threadIdx.x < S.ncoef) cf[threadIdx.x] = coef_eval(S, threadIdx.x, t);
    __syncthreads();
    for (int i = threadIdx.x; i < nv; i += blockDim.x) {
        double2 v = vtab[i];
        if (i >= S.vstart[0] && i < S.vstart[S.ncoef]) {
            int e = 0;
            while (e + 1 < S.ncoef && i >= S.vstart[e + 1]) ++e;
            v = cmul(v, cf[e]);
        }
        vt_out[i] = v;
    }
}

// Coefficient programs at many times: out[e * ntimes + k] = c_e(times[k])
// (data-layer evaluation of a compiled coefficient set).
__global__ __launch_bounds__(256) void coef_eval_kernel(CoefSet S, int ntimes,
                                                        const double *__restrict__ times,
                                                        double2 *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)S.ncoef * ntimes) return;
    const int e = (int)(i / ntimes), k = (int)(i % ntimes);
    out[i] = coef_eval(S, e, times[k]);
}

// ---- banded (DIA) Lindblad right-hand side ----------------------------------
// The matrix-form RHS k = Hm X + X Hp + sum_c C_c X Cd_c (Hm = -i H_nh,
// Hp = i H_nh^dag, Cd_c = C_c^dag) for banded operators, as LindbladMatrix-
// Form computes it with Dia products (lindblad_matrix_form.pyx:105-203 over
// matmul_dia_dense_dense, matmul.pyx:429-507), in FP64.  Every product is a
// list of terms acc += A[i, i+o] x[i+o, j-p] B[j-p, j]: a diagonal of Hm
// (B = 1), of Hp (A = 1), or a pair of diagonals of C_c and Cd_c.  Values
// are rows of `val` (row d: the diagonal's entries by row index; the last
// row all ones).  One thread per element (i, j) of the column-major N x N
// state, i fastest: the x and A loads of a wavefront are coalesced and its
// B load is one value.  Terms go in chunks of at most eight, each a body of
// exactly its length, so a chunk's loads are issued together; out-of-range
// neighbours are clamped and their terms dropped by a select.
constexpr int kDiaMaxTerms = 96;

struct DiaTerm {
    int o, p, a, b;
};

struct DiaOp {
    int N, nterms;
    const double2 *val;  // [rows][N]
    DiaTerm t[kDiaMaxTerms];
};

template <int M>
__device__ __forceinline__ void dia_terms(const DiaOp &D, const double2 *__restrict__ x, int i,
                                          int j, int t0, double2 &acc) {
    const int N = D.N;
    double2 av[M], bv[M], xv[M];
    bool ok[M];
#pragma unroll
    for (int u = 0; u < M; ++u) {
        const DiaTerm T = D.t[t0 + u];
        const int ii = i + T.o, jj = j - T.p;
        ok[u] = (unsigned)ii < (unsigned)N && (unsigned)jj < (unsigned)N;
        const int iic = ok[u] ? ii : i, jjc = ok[u] ? jj : j;
        av[u] = D.val[(int64_t)T.a * N + i];
        bv[u] = D.val[(int64_t)T.b * N + jjc];
        xv[u] = x[(int64_t)jjc * N + iic];
    }
#pragma unroll
    for (int u = 0; u < M; ++u) {
        const double2 t = cmul(cmul(av[u], xv[u]), bv[u]);
        if (ok[u]) acc = cadd(acc, t);
    }
}

__device__ __forceinline__ double2 dia_rhs_at(const DiaOp &D, const double2 *__restrict__ x,
                                              int i, int j) {
    double2 acc = make_double2(0.0, 0.0);
    int t = 0;
    for (; t + 8 <= D.nterms; t += 8) dia_terms<8>(D, x, i, j, t, acc);
    switch (D.nterms - t) {
        case 7: dia_terms<7>(D, x, i, j, t, acc); break;
        case 6: dia_terms<6>(D, x, i, j, t, acc); break;
        case 5: dia_terms<5>(D, x, i, j, t, acc); break;
        case 4: dia_terms<4>(D, x, i, j, t, acc); break;
        case 3: dia_terms<3>(D, x, i, j, t, acc); break;
        case 2: dia_terms<2>(D, x, i, j, t, acc); break;
        case 1: dia_terms<1>(D, x, i, j, t, acc); break;
        default: break;
    }
    return acc;
}

// One RK stage of the banded matrix form: the product for element r plus
// the stage epilogue (the sparse stage kernel's, element for element).
__global__ __launch_bounds__(256) void rk_dia_stage_kernel(EpiArgs P, DiaOp D) {
    RKCtl *ctl = P.ctl;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int phase = ctl->phase;
    const StageVecs v = stage_vecs(P, ctl);
    if (phase != P.run_phase) return;  // uniform over the whole grid
    double part = 0.0;
    if (r < P.nrows) {
        double2 o, er;
        epi_pre(P, v, r, o, er);
        const double2 acc = dia_rhs_at(D, v.xin, (int)(r % D.N), (int)(r / D.N));
        part = epi_post(P, ctl, v, r, acc, o, er);
    }
    if (P.epi == EPI_FINAL) {
        __shared__ double smem[256 / QA_WAVE];
        const double sum = block_sum(part, smem);
        if (threadIdx.x == 0) P.partials[blockIdx.x] = sum;
    }
}

// out = L(x) (ungated): the derivative for the first-step estimate and apply.
__global__ __launch_bounds__(256) void dia_rhs_kernel(DiaOp D, const double2 *__restrict__ x,
                                                      double2 *__restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < (int64_t)D.N * D.N) out[r] = dia_rhs_at(D, x, (int)(r % D.N), (int)(r / D.N));
}

// ---- dense Lindblad right-hand side on MFMA ----------------------------------
// drho/dt = -i H_nh rho + i rho H_nh^dag + sum_c c rho c^dag with
// H_nh = H - i/2 sum c^dag c (qutip/core/cy/lindblad_matrix_form.pyx:92-97,
// 189-201), rho column-major N x N (the vector buffers of the engine).
// Per stage: T_c = c X for every collapse operator, then one fused kernel
// k = Hm X + X Hp + sum_c T_c Cd_c with Hm = -i H_nh, Hp = i H_nh^dag,
// Cd_c = c^dag, followed by the stage epilogue.  Complex products use the
// real 4M decomposition on 16x16 MFMA tiles (one wavefront per tile):
// FP32 inputs and accumulation (v_mfma_f32_16x16x4f32), or each FP32 input
// split into bf16 pieces on v_mfma_f32_16x16x32_bf16 with FP32 accumulation:
// hi + lo with the three significant products (bf16 x3, about 16 bits), or
// hi + mid + lo with six (bf16 x6, FP32 accuracy).
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

struct DenseOp {
    int N, nc, prec;  // prec 0: FP32 MFMA, 1: bf16 x3 split, 2: bf16 x6 split
    int nsplit;       // waves per 16 x 16 tile (split K) of the full-RHS kernels
    int hsplit;       // the same for the Hermitian kernels (32-wide K steps)
    int tiled;        // 1: 32 x 32 workgroup tiles staged through LDS (large N)
    int herm;         // 1: Hermitian stage inputs, the upper tiles and their mirror
    const float2 *Hm, *Hp, *C, *Cd;  // N x N each; C / Cd: nc blocks
    float2 *T;                        // nc blocks of scratch
    float2 *Xf;                       // the stage input in complex64 (tiled path)
};

__device__ __forceinline__ float2 tof(double2 a) { return make_float2((float)a.x, (float)a.y); }
__device__ __forceinline__ float2 tof(float2 a) { return a; }

// re + i im += A[i0:i0+16, k0:k1] B[k0:k1, j0:j0+16]  (FP32 MFMA).
// MFMA u of a 16-wide K block takes k = kk + 4 kq + u on lane (r, kq), so
// each lane reads four consecutive rows of its B column (one contiguous
// 32- or 64-byte piece) and the 16 lanes of a row group read 16 consecutive
// A rows.  The operands of the next K block are loaded before the MFMAs of
// the current one (register double buffering), so one memory latency is
// exposed per product instead of one per block.  The prefetch index is
// clamped to the last block, so the loads need no branch.
template <typename TA, typename TB>
__device__ __forceinline__ void zmma_f32(v4f &re, v4f &im, const TA *__restrict__ A,
                                         const TB *__restrict__ B, int N, int i0, int j0,
                                         int lane, int k0, int k1) {
    const int r = lane & 15, kq = lane >> 4;
    const TA *pa = A + (size_t)(4 * kq) * N + i0 + r;
    const TB *pb = B + (size_t)(j0 + r) * N + 4 * kq;
    // two accumulator chains per part (four independent MFMA chains), added
    // at the end: back-to-back MFMAs on one accumulator wait for each other
    v4f re2 = {0.f, 0.f, 0.f, 0.f}, im2 = {0.f, 0.f, 0.f, 0.f};
    float2 a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        a[u] = tof(pa[(size_t)(k0 + u) * N]);
        b[u] = tof(pb[k0 + u]);
    }
    for (int kk = k0; kk < k1; kk += 16) {
        const int kn = min(kk + 16, k1 - 16);
        float2 an[4], bn[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            an[u] = tof(pa[(size_t)(kn + u) * N]);
            bn[u] = tof(pb[kn + u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            re = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].x, b[u].x, re, 0, 0, 0);
            re2 = __builtin_amdgcn_mfma_f32_16x16x4f32(-a[u].y, b[u].y, re2, 0, 0, 0);
            im = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].x, b[u].y, im, 0, 0, 0);
            im2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].y, b[u].x, im2, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            a[u] = an[u];
            b[u] = bn[u];
        }
    }
    re += re2;
    im += im2;
}

// x as NP bf16 pieces x = p_0 + p_1 + ... (+ a remainder below the last
// piece): each piece rounds what the previous ones left, and the remainder
// x - (p_0 + ... + p_q) is exact in FP32, so two pieces keep 16 bits of
// x's mantissa and three keep all 24 (Henry, Tang & Heinecke, ARITH 2019).
template <int NP>
__device__ __forceinline__ void split_bf16(float x, int u, v8bf (&p)[NP]) {
    float r = x;
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const __bf16 h = (__bf16)r;
        p[q][u] = h;
        r -= (float)h;
    }
}

template <int NP>
__device__ __forceinline__ void negate_bf16(const v8bf (&p)[NP], v8bf (&n)[NP]) {
#pragma unroll
    for (int q = 0; q < NP; ++q) n[q] = -p[q];
}

// acc += sum of the piece products a_i b_j with i + j < NP, the largest
// terms first: NP = 2 is bf16 x3 (hi hi, hi lo, lo hi), NP = 3 is bf16 x6
// (the dropped products are below 2^-24 relative: FP32 accuracy).
template <int NP>
__device__ __forceinline__ v4f mma_pieces(const v8bf (&a)[NP], const v8bf (&b)[NP], v4f acc) {
#pragma unroll
    for (int s = 0; s < NP; ++s)
#pragma unroll
        for (int i = 0; i <= s; ++i)
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[s - i], acc, 0, 0, 0);
    return acc;
}

// Same product as zmma_f32 with the inputs split into NP bf16 pieces
// (32-wide K blocks; lane (r, kq) holds k = kk + 8 kq + u, u < 8, so its B
// piece is contiguous).  The next block's operands are prefetched as in
// zmma_f32.
template <int NP, typename TA, typename TB>
__device__ __forceinline__ void zmma_bf16(v4f &re, v4f &im, const TA *__restrict__ A,
                                          const TB *__restrict__ B, int N, int i0, int j0,
                                          int lane, int k0, int k1) {
    const int r = lane & 15, kq = lane >> 4;
    const TA *pa = A + (size_t)(8 * kq) * N + i0 + r;
    const TB *pb = B + (size_t)(j0 + r) * N + 8 * kq;
    v4f re2 = {0.f, 0.f, 0.f, 0.f}, im2 = {0.f, 0.f, 0.f, 0.f};
    float2 a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        a[u] = tof(pa[(size_t)(k0 + u) * N]);
        b[u] = tof(pb[k0 + u]);
    }
    for (int kk = k0; kk < k1; kk += 32) {
        const int kn = min(kk + 32, k1 - 32);
        float2 an[8], bn[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            an[u] = tof(pa[(size_t)(kn + u) * N]);
            bn[u] = tof(pb[kn + u]);
        }
        v8bf ar[NP], ai[NP], nai[NP], br[NP], bi[NP];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            split_bf16<NP>(a[u].x, u, ar);
            split_bf16<NP>(a[u].y, u, ai);
            split_bf16<NP>(b[u].x, u, br);
            split_bf16<NP>(b[u].y, u, bi);
        }
        negate_bf16<NP>(ai, nai);
        // re += ar br - ai bi ; im += ar bi + ai br
        re = mma_pieces<NP>(ar, br, re);
        re2 = mma_pieces<NP>(nai, bi, re2);
        im = mma_pieces<NP>(ar, bi, im);
        im2 = mma_pieces<NP>(ai, br, im2);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            a[u] = an[u];
            b[u] = bn[u];
        }
    }
    re += re2;
    im += im2;
}

// The K range of wave w of a workgroup of `nsplit` waves: the tile's sum is
// split over the waves (more waves in flight, shorter serial loops) and the
// partial tiles are added through LDS in wave order (deterministic).
__device__ __forceinline__ void k_range(int N, int nsplit, int w, int &k0, int &k1) {
    const int len = N / nsplit;
    k0 = w * len;
    k1 = k0 + len;
}

template <typename TA, typename TB>
__device__ __forceinline__ void zmma(const DenseOp &D, v4f &re, v4f &im, const TA *A, const TB *B,
                                     int i0, int j0, int lane, int k0, int k1) {
    if (D.prec == 1)
        zmma_bf16<2>(re, im, A, B, D.N, i0, j0, lane, k0, k1);
    else if (D.prec == 2)
        zmma_bf16<3>(re, im, A, B, D.N, i0, j0, lane, k0, k1);
    else
        zmma_f32(re, im, A, B, D.N, i0, j0, lane, k0, k1);
}

// Element (i0 + 4 (lane / 16) + v, j0 + lane % 16) of a 16 x 16 MFMA tile
// (the accumulator layout of the 16x16 instructions), as a column-major
// index.
__device__ __forceinline__ int64_t tile_index(int N, int i0, int j0, int lane, int v) {
    return (int64_t)(j0 + (lane & 15)) * N + i0 + 4 * (lane >> 4) + v;
}

constexpr int kMaxSplit = 8;

// Sum the per-wave partial tiles into wave 0 (fixed order), via LDS.
__device__ __forceinline__ void reduce_split(v4f &re, v4f &im, int lane, int w, int nsplit,
                                             float *lds) {
    if (nsplit == 1) return;
    __syncthreads();
    if (w > 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            lds[((w - 1) * 8 + v) * 64 + lane] = re[v];
            lds[((w - 1) * 8 + 4 + v) * 64 + lane] = im[v];
        }
    }
    __syncthreads();
    if (w == 0) {
        for (int q = 1; q < nsplit; ++q) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                re[v] += lds[((q - 1) * 8 + v) * 64 + lane];
                im[v] += lds[((q - 1) * 8 + 4 + v) * 64 + lane];
            }
        }
    }
}

// T_c = c X for every collapse operator (grid: N/16, N/16, nc; nsplit waves).
__device__ __forceinline__ void dense_t_tile(const DenseOp &D, const double2 *X) {
    __shared__ float lds[(kMaxSplit - 1) * 8 * 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nsplit = blockDim.x >> 6;
    const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16, c = blockIdx.z;
    const size_t NN = (size_t)D.N * D.N;
    int k0, k1;
    k_range(D.N, nsplit, w, k0, k1);
    v4f re = {0.f, 0.f, 0.f, 0.f}, im = {0.f, 0.f, 0.f, 0.f};
    zmma(D, re, im, D.C + c * NN, X, i0, j0, lane, k0, k1);
    reduce_split(re, im, lane, w, nsplit, lds);
    if (w == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
            D.T[c * NN + tile_index(D.N, i0, j0, lane, v)] = make_float2(re[v], im[v]);
    }
}

// k tile = Hm X + X Hp + sum_c T_c Cd_c; the result is complete in wave 0.
__device__ __forceinline__ void dense_k_tile(const DenseOp &D, const double2 *X, v4f &re,
                                             v4f &im) {
    __shared__ float lds[(kMaxSplit - 1) * 8 * 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nsplit = blockDim.x >> 6;
    const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
    const size_t NN = (size_t)D.N * D.N;
    int k0, k1;
    k_range(D.N, nsplit, w, k0, k1);
    zmma(D, re, im, D.Hm, X, i0, j0, lane, k0, k1);
    zmma(D, re, im, X, D.Hp, i0, j0, lane, k0, k1);
    for (int c = 0; c < D.nc; ++c)
        zmma(D, re, im, D.T + c * NN, D.Cd + c * NN, i0, j0, lane, k0, k1);
    reduce_split(re, im, lane, w, nsplit, lds);
}

__global__ __launch_bounds__(64 * kMaxSplit) void rk_dense_t_kernel(EpiArgs P, DenseOp D) {
    if (P.ctl->phase != P.run_phase) return;
    const StageVecs v = stage_vecs(P, P.ctl);
    dense_t_tile(D, v.xin);
}

__global__ __launch_bounds__(64 * kMaxSplit) void rk_dense_stage_kernel(EpiArgs P, DenseOp D) {
    RKCtl *ctl = P.ctl;
    if (ctl->phase != P.run_phase) return;
    const StageVecs v = stage_vecs(P, ctl);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
    double2 o[4], er[4];
    if (w == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) epi_pre(P, v, tile_index(D.N, i0, j0, lane, q), o[q], er[q]);
    }
    v4f re = {0.f, 0.f, 0.f, 0.f}, im = {0.f, 0.f, 0.f, 0.f};
    dense_k_tile(D, v.xin, re, im);
    if (w != 0) return;
    double part = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        part += epi_post(P, ctl, v, tile_index(D.N, i0, j0, lane, q),
                         make_double2((double)re[q], (double)im[q]), o[q], er[q]);
    if (P.epi == EPI_FINAL) {
        part = wave_sum(part);
        if (lane == 0) P.partials[blockIdx.y * gridDim.x + blockIdx.x] = part;
    }
}

// ---- large N: 32 x 32 workgroup tiles through LDS ------------------------
// A workgroup of 4 waves owns a 32 x 32 output tile (each wave one 16 x 16
// MFMA tile).  K goes in chunks of 64: the 256 threads load the A chunk
// (32 rows x 64 k, one 256-byte column piece per k) and the B chunk (64 k x
// 32 columns, one 512-byte piece per column) into LDS, the next chunk's loads are issued before the current
// chunk's MFMAs (one chunk of register prefetch, double-buffered LDS), and
// every A and B element loaded serves two waves.  All products of a stage
// run through one flattened chunk loop, so one memory latency is exposed
// per stage.  Used for N >= kTiledN, where the one-tile-per-workgroup
// kernels above re-read every operand row and column 32 times from L2.
constexpr int kGT = 32;      // output tile edge per workgroup
constexpr int kGK = 64;      // K chunk
constexpr int kGP = 2;       // LDS row padding (float2) against bank conflicts
constexpr int kTiledN = 384;

struct GemmTile {
    float2 a[2][kGK][kGT + kGP];  // a[buf][k][i]
    float2 b[2][kGT][kGK + kGP];  // b[buf][j][k]
};

constexpr int kGQ = kGT * kGK / 256;  // chunk elements per thread, per operand

template <typename TA, typename TB>
__device__ __forceinline__ void gt_fetch(const TA *__restrict__ A, const TB *__restrict__ B,
                                         int N, int i0, int j0, int k0, int tid,
                                         float2 (&ra)[kGQ], float2 (&rb)[kGQ]) {
#pragma unroll
    for (int q = 0; q < kGQ; ++q) {
        const int e = tid + 256 * q;
        ra[q] = tof(A[(size_t)(k0 + e / kGT) * N + i0 + e % kGT]);
        rb[q] = tof(B[(size_t)(j0 + e / kGK) * N + k0 + e % kGK]);
    }
}

__device__ __forceinline__ void gt_store(GemmTile &L, int buf, int tid, const float2 (&ra)[kGQ],
                                         const float2 (&rb)[kGQ]) {
#pragma unroll
    for (int q = 0; q < kGQ; ++q) {
        const int e = tid + 256 * q;
        L.a[buf][e / kGT][e % kGT] = ra[q];
        L.b[buf][e / kGK][e % kGK] = rb[q];
    }
}

// The wave's 16 x 16 tile += A_chunk B_chunk from LDS buffer `buf`, with
// the inputs split into NP bf16 pieces (32-wide K steps).
template <int NP>
__device__ __forceinline__ void gt_mma_bf16(const GemmTile &L, int buf, int wi, int wj, int lane,
                                            v4f &re, v4f &im, v4f &re2, v4f &im2) {
    const int r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int st = 0; st < kGK / 32; ++st) {
        v8bf ar[NP], ai[NP], nai[NP], br[NP], bi[NP];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const float2 a = L.a[buf][32 * st + 8 * kq + u][16 * wi + r];
            const float2 b = L.b[buf][16 * wj + r][32 * st + 8 * kq + u];
            split_bf16<NP>(a.x, u, ar);
            split_bf16<NP>(a.y, u, ai);
            split_bf16<NP>(b.x, u, br);
            split_bf16<NP>(b.y, u, bi);
        }
        negate_bf16<NP>(ai, nai);
        re = mma_pieces<NP>(ar, br, re);
        re2 = mma_pieces<NP>(nai, bi, re2);
        im = mma_pieces<NP>(ar, bi, im);
        im2 = mma_pieces<NP>(ai, br, im2);
    }
}

// The wave's 16 x 16 tile += A_chunk B_chunk from LDS buffer `buf`; NP = 0:
// FP32 MFMA, 2 / 3: the bf16 x3 / x6 splits.
template <int NP>
__device__ __forceinline__ void gt_mma(const GemmTile &L, int buf, int wi, int wj, int lane,
                                       v4f &re, v4f &im) {
    const int r = lane & 15, kq = lane >> 4;
    v4f re2 = {0.f, 0.f, 0.f, 0.f}, im2 = {0.f, 0.f, 0.f, 0.f};
    if constexpr (NP != 0) {
        gt_mma_bf16<NP>(L, buf, wi, wj, lane, re, im, re2, im2);
    } else {
#pragma unroll
        for (int st = 0; st < kGK / 4; ++st) {
            const float2 a = L.a[buf][4 * st + kq][16 * wi + r];
            const float2 b = L.b[buf][16 * wj + r][4 * st + kq];
            re = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, re, 0, 0, 0);
            re2 = __builtin_amdgcn_mfma_f32_16x16x4f32(-a.y, b.y, re2, 0, 0, 0);
            im = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.y, im, 0, 0, 0);
            im2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.x, im2, 0, 0, 0);
        }
    }
    re += re2;
    im += im2;
}

// Operand pair of product p of a stage: Hm Xf, Xf Hp, then T_c Cd_c.
__device__ __forceinline__ void gt_operands(const DenseOp &D, int p, const float2 *&A,
                                            const float2 *&B) {
    const size_t NN = (size_t)D.N * D.N;
    if (p == 0) {
        A = D.Hm;
        B = D.Xf;
    } else if (p == 1) {
        A = D.Xf;
        B = D.Hp;
    } else {
        A = D.T + (size_t)(p - 2) * NN;
        B = D.Cd + (size_t)(p - 2) * NN;
    }
}

// The chunks of one output tile are split over the two halves of a
// 512-thread workgroup (chunk 2 s + h to half h, each half with its own LDS
// pipeline), which puts two waves on every SIMD; half 1's partial tile is
// then added to half 0's through LDS (fixed order).  Both halves run the
// same number of loop trips (a half past the end recomputes its last chunk
// and discards it), so every barrier is reached by all threads.
constexpr int kGHalves = 2;


struct GemmShared {
    GemmTile L[kGHalves];
};

// Product p's operands: p < 0 selects the T kernel's single product C_c X.
template <typename TB, int NP>
__device__ __forceinline__ void gt_run(const DenseOp &D, GemmShared &G, int np,
                                       const float2 *A0, const TB *B0, int i0, int j0,
                                       v4f &re, v4f &im) {
    const int h = threadIdx.x >> 8, tid = threadIdx.x & 255;
    const int lane = threadIdx.x & 63, w = tid >> 6, wi = w & 1, wj = w >> 1;
    const int nch = D.N / kGK, total = np * nch;
    const int trips = (total + kGHalves - 1) / kGHalves;
    GemmTile &L = G.L[h];
    auto operands = [&](int idx, const float2 *&A, const TB *&B, int &k0) {
        const int p = idx / nch;
        k0 = (idx % nch) * kGK;
        if (A0) {
            A = A0;
            B = B0;
        } else {
            const float2 *Bf;
            gt_operands(D, p, A, Bf);
            B = (const TB *)Bf;
        }
    };
    float2 ra[kGQ], rb[kGQ];
    const float2 *A;
    const TB *B;
    int k0;
    operands(min(h, total - 1), A, B, k0);
    gt_fetch(A, B, D.N, i0, j0, k0, tid, ra, rb);
    gt_store(L, 0, tid, ra, rb);
    __syncthreads();
    for (int it = 0; it < trips; ++it) {
        const int cur = kGHalves * it + h;
        const int nx = min(cur + kGHalves, total - 1);
        operands(nx, A, B, k0);
        gt_fetch(A, B, D.N, i0, j0, k0, tid, ra, rb);
        if (cur < total) gt_mma<NP>(L, it & 1, wi, wj, lane, re, im);
        gt_store(L, (it + 1) & 1, tid, ra, rb);
        __syncthreads();
    }
    // half 1 hands its partial tile to half 0 through its own LDS tile
    float *red = (float *)&G.L[1];
    if (h == 1) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            red[(w * 8 + v) * 64 + lane] = re[v];
            red[(w * 8 + 4 + v) * 64 + lane] = im[v];
        }
    }
    __syncthreads();
    if (h == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            re[v] += red[(w * 8 + v) * 64 + lane];
            im[v] += red[(w * 8 + 4 + v) * 64 + lane];
        }
    }
}

// T_c = C_c X for the workgroup's tile (grid N/32, N/32, max(nc, 1)); the
// blockIdx.z == 0 workgroups also write Xf = complex64(X) for their tile.
template <int NP>
__device__ __forceinline__ void dense_t_tiled(const DenseOp &D, const double2 *X) {
    __shared__ GemmShared G;
    const int i0 = blockIdx.x * kGT, j0 = blockIdx.y * kGT, c = blockIdx.z;
    const size_t NN = (size_t)D.N * D.N;
    if (c == 0) {
        for (int e = threadIdx.x; e < kGT * kGT; e += blockDim.x) {
            const size_t idx = (size_t)(j0 + e / kGT) * D.N + i0 + e % kGT;
            D.Xf[idx] = tof(X[idx]);
        }
    }
    if (c >= D.nc) return;  // uniform over the workgroup
    v4f re = {0.f, 0.f, 0.f, 0.f}, im = {0.f, 0.f, 0.f, 0.f};
    gt_run<double2, NP>(D, G, 1, D.C + (size_t)c * NN, X, i0, j0, re, im);
    if ((threadIdx.x >> 8) != 0) return;
    const int lane = threadIdx.x & 63, w = (threadIdx.x & 255) >> 6;
#pragma unroll
    for (int v = 0; v < 4; ++v)
        D.T[c * NN + tile_index(D.N, i0 + 16 * (w & 1), j0 + 16 * (w >> 1), lane, v)] =
            make_float2(re[v], im[v]);
}

// k tile = Hm Xf + Xf Hp + sum_c T_c Cd_c; complete in half 0.
template <int NP>
__device__ __forceinline__ void dense_k_tiled(const DenseOp &D, GemmShared &G, v4f &re,
                                              v4f &im) {
    gt_run<float2, NP>(D, G, 2 + D.nc, nullptr, (const float2 *)nullptr, blockIdx.x * kGT,
                   blockIdx.y * kGT, re, im);
}

template <int NP>
__global__ __launch_bounds__(256 * kGHalves) void rk_dense_t_tiled_kernel(EpiArgs P, DenseOp D) {
    if (P.ctl->phase != P.run_phase) return;
    const StageVecs v = stage_vecs(P, P.ctl);
    dense_t_tiled<NP>(D, v.xin);
}

template <int NP>
__global__ __launch_bounds__(256 * kGHalves) void rk_dense_stage_tiled_kernel(EpiArgs P,
                                                                             DenseOp D) {
    __shared__ GemmShared G;
    __shared__ double smem[256 * kGHalves / QA_WAVE];
    RKCtl *ctl = P.ctl;
    if (ctl->phase != P.run_phase) return;
    const StageVecs v = stage_vecs(P, ctl);
    const int h = threadIdx.x >> 8, lane = threadIdx.x & 63, w = (threadIdx.x & 255) >> 6;
    const int i0 = blockIdx.x * kGT + 16 * (w & 1), j0 = blockIdx.y * kGT + 16 * (w >> 1);
    double2 o[4], er[4];
    if (h == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) epi_pre(P, v, tile_index(D.N, i0, j0, lane, q), o[q], er[q]);
    }
    v4f re = {0.f, 0.f, 0.f, 0.f}, im = {0.f, 0.f, 0.f, 0.f};
    dense_k_tiled<NP>(D, G, re, im);
    double part = 0.0;
    if (h == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            part += epi_post(P, ctl, v, tile_index(D.N, i0, j0, lane, q),
                             make_double2((double)re[q], (double)im[q]), o[q], er[q]);
    }
    if (P.epi == EPI_FINAL) {
        const double sum = block_sum(part, smem);
        if (threadIdx.x == 0) P.partials[blockIdx.y * gridDim.x + blockIdx.x] = sum;
    }
}

// ---- Hermitian stage inputs: the upper tiles and their mirror -------------
// For a Hermitian input X the RHS k = -i H_nh X + i X H_nh^dag + sum_c c X c^dag
// is Hermitian (lindblad_matrix_form.pyx:175-189 forms A + A^dag for the same
// reason, assume_hermitian_state), so only the tiles on and above the
// diagonal are computed; each is also written conjugate-transposed below it.
// Upper tile p = bj (bj + 1) / 2 + bi has block row bi <= block column bj.
__device__ __forceinline__ void upper_tile(int p, int &bi, int &bj) {
    int j = (int)((sqrtf(8.0f * (float)p + 1.0f) - 1.0f) * 0.5f);
    while (j * (j + 1) / 2 > p) --j;
    while ((j + 1) * (j + 2) / 2 <= p) ++j;
    bj = j;
    bi = p - j * (j + 1) / 2;
}

// Item q of the upper tile of edge E at (i0, j0): the element it writes
// (column-major index) and the tile element it takes (src = row + E col,
// conjugated or with its imaginary part dropped).  Items [0, E^2) are the
// tile, [E^2, 2 E^2) the mirror tile at (j0, i0), whose element (r, c) is
// conj(tile(c, r)).  A diagonal tile (i0 == j0) has only the first E^2
// items: its strictly lower triangle mirrors the upper one and its diagonal
// is real, so the block written is exactly Hermitian (as A + A^dag is).
struct HermItem {
    int64_t idx;
    int src, conj, real;
};

template <int E>
__device__ __forceinline__ HermItem herm_item(int N, int i0, int j0, int q) {
    const int s = q % (E * E), r = s % E, c = s / E;
    HermItem it;
    if (q < E * E) {
        it.idx = (int64_t)(j0 + c) * N + i0 + r;
        const bool low = (i0 == j0) && r > c;
        it.src = low ? c + E * r : r + E * c;
        it.conj = low;
        it.real = (i0 == j0) && r == c;
    } else {
        it.idx = (int64_t)(i0 + c) * N + j0 + r;
        it.src = c + E * r;
        it.conj = 1;
        it.real = 0;
    }
    return it;
}

__device__ __forceinline__ double2 herm_value(float2 s, const HermItem &it) {
    return make_double2((double)s.x, it.real ? 0.0 : (it.conj ? -(double)s.y : (double)s.y));
}

// A 32-wide K step of a 16 x 16 tile: lane (r, kq) holds k = kk + 8 kq + u,
// u < 8 (a contiguous 64-byte piece of its B column), for the FP32 MFMAs
// (eight 16x16x4 instructions per part) and the bf16 pieces alike.  B may be
// the complex128 stage input (converted when the MFMAs use it, so the loads
// stay in flight).
template <typename TB>
__device__ __forceinline__ void step_load(const float2 *__restrict__ A, const TB *__restrict__ B,
                                          int N, int i0, int j0, int lane, int kk, float2 (&a)[8],
                                          TB (&b)[8]) {
    const int r = lane & 15, kq = lane >> 4;
    const float2 *pa = A + (size_t)(kk + 8 * kq) * N + i0 + r;
    const TB *pb = B + (size_t)(j0 + r) * N + kk + 8 * kq;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        a[u] = pa[(size_t)u * N];
        b[u] = pb[u];
    }
}

template <int NP, typename TB>
__device__ __forceinline__ void step_mma(const float2 (&a)[8], const TB (&bv)[8], v4f &re,
                                         v4f &im, v4f &re2, v4f &im2) {
    float2 b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) b[u] = tof(bv[u]);
    if constexpr (NP == 0) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            re = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].x, b[u].x, re, 0, 0, 0);
            re2 = __builtin_amdgcn_mfma_f32_16x16x4f32(-a[u].y, b[u].y, re2, 0, 0, 0);
            im = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].x, b[u].y, im, 0, 0, 0);
            im2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].y, b[u].x, im2, 0, 0, 0);
        }
    } else {
        v8bf ar[NP], ai[NP], nai[NP], br[NP], bi[NP];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            split_bf16<NP>(a[u].x, u, ar);
            split_bf16<NP>(a[u].y, u, ai);
            split_bf16<NP>(b[u].x, u, br);
            split_bf16<NP>(b[u].y, u, bi);
        }
        negate_bf16<NP>(ai, nai);
        re = mma_pieces<NP>(ar, br, re);
        re2 = mma_pieces<NP>(nai, bi, re2);
        im = mma_pieces<NP>(ar, bi, im);
        im2 = mma_pieces<NP>(ai, br, im2);
    }
}

template <typename TB>
__device__ __forceinline__ void step_mma_prec(int prec, const float2 (&a)[8], const TB (&b)[8],
                                              v4f &re, v4f &im, v4f &re2, v4f &im2) {
    if (prec == 1)
        step_mma<2>(a, b, re, im, re2, im2);
    else if (prec == 2)
        step_mma<3>(a, b, re, im, re2, im2);
    else
        step_mma<0>(a, b, re, im, re2, im2);
}

// Step s of the stage products of a wave (nb steps per product): Hm Xf,
// Xf Hp, then T_c Cd_c.
__device__ __forceinline__ void herm_step_load(const DenseOp &D, int nb, int s, int i0, int j0,
                                               int lane, int k0, float2 (&a)[8], float2 (&b)[8]) {
    const float2 *A, *B;
    gt_operands(D, s / nb, A, B);
    step_load(A, B, D.N, i0, j0, lane, k0 + 32 * (s % nb), a, b);
}

// Waves per 16 x 16 tile of the Hermitian kernels: K split so that every
// wave keeps whole 32-wide steps (N % 32 == 0).
__host__ __device__ constexpr int herm_split(int N) {
    return N % 256 == 0 ? 8 : N % 128 == 0 ? 4 : N % 64 == 0 ? 2 : 1;
}

// The wave partial tiles of a workgroup, summed in wave order (the same bits
// for a tile element wherever it is read).
__device__ __forceinline__ void herm_partial_store(float2 (*part)[256], int w, int lane,
                                                   const v4f &re, const v4f &im) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
        part[w][4 * (lane >> 4) + q + 16 * (lane & 15)] = make_float2(re[q], im[q]);
}

__device__ __forceinline__ float2 herm_partial_sum(const float2 (*part)[256], int ns, int e) {
    float2 t = part[0][e];
    for (int u = 1; u < ns; ++u) {
        t.x += part[u][e].x;
        t.y += part[u][e].y;
    }
    return t;
}

// T_c = C_c X and Xf = complex64(X) for Hermitian stages (grid N/16, N/16,
// max(nc, 1); herm_split(N) waves per tile).
__global__ __launch_bounds__(64 * kMaxSplit) void rk_dense_herm_t_kernel(EpiArgs P, DenseOp D) {
    __shared__ float2 part[kMaxSplit][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, ns = blockDim.x >> 6;
    const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16, c = blockIdx.z;
    const bool prod = c < D.nc;  // uniform: nc == 0 launches one layer for Xf
    const size_t NN = (size_t)D.N * D.N;
    int k0, k1;
    k_range(D.N, ns, w, k0, k1);
    const int nb = (k1 - k0) / 32;
    const float2 *Cc = D.C + (prod ? c : 0) * NN;
    // the tile's element e: row e % 16, column e / 16
    auto tile_idx = [&](int e) { return (size_t)(j0 + e / 16) * D.N + i0 + e % 16; };
    if (P.ctl->phase != P.run_phase) return;
    const StageVecs v = stage_vecs(P, P.ctl);
    const double2 *X = v.xin;
    if (c == 0) {
        for (int e = threadIdx.x; e < 256; e += blockDim.x) D.Xf[tile_idx(e)] = tof(X[tile_idx(e)]);
    }
    if (!prod) return;
    float2 a[8];
    double2 b[8];
    step_load(Cc, X, D.N, i0, j0, lane, k0, a, b);
    v4f re = {0.f, 0.f, 0.f, 0.f}, im = {0.f, 0.f, 0.f, 0.f};
    v4f re2 = {0.f, 0.f, 0.f, 0.f}, im2 = {0.f, 0.f, 0.f, 0.f};
    for (int s = 0; s < nb; ++s) {
        float2 an[8];
        double2 bn[8];
        step_load(Cc, X, D.N, i0, j0, lane, k0 + 32 * min(s + 1, nb - 1), an, bn);
        step_mma_prec(D.prec, a, b, re, im, re2, im2);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            a[u] = an[u];
            b[u] = bn[u];
        }
    }
    re += re2;
    im += im2;
    herm_partial_store(part, w, lane, re, im);
    __syncthreads();
    for (int e = threadIdx.x; e < 256; e += blockDim.x)
        D.T[c * NN + tile_idx(e)] = herm_partial_sum(part, ns, e);
}

// k on the upper 16 x 16 tiles (grid: one workgroup per upper tile, 8 / IPT
// waves); IPT items per thread, so every thread takes part in the epilogue.
// The first two product steps and the epilogue's loads are requested
// together, then each later step two steps ahead.
template <int IPT>
__global__ __launch_bounds__(512 / IPT) void rk_dense_herm_stage_kernel(EpiArgs P, DenseOp D) {
    __shared__ float2 part[kMaxSplit][256];
    __shared__ double smem[kMaxSplit];
    int bi, bj;
    upper_tile(blockIdx.x, bi, bj);
    const int i0 = 16 * bi, j0 = 16 * bj;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, ns = blockDim.x >> 6;
    int k0, k1;
    k_range(D.N, ns, w, k0, k1);
    const int nb = (k1 - k0) / 32, total = (2 + D.nc) * nb;
    RKCtl *ctl = P.ctl;
    if (ctl->phase != P.run_phase) return;
    float2 a0[8], b0[8], a1[8], b1[8];
    herm_step_load(D, nb, 0, i0, j0, lane, k0, a0, b0);
    herm_step_load(D, nb, min(1, total - 1), i0, j0, lane, k0, a1, b1);
    const StageVecs v = stage_vecs(P, ctl);
    const int nitems = bi == bj ? 256 : 512;
    HermItem it[IPT];
    double2 o[IPT], er[IPT];
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
        const int item = threadIdx.x + q * blockDim.x;
        it[q] = herm_item<16>(D.N, i0, j0, item);
        if (item < nitems) epi_pre(P, v, it[q].idx, o[q], er[q]);
    }
    v4f re = {0.f, 0.f, 0.f, 0.f}, im = {0.f, 0.f, 0.f, 0.f};
    v4f re2 = {0.f, 0.f, 0.f, 0.f}, im2 = {0.f, 0.f, 0.f, 0.f};
    for (int s = 0; s < total; ++s) {
        float2 a2[8], b2[8];
        herm_step_load(D, nb, min(s + 2, total - 1), i0, j0, lane, k0, a2, b2);
        step_mma_prec(D.prec, a0, b0, re, im, re2, im2);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            a0[u] = a1[u];
            b0[u] = b1[u];
            a1[u] = a2[u];
            b1[u] = b2[u];
        }
    }
    re += re2;
    im += im2;
    herm_partial_store(part, w, lane, re, im);
    __syncthreads();
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
        const int item = threadIdx.x + q * blockDim.x;
        if (item >= nitems) continue;
        const float2 t = herm_partial_sum(part, ns, it[q].src);
        sum += epi_post(P, ctl, v, it[q].idx, herm_value(t, it[q]), o[q], er[q]);
    }
    if (P.epi == EPI_FINAL) {
        sum = block_sum(sum, smem);
        if (threadIdx.x == 0) P.partials[blockIdx.x] = sum;
    }
}

// The same on the upper 32 x 32 workgroup tiles of the LDS-staged path
// (N >= kTiledN): 2 E^2 items over 512 threads.
template <int NP>
__global__ __launch_bounds__(256 * kGHalves) void rk_dense_herm_stage_tiled_kernel(EpiArgs P,
                                                                                  DenseOp D) {
    constexpr int IPT = 2 * kGT * kGT / (256 * kGHalves);
    __shared__ GemmShared G;
    __shared__ double smem[256 * kGHalves / QA_WAVE];
    RKCtl *ctl = P.ctl;
    if (ctl->phase != P.run_phase) return;
    const StageVecs v = stage_vecs(P, ctl);
    int bi, bj;
    upper_tile(blockIdx.x, bi, bj);
    const int i0 = kGT * bi, j0 = kGT * bj;
    const int nitems = bi == bj ? kGT * kGT : 2 * kGT * kGT;
    v4f re = {0.f, 0.f, 0.f, 0.f}, im = {0.f, 0.f, 0.f, 0.f};
    gt_run<float2, NP>(D, G, 2 + D.nc, nullptr, (const float2 *)nullptr, i0, j0, re, im);
    // the tile is complete in half 0; it goes through LDS to the items
    __syncthreads();
    float2 *tile = (float2 *)&G.L[0];
    const int h = threadIdx.x >> 8, lane = threadIdx.x & 63, w = (threadIdx.x & 255) >> 6;
    if (h == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            tile[16 * (w & 1) + 4 * (lane >> 4) + q + kGT * (16 * (w >> 1) + (lane & 15))] =
                make_float2(re[q], im[q]);
    }
    __syncthreads();
    // the epilogue's loads follow the products here: issued before them,
    // four items' K terms would not fit the registers beside the LDS pipeline
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
        const int item = threadIdx.x + q * blockDim.x;
        if (item >= nitems) continue;
        const HermItem it = herm_item<kGT>(D.N, i0, j0, item);
        double2 o, er;
        epi_pre(P, v, it.idx, o, er);
        sum += epi_post(P, ctl, v, it.idx, herm_value(tile[it.src], it), o, er);
    }
    if (P.epi == EPI_FINAL) {
        sum = block_sum(sum, smem);
        if (threadIdx.x == 0) P.partials[blockIdx.x] = sum;
    }
}

// Ungated tiled versions (qa_rk_apply and the first-step estimate).
template <int NP>
__global__ __launch_bounds__(256 * kGHalves) void dense_t_tiled_kernel(DenseOp D,
                                                                      const double2 *X) {
    dense_t_tiled<NP>(D, X);
}

template <int NP>
__global__ __launch_bounds__(256 * kGHalves) void dense_rhs_tiled_kernel(DenseOp D,
                                                                        double2 *out) {
    __shared__ GemmShared G;
    v4f re = {0.f, 0.f, 0.f, 0.f}, im = {0.f, 0.f, 0.f, 0.f};
    dense_k_tiled<NP>(D, G, re, im);
    if ((threadIdx.x >> 8) != 0) return;
    const int lane = threadIdx.x & 63, w = (threadIdx.x & 255) >> 6;
    const int i0 = blockIdx.x * kGT + 16 * (w & 1), j0 = blockIdx.y * kGT + 16 * (w >> 1);
#pragma unroll
    for (int q = 0; q < 4; ++q)
        out[tile_index(D.N, i0, j0, lane, q)] = make_double2((double)re[q], (double)im[q]);
}

// Ungated versions: out = L(X) (the data-layer product of the matrix form).
__global__ __launch_bounds__(64 * kMaxSplit) void dense_t_kernel(DenseOp D, const double2 *X) {
    dense_t_tile(D, X);
}

__global__ __launch_bounds__(64 * kMaxSplit) void dense_rhs_kernel(DenseOp D, const double2 *X, double2 *out) {
    v4f re = {0.f, 0.f, 0.f, 0.f}, im = {0.f, 0.f, 0.f, 0.f};
    dense_k_tile(D, X, re, im);
    if ((threadIdx.x >> 6) != 0) return;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        out[tile_index(D.N, blockIdx.x * 16, blockIdx.y * 16, threadIdx.x & 63, q)] =
            make_double2((double)re[q], (double)im[q]);
}

// Error norm and step control after the last stage: one workgroup sums the
// per-workgroup partials of the last stage kernel in a fixed order and runs
// the controller.  A kernel of its own (rather than a last-workgroup
// handshake inside the stage kernel) because a device-scope fence and atomic
// per workgroup costs more than this launch.
__global__ __launch_bounds__(256) void rk_control_kernel(RKCtl *ctl, const double *partials,
                                                         int nparts, int nrows) {
    __shared__ double smem[256 / QA_WAVE];
    // the partials and the control block are read in one round trip; the
    // sum is sum_partials' (strided per thread, then the block tree)
    double v = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) v += partials[i];
    const int phase = ctl->phase, adaptive = ctl->adaptive;
    if (phase != PH_STEP) return;
    const double tot = block_sum(v, smem);
    if (threadIdx.x == 0) after_attempt(ctl, adaptive ? sqrt(tot / (double)nrows) : 0.0);
}

struct LinCoef {
    double c[kMaxStages];
};

// x = y_prev + dt_int * sum_j coef_j k_j, for a first dense stage whose input
// is not y_front (explicit_rk.pyx:404-407).
__global__ __launch_bounds__(256) void rk_combine_kernel(RKCtl *ctl, int64_t npad, double2 *y0,
                                                         double2 *y1, const double2 *K,
                                                         double2 *xout, int nterms, LinCoef lc) {
    if (ctl->phase != PH_DENSE) return;
    const double2 *yprev = ctl->cur ? y1 : y0;
    const double dt = ctl->dt_int;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < npad;
         r += (int64_t)gridDim.x * blockDim.x) {
        double2 o = yprev[r];
        for (int j = 0; j < nterms; ++j) {
            const double f = dt * lc.c[j];
            if (f != 0.0) raxpy(o, f, K[(int64_t)j * npad + r]);
        }
        xout[r] = o;
    }
}

constexpr int kMaxDenseOrder = 8;

struct OutArgs {
    RKCtl *ctl;
    double2 *y0, *y1;
    const double2 *K;
    int64_t npad;
    int nk, dorder, has_bi;
    int nfun;
    const int *fptr, *fkind, *frow, *fcol;
    const double2 *fw;
    double2 *partials;  // [kMaxPending][nfun][nchunks]
    int nchunks;
    double2 *y_last;
    // dense-output table in HBM, [i][j] at i * kMaxDenseOrder + j (a device
    // array: a per-thread index into a kernel-argument array becomes one
    // dependent kernarg load per element)
    const double *bi;
    // every workgroup of the output grid arrives here once and the last one
    // completes the output; zero between launches
    unsigned int *arrivals;
};

// The fields of the control block the output kernels read, loaded at once
// at the start of a kernel (one round trip).
struct OutCtl {
    int phase, cur, out_idx, ntimes;
    double t_front, t_prev, dt_int;
    const double *times;
};

__device__ __forceinline__ OutCtl out_ctl(const RKCtl *c) {
    OutCtl o;
    o.phase = c->phase;
    o.cur = c->cur;
    o.out_idx = c->out_idx;
    o.ntimes = c->ntimes;
    o.t_front = c->t_front;
    o.t_prev = c->t_prev;
    o.dt_int = c->dt_int;
    o.times = c->times;
    return o;
}

// Number of pending outputs (output times inside [t_prev, t_front]) and
// their interpolation weights w[q][i] = dt_int * b_factor_i
// (explicit_rk.pyx:412-430), computed by the whole workgroup into LDS;
// front[q] = 1 when the output time equals t_front (the state is y_front).
__device__ int pending_outputs(const OutCtl &c, const OutArgs &A, double (*w)[kMaxStages],
                               int *front) {
    // Thread (q, i) reads output time q itself, so all the reads are one
    // round trip; the count is the number of leading times inside the step.
    __shared__ int okq[kMaxPending];
    const int t = threadIdx.x;
    if (t < kMaxPending * kMaxStages) {
        const int q = t / kMaxStages, i = t % kMaxStages;
        const int idx = c.out_idx + q;
        const double t_front = c.t_front, t_prev = c.t_prev, dt_int = c.dt_int;
        const bool in_range = idx < c.ntimes;
        // this thread's row of the dense-output table, every load issued at
        // once (they do not depend on the control block)
        double bv[kMaxDenseOrder];
#pragma unroll
        for (int j = 0; j < kMaxDenseOrder; ++j) bv[j] = A.bi[i * kMaxDenseOrder + j];
        const double tq = in_range ? c.times[idx] : 0.0;
        const bool ok = in_range && tq <= t_front;
        double v = 0.0;
        if (ok && tq != t_front && A.has_bi && i < A.nk) {
            const double tau = (tq - t_prev) / dt_int;
#pragma unroll
            for (int j = kMaxDenseOrder - 1; j >= 0; --j)
                if (j < A.dorder) {  // Horner from j = dorder - 1 down
                    v += bv[j];
                    v *= tau;
                }
            v *= dt_int;
        }
        w[q][i] = v;
        if (i == 0) {
            okq[q] = ok;
            front[q] = ok && tq == t_front;
        }
    }
    __syncthreads();
    int nq = 0;
    while (nq < kMaxPending && okq[nq]) ++nq;
    return nq;
}

// The stage vectors at idx, every load unconditional (the stage index is
// clamped into [0, nk), so no load sits behind a branch and all of them are
// in flight at once).
__device__ __forceinline__ void gather_stages(const OutArgs &A, int64_t idx,
                                              double2 (&kv)[kMaxStages]) {
#pragma unroll
    for (int i = 0; i < kMaxStages; ++i) {
        const int ii = i < A.nk ? i : A.nk - 1;
        kv[i] = A.K[(int64_t)ii * A.npad + idx];
    }
}

// y(t_q) from gathered values: y_front (yf), or y_prev (yp) + sum_i w_i k_i.
// Stages without a weight are dropped by a select, so what an unused slot
// holds never enters.
__device__ __forceinline__ double2 interp(const OutArgs &A, const double2 (&kv)[kMaxStages],
                                          double2 yp, double2 yf, const double *w, int front) {
    if (front) return yf;
    double2 o = yp;
#pragma unroll
    for (int i = 0; i < kMaxStages; ++i) {
        const double wi = i < A.nk ? w[i] : 0.0;
        const double2 kk = wi != 0.0 ? kv[i] : make_double2(0.0, 0.0);
        raxpy(o, wi, kk);
    }
    return o;
}

// y(t_q)[idx]: y_front, or y_prev + sum_i w_i k_i[idx].
__device__ __forceinline__ double2 state_at(const OutArgs &A, const double2 *yprev,
                                            const double2 *yfront, const double *w, int front,
                                            int64_t idx) {
    if (front) return yfront[idx];
    double2 kv[kMaxStages];
    gather_stages(A, idx, kv);
    return interp(A, kv, yprev[idx], make_double2(0.0, 0.0), w, 0);
}

// Writes the states of the pending outputs (for store_states) and the state
// at the last requested time (for get_state).
__global__ __launch_bounds__(256) void rk_store_kernel(OutArgs A) {
    RKCtl *c = A.ctl;
    const int ph = c->phase;
    if (ph != PH_DENSE && ph != PH_OUTPUT) return;
    // only the final requested output (get_state) or stored states need it
    if (!c->states_out) {
        const int last = c->ntimes - 1;
        if (last - c->out_idx >= kMaxPending || c->times[last] > c->t_front) return;
    }
    __shared__ double w[kMaxPending][kMaxStages];
    __shared__ int front[kMaxPending];
    const OutCtl oc = out_ctl(c);
    const int nq = pending_outputs(oc, A, w, front);
    const double2 *yprev = oc.cur ? A.y1 : A.y0;
    const double2 *yfront = oc.cur ? A.y0 : A.y1;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < A.npad;
         r += (int64_t)gridDim.x * blockDim.x) {
        for (int q = 0; q < nq; ++q) {
            const bool is_last = (c->out_idx + q == c->ntimes - 1);
            if (!c->states_out && !is_last) continue;
            const double2 v = state_at(A, yprev, yfront, w[q], front[q], r);
            if (c->states_out) c->states_out[(int64_t)(c->out_idx + q) * A.npad + r] = v;
            if (is_last) A.y_last[r] = v;
        }
    }
}

// Completion of an output (thread 0 of one workgroup): the time and status
// of the last output, the counters, and the next phase (route).
__device__ void finish_outputs(RKCtl *c, int ph, int nq, const int *front) {
    if (nq == 0) {
        // a non-finite step left no output time inside [t_prev, t_front]
        c->status = ST_OUTSIDE_RANGE;
        c->phase = PH_ERROR;
        return;
    }
    if (ph == PH_DENSE) {
        c->dense_ready = 1;
        c->n_dense++;
    }
    c->t_last = c->times[c->out_idx + nq - 1];
    c->status = front[nq - 1] ? ST_AT_FRONT : ST_INTERPOLATED;
    c->n_outputs += nq;
    c->out_idx += nq;
    route(c);
}

// Functionals of the pending output states: linear f = sum w y[col] (a
// vectorised density matrix: tr(O rho), tr(rho)) or quadratic
// f = sum w conj(y[row]) y[col] (a ket: <psi|O|psi>, ||psi||^2).  Grid
// (nchunks, nfun) of one wavefront each: a workgroup takes kOutChunk
// entries of one functional, one per lane, and sums them for every pending
// output time (the stage gathers of an entry serve all of them).  Many small
// workgroups, because the gathers (18 scattered lines per entry) are bound
// by what one CU can request, not by HBM.  With one chunk per functional
// the sums are the values; otherwise each workgroup writes its partial sum.
// The last workgroup to arrive adds the partial sums in a fixed order and
// completes the output (counters, next phase).
//
// Round trips: 1 the control block and the functional's range, 2 this
// lane's entry beside the output times, 3 the entry's gathers; all are
// issued before the weights are known.
__global__ __launch_bounds__(kOutThreads) void rk_output_kernel(OutArgs A) {
    RKCtl *c = A.ctl;
    const int f = blockIdx.y;
    const int lane = threadIdx.x;
    const OutCtl oc = out_ctl(c);
    const int fbeg = A.fptr[f], fend = A.fptr[f + 1];
    const int quad = A.fkind[f];
    const int ph = oc.phase;
    // an inactive launch changes nothing, so every workgroup of it reads the
    // same phase and leaves (the arrival counter stays at zero)
    if (ph != PH_DENSE && ph != PH_OUTPUT) return;
    __shared__ double w[kMaxPending][kMaxStages];
    __shared__ int front[kMaxPending];
    const double2 *yprev = oc.cur ? A.y1 : A.y0;
    const double2 *yfront = oc.cur ? A.y0 : A.y1;
    const int e = fbeg + blockIdx.x * kOutChunk + lane;
    const bool has = lane < kOutChunk && e < fend;
    const int ec = has ? e : 0;  // the entry arrays hold at least one element
    int col = A.fcol[ec], row = A.frow[ec];
    const double2 fw = A.fw[ec];
    if (!has) col = row = 0;
    double2 kc[kMaxStages], kr[kMaxStages];
    gather_stages(A, col, kc);
    const double2 pc = yprev[col], fc = yfront[col];
    double2 pr = pc, fr = fc;
    if (quad) {  // uniform over the workgroup
        gather_stages(A, row, kr);
        pr = yprev[row];
        fr = yfront[row];
    }
    const int nq = pending_outputs(oc, A, w, front);
    const bool one_pass = A.nchunks == 1;
#pragma unroll
    for (int q = 0; q < kMaxPending; ++q) {
        if (q >= nq) break;  // uniform
        double2 v = make_double2(0.0, 0.0);
        if (has) {
            v = cmul(fw, interp(A, kc, pc, fc, w[q], front[q]));
            if (quad) {
                const double2 l = interp(A, kr, pr, fr, w[q], front[q]);
                v = cmul(make_double2(l.x, -l.y), v);
            }
        }
        const double re = wave_sum(v.x), im = wave_sum(v.y);
        if (lane == 0 && f < A.nfun) {
            const double2 sq = make_double2(re, im);
            if (one_pass)
                c->fun_out[(int64_t)(oc.out_idx + q) * A.nfun + f] = sq;
            else
                A.partials[((int64_t)q * A.nfun + f) * A.nchunks + blockIdx.x] = sq;
        }
    }
    // Every workgroup arrives once, after its reads of the control block
    // have returned (their values are used above), so the last one to arrive
    // may change it.  With one chunk nothing this kernel writes is read by
    // another of its workgroups (the values go to the host, the control
    // block to the next kernel across the launch boundary) and the arrival
    // needs no fence (a device-scope fence is a write-back and invalidation
    // of the L2); with several, the partial sums are released before the
    // arrival and acquired by the last workgroup.  It resets the counter for
    // the next launch.
    __shared__ int last_s;
    if (lane == 0) {
        if (!one_pass) __threadfence();
        const unsigned total = gridDim.x * gridDim.y;
        const bool last = __hip_atomic_fetch_add(A.arrivals, 1u, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) == total - 1;
        if (last) {
            __hip_atomic_exchange(A.arrivals, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (!one_pass) __threadfence();
        }
        last_s = last;
    }
    __syncthreads();
    if (!last_s) return;
    if (!one_pass) {
        // each value: the partial sums strided over the lanes, then the wave
        // tree; the values go to LDS first, so no store sits between the
        // loads of one value and the next (four of them are in flight)
        __shared__ double2 vals[kOutThreads];
        const int npairs = nq * A.nfun;
        for (int p0 = 0; p0 < npairs; p0 += kOutThreads) {
#pragma unroll 4
            for (int p = p0; p < min(npairs, p0 + kOutThreads); ++p) {
                const double2 *part = A.partials + (int64_t)p * A.nchunks;  // p = q nfun + f
                double re = 0.0, im = 0.0;
                for (int k = lane; k < A.nchunks; k += kOutThreads) {
                    re += part[k].x;
                    im += part[k].y;
                }
                re = wave_sum(re);
                im = wave_sum(im);
                if (lane == 0) vals[p - p0] = make_double2(re, im);
            }
            __syncthreads();
            const int p = p0 + lane;
            if (p < npairs) {
                const int q = p / A.nfun, ff = p % A.nfun;
                c->fun_out[(int64_t)(oc.out_idx + q) * A.nfun + ff] = vals[lane];
            }
            __syncthreads();
        }
    }
    if (lane == 0) finish_outputs(c, ph, nq, front);
}

// The state at the last output time of a finished run into y_last, for a
// run whose iterations carry no rk_store_kernel: after the last output
// nothing moves the control block, the stage vectors or y_prev / y_front,
// so this is the state rk_store_kernel would have written (the same
// weights and the same sums).
__global__ __launch_bounds__(256) void rk_final_state_kernel(OutArgs A) {
    RKCtl *c = A.ctl;
    __shared__ double w[kMaxStages];
    __shared__ int front_s;
    const int tid = threadIdx.x;
    const RKCtl &c0 = *c;
    if (c0.phase != PH_DONE || c0.out_idx != c0.ntimes || c0.ntimes < 1) return;
    if (tid < kMaxStages) {
        double bv[kMaxDenseOrder];
#pragma unroll
        for (int j = 0; j < kMaxDenseOrder; ++j) bv[j] = A.bi[tid * kMaxDenseOrder + j];
        const double tq = c0.times[c0.ntimes - 1];
        double v = 0.0;
        if (tq != c0.t_front && A.has_bi && tid < A.nk) {
            const double tau = (tq - c0.t_prev) / c0.dt_int;
#pragma unroll
            for (int j = kMaxDenseOrder - 1; j >= 0; --j)
                if (j < A.dorder) {
                    v += bv[j];
                    v *= tau;
                }
            v *= c0.dt_int;
        }
        w[tid] = v;
        if (tid == 0) front_s = tq == c0.t_front;
    }
    __syncthreads();
    const double2 *yprev = c0.cur ? A.y1 : A.y0;
    const double2 *yfront = c0.cur ? A.y0 : A.y1;
    const int front = front_s;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + tid; r < A.npad;
         r += (int64_t)gridDim.x * blockDim.x)
        A.y_last[r] = state_at(A, yprev, yfront, w, front, r);
}

// Start of a run over new output times.
__global__ void rk_begin_kernel(RKCtl *c, int ntimes, double2 *states_out) {
    c->ntimes = ntimes;
    c->states_out = states_out;
    c->out_idx = 0;
    route(c);
}

// Initial state (explicit_rk.pyx:204-230): t_prev = t_front = t_last = t0,
// y_front = y_last = y0, step size from the host estimate.
__global__ void rk_init_kernel(RKCtl *c, double t0, double dt_safe) {
    c->t_prev = c->t_front = c->t_last = t0;
    c->dt_int = 0.0;
    c->dt_safe = dt_safe;
    c->dt_try = dt_safe;
    c->cur = 0;  // y_prev = y0 buffer, y_front = y1 buffer
    c->dense_ready = 0;
    c->status = ST_NORMAL;
    c->phase = PH_IDLE;
    c->n_attempts = c->n_accepted = c->n_dense = c->n_outputs = 0;
}

// Profiling helper: force the phase a stage kernel waits for.
__global__ void rk_set_phase_kernel(RKCtl *c, int phase) { c->phase = phase; }

}  // namespace

// ---------------------------------------------------------------------------
// host-side engine object
// ---------------------------------------------------------------------------
struct qa_rk_solver {
    int kind = 0;  // 0: sparse SELL-C generator, 1: dense Lindblad (matrix form), 2: banded
    DiaOp dia = {};  // kind 2: term list and the diagonals in HBM (dense_block)
    // buffers replaced while the engine lives; freed with it (hipFree waits
    // for the whole device, which would stall other engines' streams)
    std::vector<void *> retired;
    int n = 0, nslices = 0, vals_fp32 = 0, lanes = 1;
    int nparts = 0;      // per-workgroup error partials of the last stage
    DenseOp dop = {};    // kind 1: operators in HBM (complex64)
    void *dense_block = nullptr;
    int64_t npad = 0;
    const int *slice_ptr = nullptr;
    const int *col = nullptr;
    const void *vals = nullptr;
    int coded = 0;  // coded SELL-C operator: codes / value table / offset table
    int vbits = 8;
    int uwords = 0;  // coded, every slice the same width: words per lane
    const void *code = nullptr;
    const double2 *vtab = nullptr;
    const int *otab = nullptr;
    // time-dependent coded generator: coefficient programs (td.code and
    // td.consts in td_block) and a value table at a given time (vtab_t)
    int td_on = 0;
    CoefSet td = {};
    void *td_block = nullptr;
    double2 *vtab_t = nullptr;
    // tableau
    int order = 0, rk_step = 0, rk_extra = 0, dorder = 0, interpolate = 1, adaptive = 1;
    int first_dense_from_front = 0;
    std::vector<double> a, b, c, e, bi;
    double rtol = 1e-6, atol = 1e-8, first_step = 0, min_step = 0, max_step = 0;
    int nsteps = 1000;
    hipStream_t stream = nullptr;
    // device buffers
    void *block = nullptr;  // one allocation for vectors
    double2 *y0 = nullptr, *y1 = nullptr, *x0 = nullptr, *x1 = nullptr, *K = nullptr;
    double2 *y_last = nullptr, *tmp = nullptr;
    double *partials = nullptr;
    RKCtl *ctl = nullptr;
    unsigned int *arrivals = nullptr;  // the output kernel's arrival counter
    double *bi_dev = nullptr;
    double *times_dev = nullptr;
    int times_cap = 0;
    double2 *fun_out_dev = nullptr;
    int64_t fun_out_cap = 0;
    // functionals
    int nfun = 0, nchunks = 1;
    int *fptr_dev = nullptr, *fkind_dev = nullptr, *frow_dev = nullptr, *fcol_dev = nullptr;
    double2 *fw_dev = nullptr, *fpart_dev = nullptr;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_dirty = true;
    // whether the iteration carries rk_store_kernel: a run that stores
    // states does; otherwise rk_final_state_kernel writes y_last once at
    // the end of the run.  graph_store is the form the graph was captured in.
    bool run_store = true, graph_store = true;
    bool use_graph = true;  // QUTIP_AMD_GRAPH=0: launch the iteration directly
    // The step-only iteration: the step stages and rk_control_kernel, none of
    // the launches that only an output needs (QUTIP_AMD_STEP_ONLY=0: off).
    hipGraph_t graph_step = nullptr;
    hipGraphExec_t exec_step = nullptr;
    // kGraphBatch iterations in one graph, full and step-only: one graph
    // launch instead of kGraphBatch inside a chunk
    hipGraph_t graph_batch = nullptr, graph_step_batch = nullptr;
    hipGraphExec_t exec_batch = nullptr, exec_step_batch = nullptr;
    bool step_only = true;
    RKCtl *ctl_host = nullptr;  // pinned mirror
    long long graph_launches = 0;
    long long setup_rhs = 0;  // derivative evaluations of the first-step estimate
    int state_set = 0;
    // asynchronous run (qa_rk_begin / qa_rk_advance / qa_rk_finish)
    hipEvent_t event = nullptr;
    int run_state = 0;  // 0 idle, 1 running, 2 done, 3 failed
    int run_ntimes = 0, chunk = 4;
    long long run_iters = 0, last_attempts = -1, last_outputs = -1;
    bool in_flight = false;
    // step-only iterations: the run's output times (host copy), whether the
    // run takes single steps, the longest step-only head allowed next, the
    // previous chunk's step-only head and starting time
    std::vector<double> times_host;
    bool run_single = false;
    int ramp = 8, last_head = 0;
    double last_front = 0.0;
};

namespace {

void fill_epi(EpiArgs &P, qa_rk_solver *S, int i) {
    P.nrows = S->n;
    P.npad = S->npad;
    P.y0 = S->y0;
    P.y1 = S->y1;
    P.x0 = S->x0;
    P.x1 = S->x1;
    P.K = S->K;
    P.partials = S->partials;
    P.ctl = S->ctl;
    P.stage = i;
    P.cstage = i < (int)S->c.size() ? S->c[i] : 0.0;
    double coef[kMaxStages], ecoef[kMaxStages];
    for (int j = 0; j < kMaxStages; ++j) coef[j] = ecoef[j] = 0.0;
    const int ne = S->rk_extra;
    if (i < S->rk_step) {
        P.run_phase = PH_STEP;
        P.input = i == 0 ? IN_YPREV : IN_X;
        if (i < S->rk_step - 1) {
            P.epi = EPI_NEXT;
            P.nterms = i + 1;
            for (int j = 0; j <= i; ++j) coef[j] = S->a[(size_t)(i + 1) * ne + j];
        } else {
            P.epi = EPI_FINAL;
            P.nterms = S->rk_step;
            for (int j = 0; j < S->rk_step; ++j) {
                coef[j] = S->b[j];
                ecoef[j] = S->adaptive ? S->e[j] : 0.0;
            }
        }
    } else {
        P.run_phase = PH_DENSE;
        P.input = (i == S->rk_step && S->first_dense_from_front) ? IN_YFRONT : IN_X;
        if (i < S->rk_extra - 1) {
            P.epi = EPI_NEXT;
            P.nterms = i + 1;
            for (int j = 0; j <= i; ++j) coef[j] = S->a[(size_t)(i + 1) * ne + j];
        } else {
            P.epi = EPI_NONE;
            P.nterms = 0;
        }
    }
    P.coef_self = coef[i];
    P.ecoef_self = ecoef[i];
    P.nt = 0;
    for (int j = 0; j < kMaxStages; ++j) {
        P.tj[j] = 0;
        P.tc[j] = P.te[j] = 0.0;
    }
    if (P.epi != EPI_NONE) {
        for (int j = 0; j < i; ++j) {
            if (coef[j] != 0.0 || ecoef[j] != 0.0) {
                P.tj[P.nt] = j;
                P.tc[P.nt] = coef[j];
                P.te[P.nt] = ecoef[j];
                ++P.nt;
            }
        }
    }
}

template <typename V>
StageArgs<V> stage_args(qa_rk_solver *S, int i) {
    StageArgs<V> P;
    fill_epi(P, S, i);
    P.slice_ptr = S->slice_ptr;
    P.col = S->col;
    P.vals = (const V *)S->vals;
    P.nslices = S->nslices;
    P.code = (const uint2 *)S->code;
    P.vtab = S->vtab;
    P.otab = S->otab;
    P.vbits = S->vbits;
    P.uwords = S->uwords;
    P.td = S->td;
    return P;
}

int stage_blocks(const qa_rk_solver *S) {
    const int64_t threads = (int64_t)S->nslices * QA_WAVE;
    return (int)((threads + kStepBlocks - 1) / kStepBlocks);
}

template <typename V, int FORM, bool TD = false>
int launch_sparse_stage(qa_rk_solver *S, int i) {
    StageArgs<V> P = stage_args<V>(S, i);
    const dim3 grid((unsigned)stage_blocks(S)), block(kStepBlocks);
    switch (S->lanes) {
        case 1: hipLaunchKernelGGL((rk_stage_kernel<V, 1, FORM, TD>), grid, block, 0, S->stream, P); break;
        case 2: hipLaunchKernelGGL((rk_stage_kernel<V, 2, FORM, TD>), grid, block, 0, S->stream, P); break;
        case 4: hipLaunchKernelGGL((rk_stage_kernel<V, 4, FORM, TD>), grid, block, 0, S->stream, P); break;
        case 8: hipLaunchKernelGGL((rk_stage_kernel<V, 8, FORM, TD>), grid, block, 0, S->stream, P); break;
        default:
            qa::set_error("lanes per row must be 1, 2, 4 or 8");
            return QA_ERR_ARG;
    }
    QA_CHECK_LAUNCH();
    return QA_OK;
}

// Launches KERNEL<NP> for the dense engine's precision (prec 0: FP32 MFMA,
// NP 0; 1: bf16 x3, NP 2; 2: bf16 x6, NP 3).
#define QA_DENSE_LAUNCH(PREC, KERNEL, GRID, BLOCK, STREAM, ...)                             \
    do {                                                                                 \
        switch (PREC) {                                                                  \
            case 1: hipLaunchKernelGGL(KERNEL<2>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
            case 2: hipLaunchKernelGGL(KERNEL<3>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
            default: hipLaunchKernelGGL(KERNEL<0>, GRID, BLOCK, 0, STREAM, __VA_ARGS__); break; \
        }                                                                                \
    } while (0)

// Hermitian stage inputs: the T layer (which also writes Xf), then the
// upper-tile stage kernel.
int launch_dense_herm_stage(qa_rk_solver *S, const EpiArgs &P) {
    const int N = S->dop.N, nz = std::max(S->dop.nc, 1);
    if (S->dop.tiled) {
        const int nt = N / kGT;
        QA_DENSE_LAUNCH(S->dop.prec, rk_dense_t_tiled_kernel, dim3(nt, nt, nz),
                        dim3(256 * kGHalves), S->stream, P, S->dop);
        QA_CHECK_LAUNCH();
        QA_DENSE_LAUNCH(S->dop.prec, rk_dense_herm_stage_tiled_kernel, dim3(nt * (nt + 1) / 2),
                        dim3(256 * kGHalves), S->stream, P, S->dop);
        QA_CHECK_LAUNCH();
        return QA_OK;
    }
    const int nt = N / 16;
    const int ns = S->dop.hsplit;
    hipLaunchKernelGGL(rk_dense_herm_t_kernel, dim3(nt, nt, nz), dim3(64 * ns), 0, S->stream, P,
                       S->dop);
    QA_CHECK_LAUNCH();
    const dim3 grid(nt * (nt + 1) / 2), block(64 * ns);
    switch (ns) {
        case 8: hipLaunchKernelGGL(rk_dense_herm_stage_kernel<1>, grid, block, 0, S->stream, P, S->dop); break;
        case 4: hipLaunchKernelGGL(rk_dense_herm_stage_kernel<2>, grid, block, 0, S->stream, P, S->dop); break;
        case 2: hipLaunchKernelGGL(rk_dense_herm_stage_kernel<4>, grid, block, 0, S->stream, P, S->dop); break;
        default: hipLaunchKernelGGL(rk_dense_herm_stage_kernel<8>, grid, block, 0, S->stream, P, S->dop); break;
    }
    QA_CHECK_LAUNCH();
    return QA_OK;
}

// Workgroups of the last stage kernel, i.e. the error partials per attempt.
int stage_parts(const qa_rk_solver *S) {
    if (S->kind == 1) {
        const int nt = S->dop.N / (S->dop.tiled ? kGT : 16);
        return S->dop.herm ? nt * (nt + 1) / 2 : nt * nt;
    }
    return S->nparts;
}

int launch_dense_stage(qa_rk_solver *S, int i) {
    EpiArgs P;
    fill_epi(P, S, i);
    if (S->dop.herm) return launch_dense_herm_stage(S, P);
    if (S->dop.tiled) {
        const int nt = S->dop.N / kGT;
        QA_DENSE_LAUNCH(S->dop.prec, rk_dense_t_tiled_kernel, dim3(nt, nt, std::max(S->dop.nc, 1)),
                        dim3(256 * kGHalves), S->stream, P, S->dop);
        QA_CHECK_LAUNCH();
        QA_DENSE_LAUNCH(S->dop.prec, rk_dense_stage_tiled_kernel, dim3(nt, nt),
                        dim3(256 * kGHalves), S->stream, P, S->dop);
        QA_CHECK_LAUNCH();
        return QA_OK;
    }
    const int nt = S->dop.N / 16;
    const dim3 block(64 * S->dop.nsplit);
    if (S->dop.nc > 0) {
        hipLaunchKernelGGL(rk_dense_t_kernel, dim3(nt, nt, S->dop.nc), block, 0, S->stream, P,
                           S->dop);
        QA_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(rk_dense_stage_kernel, dim3(nt, nt), block, 0, S->stream, P, S->dop);
    QA_CHECK_LAUNCH();
    return QA_OK;
}

int launch_stage(qa_rk_solver *S, int i) {
    if (S->kind == 1) return launch_dense_stage(S, i);
    if (S->kind == 2) {
        EpiArgs P;
        fill_epi(P, S, i);
        hipLaunchKernelGGL(rk_dia_stage_kernel, dim3((unsigned)S->nparts), dim3(256), 0, S->stream,
                           P, S->dia);
        QA_CHECK_LAUNCH();
        return QA_OK;
    }
    if (S->coded && S->td_on)
        return S->uwords ? launch_sparse_stage<double2, FORM_CODED_UNIFORM, true>(S, i)
                         : launch_sparse_stage<double2, FORM_CODED, true>(S, i);
    if (S->coded)
        return S->uwords ? launch_sparse_stage<double2, FORM_CODED_UNIFORM>(S, i)
                         : launch_sparse_stage<double2, FORM_CODED>(S, i);
    return S->vals_fp32 ? launch_sparse_stage<float2, FORM_PLAIN>(S, i)
                        : launch_sparse_stage<double2, FORM_PLAIN>(S, i);
}

// out = A(t) x (ungated): the derivative for the first-step estimate and
// apply (t only matters for a time-dependent generator).
int engine_rhs(qa_rk_solver *S, const double2 *x, double2 *out, double t = 0.0) {
    if (S->kind == 2) {
        hipLaunchKernelGGL(dia_rhs_kernel, dim3((unsigned)S->nparts), dim3(256), 0, S->stream,
                           S->dia, x, out);
        QA_CHECK_LAUNCH();
        return QA_OK;
    }
    if (S->kind == 1 && S->dop.tiled) {
        const int nt = S->dop.N / kGT;
        QA_DENSE_LAUNCH(S->dop.prec, dense_t_tiled_kernel, dim3(nt, nt, std::max(S->dop.nc, 1)),
                        dim3(256 * kGHalves), S->stream, S->dop, x);
        QA_CHECK_LAUNCH();
        QA_DENSE_LAUNCH(S->dop.prec, dense_rhs_tiled_kernel, dim3(nt, nt),
                        dim3(256 * kGHalves), S->stream, S->dop, out);
        QA_CHECK_LAUNCH();
        return QA_OK;
    }
    if (S->kind == 1) {
        const int nt = S->dop.N / 16;
        const dim3 block(64 * S->dop.nsplit);
        if (S->dop.nc > 0) {
            hipLaunchKernelGGL(dense_t_kernel, dim3(nt, nt, S->dop.nc), block, 0, S->stream,
                               S->dop, x);
            QA_CHECK_LAUNCH();
        }
        hipLaunchKernelGGL(dense_rhs_kernel, dim3(nt, nt), block, 0, S->stream, S->dop, x, out);
        QA_CHECK_LAUNCH();
        return QA_OK;
    }
    if (S->coded && S->td_on) {
        const int nv = 1 << S->vbits;
        hipLaunchKernelGGL(coef_table_kernel, dim3(1), dim3(256), 0, S->stream, S->td, t, S->vtab,
                           S->vtab_t, nv);
        QA_CHECK_LAUNCH();
        return qa_sell_coded_matvec_z(S->n, S->n, S->lanes, S->nslices, S->slice_ptr, S->code,
                                      S->vtab_t, S->otab, S->vbits, x, out, 1.0, 0.0, 0,
                                      S->stream);
    }
    if (S->coded)
        return qa_sell_coded_matvec_z(S->n, S->n, S->lanes, S->nslices, S->slice_ptr, S->code,
                                      S->vtab, S->otab, S->vbits, x, out, 1.0, 0.0, 0,
                                      S->stream);
    return qa_sell_matvec_z(S->n, S->lanes, S->nslices, S->slice_ptr, S->col, S->vals,
                            S->vals_fp32, x, out, 1.0, 0.0, 0, S->stream);
}

OutArgs out_args(qa_rk_solver *S) {
    OutArgs A;
    A.ctl = S->ctl;
    A.y0 = S->y0;
    A.y1 = S->y1;
    A.K = S->K;
    A.npad = S->npad;
    A.nk = S->interpolate ? S->rk_extra : S->rk_step;
    A.dorder = S->dorder;
    A.has_bi = S->interpolate ? 1 : 0;
    A.bi = S->bi_dev;
    A.nfun = S->nfun;
    A.fptr = S->fptr_dev;
    A.fkind = S->fkind_dev;
    A.frow = S->frow_dev;
    A.fcol = S->fcol_dev;
    A.fw = S->fw_dev;
    A.partials = S->fpart_dev;
    A.nchunks = S->nchunks;
    A.y_last = S->y_last;
    A.arrivals = S->arrivals;
    return A;
}

// Enqueue one loop iteration (used both for graph capture and direct launch);
// with_store: include rk_store_kernel (stored states, or y_last per output).
// step_only: end after rk_control_kernel.  Every later launch of the full
// iteration runs only in the dense-output and output phases, so a step-only
// iteration whose attempt reaches an output leaves the loop in that phase:
// the step-only iterations after it return at once, and the next full
// iteration carries on from there exactly as the full iteration would have.
int enqueue_iteration(qa_rk_solver *S, bool with_store, bool step_only = false) {
    int rc;
    for (int i = 0; i < S->rk_step; ++i) {
        rc = launch_stage(S, i);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(rk_control_kernel, dim3(1), dim3(256), 0, S->stream, S->ctl, S->partials,
                       stage_parts(S), S->n);
    QA_CHECK_LAUNCH();
    if (step_only) return QA_OK;
    if (S->interpolate && S->rk_extra > S->rk_step) {
        if (!S->first_dense_from_front) {
            LinCoef lc;
            const int i = S->rk_step;
            for (int j = 0; j < kMaxStages; ++j)
                lc.c[j] = j < i ? S->a[(size_t)i * S->rk_extra + j] : 0.0;
            double2 *xout = (i & 1) ? S->x1 : S->x0;
            unsigned blocks = (unsigned)std::min<int64_t>((S->npad + 255) / 256, 4096);
            hipLaunchKernelGGL(rk_combine_kernel, dim3(blocks), dim3(256), 0, S->stream, S->ctl,
                               S->npad, S->y0, S->y1, S->K, xout, i, lc);
            QA_CHECK_LAUNCH();
        }
        for (int i = S->rk_step; i < S->rk_extra; ++i) {
            rc = launch_stage(S, i);
            if (rc) return rc;
        }
    }
    OutArgs A = out_args(S);
    if (with_store) {
        unsigned sblocks = (unsigned)std::min<int64_t>((S->npad + 255) / 256, 2048);
        hipLaunchKernelGGL(rk_store_kernel, dim3(sblocks), dim3(256), 0, S->stream, A);
        QA_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(rk_output_kernel, dim3(S->nchunks, std::max(S->nfun, 1)),
                       dim3(kOutThreads), 0, S->stream, A);
    QA_CHECK_LAUNCH();
    return QA_OK;
}

// Capture `count` iterations (full or step-only) into *graph / *exec.
int capture_iteration(qa_rk_solver *S, bool step_only, int count, hipGraph_t *graph,
                      hipGraphExec_t *exec) {
    if (*exec) {
        QA_HIP(hipGraphExecDestroy(*exec));
        *exec = nullptr;
    }
    if (*graph) {
        QA_HIP(hipGraphDestroy(*graph));
        *graph = nullptr;
    }
    QA_HIP(hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal));
    int rc = QA_OK;
    for (int k = 0; k < count && rc == QA_OK; ++k) rc = enqueue_iteration(S, S->run_store, step_only);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(S->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    QA_HIP(e2);
    *graph = g;
    QA_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return QA_OK;
}

int rebuild_graph(qa_rk_solver *S) {
    int rc = capture_iteration(S, false, 1, &S->graph, &S->exec);
    if (rc) return rc;
    if (S->step_only) {
        rc = capture_iteration(S, true, 1, &S->graph_step, &S->exec_step);
        if (rc) return rc;
    }
    // the batch graphs are captured again when a chunk first needs them
    for (hipGraphExec_t *e : {&S->exec_batch, &S->exec_step_batch})
        if (*e) {
            QA_HIP(hipGraphExecDestroy(*e));
            *e = nullptr;
        }
    for (hipGraph_t *g : {&S->graph_batch, &S->graph_step_batch})
        if (*g) {
            QA_HIP(hipGraphDestroy(*g));
            *g = nullptr;
        }
    S->graph_dirty = false;
    S->graph_store = S->run_store;
    return QA_OK;
}

// Launch `count` iterations of one kind: graph replays (kGraphBatch at a
// time while that many are left and the batch graph is there) or direct
// launches.
int launch_iterations(qa_rk_solver *S, bool step_only, int count) {
    const bool batch = (step_only ? S->exec_step_batch : S->exec_batch) != nullptr;
    while (count > 0) {
        if (!S->use_graph) {
            int rc = enqueue_iteration(S, S->run_store, step_only);
            if (rc) return rc;
            --count;
        } else if (batch && count >= kGraphBatch) {
            QA_HIP(hipGraphLaunch(step_only ? S->exec_step_batch : S->exec_batch, S->stream));
            count -= kGraphBatch;
        } else {
            QA_HIP(hipGraphLaunch(step_only ? S->exec_step : S->exec, S->stream));
            --count;
        }
    }
    return QA_OK;
}

// Retired buffers are freed in batches of kRetiredBatch (one device-wide
// wait per batch instead of one per buffer).
constexpr size_t kRetiredBatch = 64;
int release_retired(qa_rk_solver *S) {
    if (S->retired.size() < kRetiredBatch) return QA_OK;
    for (void *p : S->retired) QA_HIP(hipFree(p));
    S->retired.clear();
    return QA_OK;
}

void free_all(qa_rk_solver *S) {
    if (S->exec) (void)hipGraphExecDestroy(S->exec);
    if (S->graph) (void)hipGraphDestroy(S->graph);
    if (S->exec_step) (void)hipGraphExecDestroy(S->exec_step);
    if (S->graph_step) (void)hipGraphDestroy(S->graph_step);
    for (hipGraphExec_t e : {S->exec_batch, S->exec_step_batch})
        if (e) (void)hipGraphExecDestroy(e);
    for (hipGraph_t g : {S->graph_batch, S->graph_step_batch})
        if (g) (void)hipGraphDestroy(g);
    if (S->event) (void)hipEventDestroy(S->event);
    void *ptrs[] = {S->block, S->dense_block, S->ctl, S->bi_dev, S->times_dev, S->fun_out_dev, S->fptr_dev,
                    S->fkind_dev, S->frow_dev, S->fcol_dev, S->fw_dev, S->fpart_dev, S->td_block,
                    S->vtab_t, S->arrivals};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (void *p : S->retired) (void)hipFree(p);
    S->retired.clear();
    if (S->ctl_host) (void)hipHostFree(S->ctl_host);
    if (S->stream) (void)hipStreamDestroy(S->stream);
}

}  // namespace

namespace {

// Tableau, options, stream, vectors and control block of a new engine whose
// operator-specific fields (kind, n, npad, nparts, operator pointers) are
// already set.  Frees the engine and returns the error code on failure.
int rk_finish_create(qa_rk_solver *S, int order, int rk_step, int rk_extra, int dorder,
                     const double *a, const double *b, const double *c, const double *e,
                     const double *bi, double rtol, double atol, int nsteps, double first_step,
                     double min_step, double max_step, int interpolate) {
    S->order = order;
    S->rk_step = rk_step;
    S->rk_extra = rk_extra;
    S->dorder = bi ? dorder : 0;
    S->a.assign(a, a + (size_t)rk_extra * rk_extra);
    S->b.assign(b, b + rk_step);
    S->c.assign(c, c + rk_extra);
    S->adaptive = e != nullptr;
    if (e) S->e.assign(e, e + rk_step);
    S->interpolate = (bi != nullptr) && interpolate;
    {
        const char *g = getenv("QUTIP_AMD_GRAPH");
        S->use_graph = !(g && g[0] == '0');
        const char *so = getenv("QUTIP_AMD_STEP_ONLY");
        S->step_only = !(so && so[0] == '0');
    }
    if (bi) S->bi.assign(bi, bi + (size_t)rk_extra * dorder);
    S->rtol = rtol;
    S->atol = atol;
    S->nsteps = nsteps;
    S->first_step = first_step;
    S->min_step = min_step != 0.0 ? min_step : 1e-15;
    S->max_step = max_step;
    // First dense stage input equals y_front when its a-row is b (vern7).
    S->first_dense_from_front = 0;
    if (S->interpolate && rk_extra > rk_step) {
        bool same = true;
        for (int j = 0; j < rk_extra; ++j) {
            const double aj = S->a[(size_t)rk_step * rk_extra + j];
            const double bj = j < rk_step ? S->b[j] : 0.0;
            if (aj != bj) same = false;
        }
        S->first_dense_from_front = same;
    }
    int rc = QA_OK;
    auto fail = [&](int code) {
        free_all(S);
        delete S;
        return code;
    };
    hipError_t he;
    if (!S->stream) {
        he = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
        if (he != hipSuccess) return fail(qa::hip_fail(he, "hipStreamCreate", __FILE__, __LINE__));
    }
    const int nk = S->interpolate ? rk_extra : rk_step;
    const int64_t nvec = 6 + nk;  // y0 y1 x0 x1 y_last tmp + K
    const size_t vbytes = (size_t)S->npad * sizeof(double2);
    const size_t nparts = (size_t)S->nparts + 1;
    size_t bytes = vbytes * nvec + nparts * sizeof(double) + 256;
    he = hipMalloc(&S->block, bytes);
    if (he != hipSuccess) return fail(qa::hip_fail(he, "hipMalloc(rk vectors)", __FILE__, __LINE__));
    char *p = (char *)S->block;
    S->y0 = (double2 *)p; p += vbytes;
    S->y1 = (double2 *)p; p += vbytes;
    S->x0 = (double2 *)p; p += vbytes;
    S->x1 = (double2 *)p; p += vbytes;
    S->y_last = (double2 *)p; p += vbytes;
    S->tmp = (double2 *)p; p += vbytes;
    S->K = (double2 *)p; p += vbytes * nk;
    S->partials = (double *)p;
    rc = qa_fill_zero(S->block, bytes - 256 - nparts * sizeof(double), S->stream);
    if (rc) return fail(rc);
    he = hipMalloc((void **)&S->ctl, sizeof(RKCtl));
    if (he != hipSuccess) return fail(qa::hip_fail(he, "hipMalloc(ctl)", __FILE__, __LINE__));
    he = hipMalloc((void **)&S->arrivals, 256);
    if (he != hipSuccess) return fail(qa::hip_fail(he, "hipMalloc(arrivals)", __FILE__, __LINE__));
    rc = qa_fill_zero(S->arrivals, 256, S->stream);
    if (rc) return fail(rc);
    {
        // dense-output table, rows padded to kMaxDenseOrder (zeros without
        // interpolation: the output kernels then never read it)
        std::vector<double> bt((size_t)kMaxStages * kMaxDenseOrder, 0.0);
        if (S->interpolate)
            for (int i = 0; i < S->rk_extra && i < kMaxStages; ++i)
                for (int j = 0; j < S->dorder && j < kMaxDenseOrder; ++j)
                    bt[(size_t)i * kMaxDenseOrder + j] = S->bi[(size_t)i * S->dorder + j];
        he = hipMalloc((void **)&S->bi_dev, sizeof(double) * bt.size());
        if (he != hipSuccess) return fail(qa::hip_fail(he, "hipMalloc(bi)", __FILE__, __LINE__));
        he = hipMemcpy(S->bi_dev, bt.data(), sizeof(double) * bt.size(), hipMemcpyHostToDevice);
        if (he != hipSuccess) return fail(qa::hip_fail(he, "hipMemcpy(bi)", __FILE__, __LINE__));
    }
    he = hipHostMalloc((void **)&S->ctl_host, sizeof(RKCtl), hipHostMallocDefault);
    if (he != hipSuccess) return fail(qa::hip_fail(he, "hipHostMalloc", __FILE__, __LINE__));
    RKCtl init;
    memset(&init, 0, sizeof(init));
    init.rtol = rtol;
    init.atol = atol;
    init.min_step = S->min_step;
    init.max_step = max_step;
    init.order = order;
    init.max_numsteps = nsteps;
    init.interpolate = S->interpolate;
    init.adaptive = S->adaptive;
    init.status = ST_NOT_INITIATED;
    init.phase = PH_IDLE;
    he = hipMemcpyAsync(S->ctl, &init, sizeof(init), hipMemcpyHostToDevice, S->stream);
    if (he != hipSuccess) return fail(qa::hip_fail(he, "ctl upload", __FILE__, __LINE__));
    he = hipStreamSynchronize(S->stream);
    if (he != hipSuccess) return fail(qa::hip_fail(he, "sync", __FILE__, __LINE__));
    // no functionals until set
    int zero_ptr[2] = {0, 0};
    rc = qa_rk_set_functionals(S, 0, nullptr, zero_ptr, nullptr, nullptr, nullptr);
    if (rc) return fail(rc);
    return QA_OK;
}

int check_tableau(int rk_step, int rk_extra, int dorder, const double *a, const double *b,
                  const double *c, const double *bi, int nsteps) {
    QA_REQUIRE(rk_step >= 1 && rk_step <= kMaxStages && rk_extra >= rk_step &&
                   rk_extra <= kMaxStages,
               "1 <= rk_step <= rk_extra <= %d", kMaxStages);
    QA_REQUIRE(a && b && c, "tableau a, b, c required");
    QA_REQUIRE(!bi || (dorder >= 1 && dorder <= kMaxDenseOrder),
               "dense-output order must be 1..%d", kMaxDenseOrder);
    QA_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
    return QA_OK;
}

}  // namespace

extern "C" {

int qa_rk_create(qa_rk_solver **out, int n, int lanes, int nslices, const int *slice_ptr,
                 const int *col, const void *vals, int vals_fp32, int order, int rk_step,
                 int rk_extra, int dorder, const double *a, const double *b, const double *c,
                 const double *e, const double *bi, double rtol, double atol, int nsteps,
                 double first_step, double min_step, double max_step, int interpolate) {
    QA_REQUIRE(out != nullptr, "qa_rk_create: null output");
    *out = nullptr;
    QA_REQUIRE(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8,
               "qa_rk_create: lanes must be 1, 2, 4 or 8");
    QA_REQUIRE(n > 0 && nslices == (n + QA_WAVE / lanes - 1) / (QA_WAVE / lanes),
               "qa_rk_create: need n > 0 and nslices = ceil(n / (64 / lanes))");
    QA_REQUIRE(slice_ptr && col && vals, "qa_rk_create: null operator pointer");
    int rc = check_tableau(rk_step, rk_extra, dorder, a, b, c, bi, nsteps);
    if (rc) return rc;
    qa_rk_solver *S = new (std::nothrow) qa_rk_solver();
    QA_REQUIRE(S != nullptr, "qa_rk_create: out of host memory");
    S->kind = 0;
    S->n = n;
    S->lanes = lanes;
    S->nslices = nslices;
    S->npad = (int64_t)nslices * (QA_WAVE / lanes);
    S->nparts = stage_blocks(S);
    S->slice_ptr = slice_ptr;
    S->col = col;
    S->vals = vals;
    S->vals_fp32 = vals_fp32;
    rc = rk_finish_create(S, order, rk_step, rk_extra, dorder, a, b, c, e, bi, rtol, atol, nsteps,
                          first_step, min_step, max_step, interpolate);
    if (rc) return rc;
    *out = S;
    return QA_OK;
}

int qa_rk_create_coded(qa_rk_solver **out, int n, int lanes, int nslices, const int *slice_ptr,
                       const void *codes, const void *vtab, const int *otab, int vbits, int order,
                       int rk_step, int rk_extra, int dorder, const double *a, const double *b,
                       const double *c, const double *e, const double *bi, double rtol,
                       double atol, int nsteps, double first_step, double min_step,
                       double max_step, int interpolate) {
    QA_REQUIRE(out != nullptr, "qa_rk_create_coded: null output");
    *out = nullptr;
    QA_REQUIRE(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8,
               "qa_rk_create_coded: lanes must be 1, 2, 4 or 8");
    QA_REQUIRE(n > 0 && nslices == (n + QA_WAVE / lanes - 1) / (QA_WAVE / lanes),
               "qa_rk_create_coded: need n > 0 and nslices = ceil(n / (64 / lanes))");
    QA_REQUIRE(slice_ptr && codes && vtab && otab, "qa_rk_create_coded: null operator pointer");
    QA_REQUIRE(vbits >= 6 && vbits <= 10, "qa_rk_create_coded: vbits must be 6..10");
    int rc = check_tableau(rk_step, rk_extra, dorder, a, b, c, bi, nsteps);
    if (rc) return rc;
    qa_rk_solver *S = new (std::nothrow) qa_rk_solver();
    QA_REQUIRE(S != nullptr, "qa_rk_create_coded: out of host memory");
    S->kind = 0;
    S->n = n;
    S->lanes = lanes;
    S->nslices = nslices;
    S->npad = (int64_t)nslices * (QA_WAVE / lanes);
    S->nparts = stage_blocks(S);
    S->slice_ptr = slice_ptr;
    S->coded = 1;
    S->code = codes;
    S->vtab = (const double2 *)vtab;
    S->otab = otab;
    S->vbits = vbits;
    // Slices of one width (words per lane <= kUWordsMax) take the kernel form
    // that requests its code words without the slice pointers.
    {
        std::vector<int> sp(nslices + 1);
        hipError_t he = hipMemcpy(sp.data(), slice_ptr, sizeof(int) * (nslices + 1),
                                  hipMemcpyDeviceToHost);
        if (he != hipSuccess) {
            delete S;
            return qa::hip_fail(he, "hipMemcpy(slice_ptr)", __FILE__, __LINE__);
        }
        const int w = sp[1] - sp[0];
        bool uniform = w > 0 && w % QA_WAVE == 0 && w / QA_WAVE <= kUWordsMax;
        for (int k = 0; uniform && k <= nslices; ++k) uniform = sp[k] == k * w;
        S->uwords = uniform ? w / QA_WAVE : 0;
    }
    rc = rk_finish_create(S, order, rk_step, rk_extra, dorder, a, b, c, e, bi, rtol, atol, nsteps,
                          first_step, min_step, max_step, interpolate);
    if (rc) return rc;
    *out = S;
    return QA_OK;
}

int qa_rk_create_dia(qa_rk_solver **out, int N, int nc, int ndiag, const int *diag_op,
                     const int *diag_off, const double *diag_vals, int order, int rk_step,
                     int rk_extra, int dorder, const double *a, const double *b, const double *c,
                     const double *e, const double *bi, double rtol, double atol, int nsteps,
                     double first_step, double min_step, double max_step, int interpolate) {
    QA_REQUIRE(out != nullptr, "qa_rk_create_dia: null output");
    *out = nullptr;
    QA_REQUIRE(N >= 1 && (int64_t)N * N < (1LL << 31), "qa_rk_create_dia: 1 <= N, N^2 < 2^31");
    QA_REQUIRE(nc >= 0 && ndiag >= 1 && diag_op && diag_off && diag_vals,
               "qa_rk_create_dia: operators");
    for (int d = 0; d < ndiag; ++d)
        QA_REQUIRE(diag_op[d] >= 0 && diag_op[d] <= nc && diag_off[d] > -N && diag_off[d] < N,
                   "qa_rk_create_dia: diagonal %d: operator index in [0, nc], |offset| < N", d);
    int rc = check_tableau(rk_step, rk_extra, dorder, a, b, c, bi, nsteps);
    if (rc) return rc;
    // value rows: Hm = -i H_nh and Hp = i H_nh^dag (one row per H_nh
    // diagonal each), C_c and Cd_c = C_c^dag (one row per c diagonal each),
    // then a row of ones.  Operator 0 is H_nh, 1 + c is c_c; diagonal d holds
    // A[r, r + off] at diag_vals[d][r].
    const size_t Nz = (size_t)N;
    auto src = [&](int d, int r) {  // the entry of diagonal d in row r (0 outside)
        const int col = r + diag_off[d];
        if (r < 0 || r >= N || col < 0 || col >= N) return make_double2(0.0, 0.0);
        return make_double2(diag_vals[2 * ((size_t)d * Nz + r)], diag_vals[2 * ((size_t)d * Nz + r) + 1]);
    };
    std::vector<double2> rows;
    std::vector<DiaTerm> terms;
    std::vector<int> hm_row(ndiag, -1), hp_row(ndiag, -1), c_row(ndiag, -1), cd_row(ndiag, -1);
    for (int d = 0; d < ndiag; ++d) {
        const int o = diag_off[d];
        if (diag_op[d] == 0) {
            hm_row[d] = (int)(rows.size() / Nz);
            for (int r = 0; r < N; ++r) {  // -i h
                const double2 h = src(d, r);
                rows.push_back(make_double2(h.y, -h.x));
            }
            hp_row[d] = (int)(rows.size() / Nz);
            for (int r = 0; r < N; ++r) {  // Hp[r, r - o] = i conj(H[r - o, r])
                const double2 h = src(d, r - o);
                rows.push_back(make_double2(h.y, h.x));
            }
        } else {
            c_row[d] = (int)(rows.size() / Nz);
            for (int r = 0; r < N; ++r) rows.push_back(src(d, r));
            cd_row[d] = (int)(rows.size() / Nz);
            for (int r = 0; r < N; ++r) {  // Cd[r, r - o] = conj(C[r - o, r])
                const double2 cv = src(d, r - o);
                rows.push_back(make_double2(cv.x, -cv.y));
            }
        }
    }
    const int ones = (int)(rows.size() / Nz);
    for (int r = 0; r < N; ++r) rows.push_back(make_double2(1.0, 0.0));
    for (int d = 0; d < ndiag; ++d)
        if (diag_op[d] == 0) {
            terms.push_back(DiaTerm{diag_off[d], 0, hm_row[d], ones});   // Hm X
            terms.push_back(DiaTerm{0, -diag_off[d], ones, hp_row[d]});  // X Hp
        }
    for (int k = 1; k <= nc; ++k)
        for (int d1 = 0; d1 < ndiag; ++d1)
            if (diag_op[d1] == k)
                for (int d2 = 0; d2 < ndiag; ++d2)
                    if (diag_op[d2] == k)  // C_k[i, i + o1] X Cd_k[j - p2, j], p2 = -o2
                        terms.push_back(DiaTerm{diag_off[d1], -diag_off[d2], c_row[d1], cd_row[d2]});
    QA_REQUIRE((int)terms.size() <= kDiaMaxTerms,
               "qa_rk_create_dia: %d products of diagonals, at most %d", (int)terms.size(),
               kDiaMaxTerms);
    qa_rk_solver *S = new (std::nothrow) qa_rk_solver();
    QA_REQUIRE(S != nullptr, "qa_rk_create_dia: out of host memory");
    S->kind = 2;
    S->n = N * N;
    S->npad = (int64_t)N * N;
    S->nparts = (int)(((int64_t)N * N + 255) / 256);
    hipError_t he = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipMalloc(&S->dense_block, sizeof(double2) * rows.size());
    if (he == hipSuccess)
        he = hipMemcpy(S->dense_block, rows.data(), sizeof(double2) * rows.size(),
                       hipMemcpyHostToDevice);
    if (he != hipSuccess) {
        free_all(S);
        delete S;
        return qa::hip_fail(he, "banded operator upload", __FILE__, __LINE__);
    }
    S->dia.N = N;
    S->dia.nterms = (int)terms.size();
    S->dia.val = (const double2 *)S->dense_block;
    for (size_t k = 0; k < terms.size(); ++k) S->dia.t[k] = terms[k];
    rc = rk_finish_create(S, order, rk_step, rk_extra, dorder, a, b, c, e, bi, rtol, atol, nsteps,
                          first_step, min_step, max_step, interpolate);
    if (rc) return rc;
    *out = S;
    return QA_OK;
}

int qa_rk_create_dense(qa_rk_solver **out, int N, int nc, const double *H_nh, const double *c_ops,
                       int precision, int order, int rk_step, int rk_extra, int dorder,
                       const double *a, const double *b, const double *c, const double *e,
                       const double *bi, double rtol, double atol, int nsteps, double first_step,
                       double min_step, double max_step, int interpolate) {
    QA_REQUIRE(out != nullptr, "qa_rk_create_dense: null output");
    *out = nullptr;
    QA_REQUIRE(N >= 16 && N % 16 == 0, "qa_rk_create_dense: N must be a positive multiple of 16");
    QA_REQUIRE(precision >= 0 && precision <= 2,
               "qa_rk_create_dense: precision 0 (fp32), 1 (bf16x3) or 2 (bf16x6)");
    QA_REQUIRE(precision == 0 || N % 32 == 0, "qa_rk_create_dense: the bf16 forms need N % 32 == 0");
    QA_REQUIRE(nc >= 0 && nc <= 64 && H_nh && (nc == 0 || c_ops), "qa_rk_create_dense: operators");
    int rc = check_tableau(rk_step, rk_extra, dorder, a, b, c, bi, nsteps);
    if (rc) return rc;
    qa_rk_solver *S = new (std::nothrow) qa_rk_solver();
    QA_REQUIRE(S != nullptr, "qa_rk_create_dense: out of host memory");
    S->kind = 1;
    S->n = N * N;
    S->npad = (int64_t)N * N;
    S->nparts = (N / 16) * (N / 16);
    // Hm = -i H_nh, Hp = i H_nh^dag, C, Cd = C^dag, as complex64, column-major
    const size_t NN = (size_t)N * N;
    std::vector<float2> h((2 + 2 * (size_t)nc) * NN);
    float2 *Hm = h.data(), *Hp = Hm + NN, *Cs = Hp + NN, *Cds = Cs + nc * NN;
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            const double re = H_nh[2 * ((size_t)j * N + i)], im = H_nh[2 * ((size_t)j * N + i) + 1];
            Hm[(size_t)j * N + i] = make_float2((float)im, (float)-re);  // -i (re + i im)
            // (H_nh^dag)[j][i] = conj(H_nh[i][j]); times i
            Hp[(size_t)i * N + j] = make_float2((float)im, (float)re);   // i (re - i im)
        }
    for (int k = 0; k < nc; ++k)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                const double *src = c_ops + 2 * (k * NN + (size_t)j * N + i);
                Cs[k * NN + (size_t)j * N + i] = make_float2((float)src[0], (float)src[1]);
                Cds[k * NN + (size_t)i * N + j] = make_float2((float)src[0], (float)-src[1]);
            }
    hipError_t he = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking);
    if (he == hipSuccess)
        he = hipMalloc(&S->dense_block, sizeof(float2) * ((3 + 3 * (size_t)nc) * NN + 1));
    if (he == hipSuccess)
        he = hipMemcpy(S->dense_block, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice);
    if (he != hipSuccess) {
        free_all(S);
        delete S;
        return qa::hip_fail(he, "dense operator upload", __FILE__, __LINE__);
    }
    float2 *d = (float2 *)S->dense_block;
    S->dop.N = N;
    S->dop.nc = nc;
    S->dop.prec = precision;
    // split K over up to kMaxSplit waves while every wave keeps whole K steps
    // (16 for FP32, 32 for the bf16 splits)
    {
        const int kstep = precision >= 1 ? 32 : 16;
        int ns = kMaxSplit;
        while (ns > 1 && N % (ns * kstep) != 0) ns /= 2;
        S->dop.nsplit = ns;
    }
    S->dop.Hm = d;
    S->dop.Hp = d + NN;
    S->dop.C = d + 2 * NN;
    S->dop.Cd = d + (2 + (size_t)nc) * NN;
    S->dop.T = d + (2 + 2 * (size_t)nc) * NN;
    S->dop.Xf = d + (2 + 3 * (size_t)nc) * NN;
    S->dop.hsplit = N % 32 == 0 ? herm_split(N) : 1;
    S->dop.tiled = (N >= kTiledN && N % kGK == 0 && N % kGT == 0) ? 1 : 0;
    // S->nparts stays (N / 16)^2, the most any tile shape needs (the shape
    // can be changed later, qa_rk_set_dense_shape)
    rc = rk_finish_create(S, order, rk_step, rk_extra, dorder, a, b, c, e, bi, rtol, atol, nsteps,
                          first_step, min_step, max_step, interpolate);
    if (rc) return rc;
    *out = S;
    return QA_OK;
}

// out = L(x) with the engine's generator (device vectors of n entries).
int qa_rk_apply(qa_rk_solver *S, const void *x, void *out) { return qa_rk_apply_at(S, 0.0, x, out); }

int qa_rk_apply_at(qa_rk_solver *S, double t, const void *x, void *out) {
    QA_REQUIRE(S && x && out, "qa_rk_apply: null argument");
    int rc = engine_rhs(S, (const double2 *)x, (double2 *)out, t);
    if (rc) return rc;
    QA_HIP(hipStreamSynchronize(S->stream));
    return QA_OK;
}

namespace {
// Host-side checks of a coefficient set: program bounds, the constants an
// instruction names, the stack depth, and value-table ranges inside a table
// of nv entries.  Fills the CoefSet's host fields.
int check_coefficients(int ncoef, const int *start, const int *vstart, int ncode, const int *code,
                       int nconst, const double *consts, int nv, CoefSet &T) {
    QA_REQUIRE(ncoef >= 1 && ncoef <= kMaxCoef, "coefficients: 1 <= ncoef <= %d", kMaxCoef);
    QA_REQUIRE(start && code && start[0] == 0 && start[ncoef] == ncode,
               "coefficients: start[0] must be 0 and start[ncoef] the code length");
    QA_REQUIRE(nconst >= 0 && (nconst == 0 || consts), "coefficients: constants");
    for (int e = 0; e < ncoef; ++e) {
        QA_REQUIRE(start[e + 1] > start[e], "coefficients: program %d is empty", e);
        int depth = 0;
        for (int pc = start[e]; pc < start[e + 1]; ++pc) {
            const int op = code[pc] & 0xff, arg = code[pc] >> 8;
            QA_REQUIRE(op >= CO_CONST && op <= CO_SELECT, "coefficients: bad op %d", op);
            if (op == CO_CONST) QA_REQUIRE(arg >= 0 && arg < nconst, "coefficients: constant %d", arg);
            if (op == CO_FN) QA_REQUIRE(arg >= FN_SIN && arg <= FN_ERF, "coefficients: function %d", arg);
            if (op == CO_INTERP) {
                QA_REQUIRE(arg >= 0 && arg + 2 <= nconst, "coefficients: table at %d", arg);
                const int npts = (int)consts[2 * arg], nrow = (int)consts[2 * arg + 1];
                QA_REQUIRE(npts >= 1 && nrow >= 1 && nrow <= 8 &&
                               (long long)arg + 2 + npts + (long long)nrow * npts <= nconst,
                           "coefficients: table at %d out of range", arg);
            }
            if (op == CO_CONST || op == CO_T || op == CO_INTERP) ++depth;
            else if (op == CO_SELECT) depth -= 2;
            else if (op != CO_NEG && op != CO_FN) depth -= 1;
            QA_REQUIRE(depth >= 1 && depth < kCoefStack, "coefficients: program %d stack", e);
        }
        QA_REQUIRE(depth == 1, "coefficients: program %d leaves %d values", e, depth);
    }
    T.ncoef = ncoef;
    for (int e = 0; e <= kMaxCoef; ++e) {
        T.start[e] = e <= ncoef ? start[e] : ncode;
        T.vstart[e] = 0;
    }
    if (vstart) {
        for (int e = 0; e <= ncoef; ++e) {
            QA_REQUIRE(vstart[e] >= 0 && vstart[e] <= nv && (e == 0 || vstart[e] >= vstart[e - 1]),
                       "coefficients: value ranges must be increasing inside the table");
            T.vstart[e] = vstart[e];
        }
        for (int e = ncoef + 1; e <= kMaxCoef; ++e) T.vstart[e] = vstart[ncoef];
    }
    return QA_OK;
}

int upload_coefficients(CoefSet &T, int ncode, const int *code, int nconst, const double *consts,
                        void **block, hipStream_t s) {
    const size_t cb = ((sizeof(int) * (size_t)ncode + 255) / 256) * 256;
    const size_t kb = sizeof(double2) * (size_t)std::max(nconst, 1);
    QA_HIP(hipMalloc(block, cb + kb));
    QA_HIP(hipMemcpyAsync(*block, code, sizeof(int) * ncode, hipMemcpyHostToDevice, s));
    if (nconst)
        QA_HIP(hipMemcpyAsync((char *)*block + cb, consts, sizeof(double2) * nconst,
                              hipMemcpyHostToDevice, s));
    QA_HIP(hipStreamSynchronize(s));
    T.code = (const int *)*block;
    T.consts = (const double2 *)((char *)*block + cb);
    return QA_OK;
}
}  // namespace

int qa_rk_set_coefficients(qa_rk_solver *S, int ncoef, const int *start, const int *vstart,
                           int ncode, const int *code, int nconst, const double *consts) {
    QA_REQUIRE(S != nullptr, "qa_rk_set_coefficients: null solver");
    QA_REQUIRE(S->kind == 0 && S->coded, "qa_rk_set_coefficients: needs a coded SELL-C engine");
    QA_REQUIRE(vstart != nullptr, "qa_rk_set_coefficients: null value ranges");
    CoefSet T = {};
    QA_CHECK(check_coefficients(ncoef, start, vstart, ncode, code, nconst, consts, 1 << S->vbits, T));
    QA_HIP(hipStreamSynchronize(S->stream));
    if (S->td_block) S->retired.push_back(S->td_block);
    S->td_block = nullptr;
    QA_CHECK(upload_coefficients(T, ncode, code, nconst, consts, &S->td_block, S->stream));
    if (!S->vtab_t) {
        void *p = nullptr;
        QA_HIP(hipMalloc(&p, sizeof(double2) * kCodeTabMax));
        S->vtab_t = (double2 *)p;
    }
    S->td = T;
    S->td_on = 1;
    S->graph_dirty = true;
    return QA_OK;
}

int qa_coef_eval(int ncoef, const int *start, int ncode, const int *code, int nconst,
                 const double *consts, int ntimes, const double *times_dev, void *out_dev,
                 void *stream) {
    QA_REQUIRE(ntimes >= 0 && (ntimes == 0 || (times_dev && out_dev)), "qa_coef_eval: buffers");
    CoefSet T = {};
    QA_CHECK(check_coefficients(ncoef, start, nullptr, ncode, code, nconst, consts, 0, T));
    if (ntimes == 0) return QA_OK;
    hipStream_t s = (hipStream_t)stream;
    void *block = nullptr;
    QA_CHECK(upload_coefficients(T, ncode, code, nconst, consts, &block, s));
    const int64_t total = (int64_t)ncoef * ntimes;
    hipLaunchKernelGGL(coef_eval_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       T, ntimes, times_dev, (double2 *)out_dev);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(block);
    if (e != hipSuccess) return qa::hip_fail(e, "coef_eval_kernel", __FILE__, __LINE__);
    return QA_OK;
}

int qa_rk_destroy(qa_rk_solver *S) {
    if (!S) return QA_OK;
    (void)hipStreamSynchronize(S->stream);
    free_all(S);
    delete S;
    return QA_OK;
}

void *qa_rk_stream(qa_rk_solver *S) { return S ? (void *)S->stream : nullptr; }

int qa_functionals_check(int n, int nfun, const int *kinds, const int *fptr, const int *rows,
                         const int *cols) {
    QA_REQUIRE(nfun >= 0 && fptr && fptr[0] == 0, "functionals: bad fptr");
    QA_REQUIRE(nfun == 0 || kinds, "functionals: null kinds");
    for (int f = 0; f < nfun; ++f)
        QA_REQUIRE(fptr[f + 1] >= fptr[f], "functionals: fptr must be non-decreasing");
    const int nnz = fptr[nfun];
    QA_REQUIRE(nnz == 0 || (rows && cols), "functionals: null entries");
    for (int p = 0; p < nnz; ++p)
        QA_REQUIRE(cols[p] >= 0 && cols[p] < n, "functional entry %d: column %d out of range",
                   p, cols[p]);
    for (int f = 0; f < nfun; ++f) {
        if (kinds[f]) {
            for (int p = fptr[f]; p < fptr[f + 1]; ++p)
                QA_REQUIRE(rows[p] >= 0 && rows[p] < n,
                           "functional entry %d: row %d out of range", p, rows[p]);
        }
    }
    return QA_OK;
}

int qa_rk_set_hermitian(qa_rk_solver *S, int on) {
    QA_REQUIRE(S != nullptr, "qa_rk_set_hermitian: null solver");
    // the upper-tile kernels split K in 32-wide steps
    const int herm = (on && S->kind == 1 && S->dop.N % 32 == 0) ? 1 : 0;
    if (herm != S->dop.herm) {
        S->dop.herm = herm;
        S->graph_dirty = true;
    }
    return QA_OK;
}

int qa_rk_set_dense_shape(qa_rk_solver *S, int tiled, int nsplit, int hsplit) {
    QA_REQUIRE(S != nullptr, "qa_rk_set_dense_shape: null solver");
    QA_REQUIRE(S->kind == 1, "qa_rk_set_dense_shape: not a dense (MFMA) engine");
    const int N = S->dop.N;
    const bool pow2 = [](int v) { return v == 1 || v == 2 || v == 4 || v == 8; }(nsplit) &&
                      [](int v) { return v == 1 || v == 2 || v == 4 || v == 8; }(hsplit);
    QA_REQUIRE(pow2, "qa_rk_set_dense_shape: nsplit and hsplit must be 1, 2, 4 or 8");
    QA_REQUIRE(tiled == 0 || (tiled == 1 && N % kGK == 0 && N % kGT == 0),
               "qa_rk_set_dense_shape: tiled needs N %% %d == 0", kGK);
    const int kstep = S->dop.prec >= 1 ? 32 : 16;
    QA_REQUIRE(N % (nsplit * kstep) == 0,
               "qa_rk_set_dense_shape: N must be a multiple of %d x nsplit", kstep);
    QA_REQUIRE(N % (hsplit * 32) == 0 || N % 32 != 0,
               "qa_rk_set_dense_shape: N must be a multiple of 32 x hsplit");
    if (tiled != S->dop.tiled || nsplit != S->dop.nsplit || hsplit != S->dop.hsplit) {
        S->dop.tiled = tiled;
        S->dop.nsplit = nsplit;
        S->dop.hsplit = N % 32 == 0 ? hsplit : 1;
        S->graph_dirty = true;
    }
    return QA_OK;
}

int qa_rk_get_dense_shape(qa_rk_solver *S, int *tiled, int *nsplit, int *hsplit) {
    QA_REQUIRE(S != nullptr && tiled && nsplit && hsplit, "qa_rk_get_dense_shape: null argument");
    QA_REQUIRE(S->kind == 1, "qa_rk_get_dense_shape: not a dense (MFMA) engine");
    *tiled = S->dop.tiled;
    *nsplit = S->dop.nsplit;
    *hsplit = S->dop.hsplit;
    return QA_OK;
}

int qa_rk_get_hermitian(qa_rk_solver *S, int *on) {
    QA_REQUIRE(S != nullptr && on != nullptr, "qa_rk_get_hermitian: null argument");
    *on = S->kind == 1 ? S->dop.herm : 0;
    return QA_OK;
}

int qa_rk_set_functionals(qa_rk_solver *S, int nfun, const int *kinds, const int *fptr,
                          const int *rows, const int *cols, const double *w) {
    QA_REQUIRE(S != nullptr, "qa_rk_set_functionals: null solver");
    QA_CHECK(qa_functionals_check(S->n, nfun, kinds, fptr, rows, cols));
    const int nnz = fptr[nfun];
    int maxlen = 0;
    for (int f = 0; f < nfun; ++f) maxlen = std::max(maxlen, fptr[f + 1] - fptr[f]);
    QA_HIP(hipStreamSynchronize(S->stream));
    void *olds[] = {S->fptr_dev, S->fkind_dev, S->frow_dev, S->fcol_dev, S->fw_dev, S->fpart_dev};
    QA_HIP(hipStreamSynchronize(S->stream));  // no launch of this engine still reads them
    for (void *p : olds)
        if (p) S->retired.push_back(p);
    QA_CHECK(release_retired(S));
    S->fptr_dev = S->fkind_dev = S->frow_dev = S->fcol_dev = nullptr;
    S->fw_dev = S->fpart_dev = nullptr;
    S->nfun = nfun;
    S->nchunks = std::max(1, (maxlen + kOutChunk - 1) / kOutChunk);
    const int nf1 = std::max(nfun, 1);
    const int nz1 = std::max(nnz, 1);
    QA_HIP(hipMalloc((void **)&S->fptr_dev, sizeof(int) * (nf1 + 1)));
    QA_HIP(hipMalloc((void **)&S->fkind_dev, sizeof(int) * nf1));
    QA_HIP(hipMalloc((void **)&S->frow_dev, sizeof(int) * nz1));
    QA_HIP(hipMalloc((void **)&S->fcol_dev, sizeof(int) * nz1));
    QA_HIP(hipMalloc((void **)&S->fw_dev, sizeof(double2) * nz1));
    QA_HIP(hipMalloc((void **)&S->fpart_dev,
                     sizeof(double2) * (size_t)kMaxPending * nf1 * S->nchunks));
    std::vector<int> zeros_i(nz1, 0);
    std::vector<int> fp(nf1 + 1, 0), fk(nf1, 0);
    for (int f = 0; f <= nfun; ++f) fp[f] = fptr[f];
    for (int f = 0; f < nfun; ++f) fk[f] = kinds[f] ? 1 : 0;
    QA_HIP(hipMemcpy(S->fptr_dev, fp.data(), sizeof(int) * (nf1 + 1), hipMemcpyHostToDevice));
    QA_HIP(hipMemcpy(S->fkind_dev, fk.data(), sizeof(int) * nf1, hipMemcpyHostToDevice));
    if (nnz > 0) {
        QA_HIP(hipMemcpy(S->frow_dev, rows ? rows : zeros_i.data(), sizeof(int) * nnz,
                         hipMemcpyHostToDevice));
        QA_HIP(hipMemcpy(S->fcol_dev, cols, sizeof(int) * nnz, hipMemcpyHostToDevice));
        QA_HIP(hipMemcpy(S->fw_dev, w, sizeof(double2) * nnz, hipMemcpyHostToDevice));
    }
    S->graph_dirty = true;
    return QA_OK;
}

// set_initial_value + _estimate_first_step (explicit_rk.pyx:204-276).
int qa_rk_set_state(qa_rk_solver *S, double t0, const void *y0_dev) {
    QA_REQUIRE(S != nullptr && y0_dev != nullptr, "qa_rk_set_state: null argument");
    hipStream_t s = S->stream;
    const size_t nb = (size_t)S->n * sizeof(double2);
    int rc = qa_fill_zero(S->y1, (size_t)S->npad * sizeof(double2), s);
    if (rc) return rc;
    QA_HIP(hipMemcpyAsync(S->y1, y0_dev, nb, hipMemcpyDeviceToDevice, s));
    QA_HIP(hipMemcpyAsync(S->y_last, S->y1, (size_t)S->npad * sizeof(double2),
                          hipMemcpyDeviceToDevice, s));
    double dt = S->first_step;
    S->setup_rhs = 0;
    S->graph_launches = 0;
    if (dt == 0.0 && S->adaptive) {
        double norm2 = 0.0, kn2 = 0.0;
        rc = qa_norm2_z(S->n, S->y1, &norm2, s);
        if (rc) return rc;
        double norm = sqrt(norm2);
        const double tol = S->atol + norm * S->rtol;
        double2 *k0 = S->K, *k1 = S->K + S->npad;
        rc = engine_rhs(S, S->y1, k0, t0);
        if (rc) return rc;
        if (norm <= S->atol) norm = 1.0;
        rc = qa_norm2_z(S->n, k0, &kn2, s);
        if (rc) return rc;
        double tmp_norm = sqrt(kn2);
        double factorial = 1.0;
        for (int i = 1; i <= S->order; ++i) factorial *= i;
        double dt1, dt2;
        if (tmp_norm >= S->atol * 1e-6)
            dt1 = pow(tol * factorial * pow(norm, S->order), 1.0 / (S->order + 1)) / tmp_norm;
        else
            dt1 = pow(tol * factorial, 1.0 / (S->order + 1)) * norm * 0.5;
        const void *vecs[1] = {k0};
        const double cf[1] = {dt1 / 100};
        rc = qa_lincomb_z(S->npad, S->y1, 1, vecs, cf, S->tmp, s);
        if (rc) return rc;
        rc = engine_rhs(S, S->tmp, k1, t0 + dt1 / 100);
        if (rc) return rc;
        rc = qa_norm2_z(S->n, k1, &kn2, s);
        if (rc) return rc;
        S->setup_rhs += 2;
        tmp_norm = sqrt(kn2);
        if (tmp_norm >= S->atol * 1e-6)
            dt2 = pow(tol * factorial * pow(norm, S->order), 1.0 / (S->order + 1)) / tmp_norm;
        else
            dt2 = dt1;
        dt = std::min(dt1, dt2);
        if (S->max_step != 0.0) dt = std::min(S->max_step, dt);
        if (S->min_step != 0.0) dt = std::max(S->min_step, dt);
    } else if (!S->adaptive) {
        dt = 0.0;
    }
    hipLaunchKernelGGL(rk_init_kernel, dim3(1), dim3(1), 0, s, S->ctl, t0, dt);
    QA_CHECK_LAUNCH();
    QA_HIP(hipStreamSynchronize(s));
    S->state_set = 1;
    return QA_OK;
}

int qa_rk_begin(qa_rk_solver *S, int ntimes, const double *times, int single_step,
                void *states_out) {
    QA_REQUIRE(S != nullptr, "qa_rk_begin: null solver");
    QA_REQUIRE(S->state_set, "qa_rk_begin: set_state was not called");
    QA_REQUIRE(ntimes >= 1 && times != nullptr, "qa_rk_begin: need at least one time");
    for (int i = 1; i < ntimes; ++i)
        QA_REQUIRE(times[i] > times[i - 1], "qa_rk_begin: times must be increasing");
    hipStream_t s = S->stream;
    // the previous run may still be draining its last chunk
    QA_HIP(hipStreamSynchronize(s));
    if (ntimes > S->times_cap) {
        if (S->times_dev) S->retired.push_back(S->times_dev);
        QA_CHECK(release_retired(S));
        S->times_dev = nullptr;
        S->times_cap = std::max(ntimes, 2 * S->times_cap);
        QA_HIP(hipMalloc((void **)&S->times_dev, sizeof(double) * S->times_cap));
    }
    const int64_t need = (int64_t)ntimes * std::max(S->nfun, 1);
    if (need > S->fun_out_cap) {
        if (S->fun_out_dev) S->retired.push_back(S->fun_out_dev);
        QA_CHECK(release_retired(S));
        S->fun_out_dev = nullptr;
        S->fun_out_cap = std::max(need, 2 * S->fun_out_cap);
        QA_HIP(hipMalloc((void **)&S->fun_out_dev, sizeof(double2) * S->fun_out_cap));
    }
    QA_HIP(hipMemcpyAsync(S->times_dev, times, sizeof(double) * ntimes, hipMemcpyHostToDevice, s));
    // pointers and flags that can change between runs go through the ctl block
    struct {
        double *times;
        double2 *fun_out;
    } ptrs = {S->times_dev, S->fun_out_dev};
    QA_HIP(hipMemcpyAsync((char *)S->ctl + offsetof(RKCtl, times), &ptrs, sizeof(ptrs),
                          hipMemcpyHostToDevice, s));
    const int nf = S->nfun;
    QA_HIP(hipMemcpyAsync((char *)S->ctl + offsetof(RKCtl, nfun), &nf, sizeof(int),
                          hipMemcpyHostToDevice, s));
    QA_HIP(hipMemcpyAsync((char *)S->ctl + offsetof(RKCtl, single_step), &single_step,
                          sizeof(int), hipMemcpyHostToDevice, s));
    // the sources above are pageable host memory: complete the copies here
    QA_HIP(hipStreamSynchronize(s));
    S->run_store = states_out != nullptr;
    if (S->use_graph && (S->graph_dirty || S->graph_store != S->run_store)) {
        int rc = rebuild_graph(S);
        if (rc) return rc;
    }
    if (!S->event) QA_HIP(hipEventCreateWithFlags(&S->event, hipEventDisableTiming));
    hipLaunchKernelGGL(rk_begin_kernel, dim3(1), dim3(1), 0, s, S->ctl, ntimes,
                       (double2 *)states_out);
    QA_CHECK_LAUNCH();
    S->run_ntimes = ntimes;
    S->run_iters = 0;
    S->chunk = 4;
    S->last_attempts = -1;
    S->last_outputs = -1;
    S->in_flight = false;
    S->times_host.assign(times, times + ntimes);
    S->run_single = single_step != 0;
    S->ramp = 8;
    S->last_head = 0;
    S->run_state = 1;
    return QA_OK;
}

namespace {

// The next chunk after the control block mirror `h`: its first *head
// iterations step-only, the rest full; returns the chunk length (`chunk`,
// the full iterations the outputs left need at the recent rate, when
// step-only iterations do not pay).  They pay on output intervals of 32 or
// more attempts, at the larger of the trial step and the last chunk's mean
// advance per attempt: with p the attempts still to the next output, the
// chunk is min(0.75 p, ramp) step-only iterations and two full ones, or,
// once 0.75 p < 4, ceil(p) + 2 full ones; every chunk of such an interval
// ends there and the next one plans from where it ends up.  `ramp` doubles
// (to 64) after a chunk that made at least as many attempts as it had
// step-only iterations and halves (to 4) after one that did not (the output
// came earlier than planned).  A step-only iteration after its output was
// reached costs only launches that return at once: the results do not
// depend on the plan.
int plan_chunk(qa_rk_solver *S, const RKCtl &h, long long attempts_before, double front_before,
               int chunk, int *head) {
    *head = 0;
    if (S->last_head > 0) {
        const bool full = h.n_attempts - attempts_before >= S->last_head;
        S->ramp = full ? std::min(kMaxChunk, 2 * S->ramp) : std::max(4, S->ramp / 2);
    }
    if (!S->step_only || S->run_single || h.phase != PH_STEP || h.out_idx >= S->run_ntimes)
        return chunk;
    double dt = std::max(h.dt_try, h.dt_safe);
    if (attempts_before >= 0 && h.n_attempts > attempts_before)
        dt = std::max(dt, (h.t_front - front_before) / (double)(h.n_attempts - attempts_before));
    const int j = h.out_idx;
    const double start = j > 0 ? S->times_host[j - 1] : h.t_prev;
    if (!(dt > 0.0) || !((S->times_host[j] - start) / dt >= 32.0)) return chunk;
    const double p = (S->times_host[j] - h.t_front) / dt;
    if (!(p >= 0.0)) return chunk;
    const int n = (int)std::min<double>(0.75 * p, (double)S->ramp);
    if (n >= 4) {
        *head = n;
        return n + 2;
    }
    return std::min(chunk, (int)std::ceil(p) + 2);
}

}  // namespace

// Non-blocking progress: when the previous chunk has finished (event
// query), read the device phase from the pinned mirror and either report
// completion or enqueue the next chunk of loop iterations (graph replays)
// followed by an asynchronous copy of the control block.  *done: 0 running,
// 1 finished, -1 failed (the integrator status is in qa_rk_finish).
// Every iteration either attempts a step or emits outputs, so a chunk that
// changes neither counter means the device loop is stuck: that is an error.
int qa_rk_advance(qa_rk_solver *S, int *done) {
    QA_REQUIRE(S != nullptr && done != nullptr, "qa_rk_advance: null argument");
    QA_REQUIRE(S->run_state != 0, "qa_rk_advance: no run in progress (qa_rk_begin)");
    *done = 0;
    if (S->run_state == 2) {
        *done = 1;
        return QA_OK;
    }
    if (S->run_state == 3) {
        *done = -1;
        return QA_OK;
    }
    hipStream_t s = S->stream;
    int head = 0;        // step-only iterations at the start of the next chunk
    bool idle = false;   // the stream has no work left (the last chunk completed)
    if (S->in_flight) {
        const hipError_t q = hipEventQuery(S->event);
        if (q == hipErrorNotReady) return QA_OK;
        QA_HIP(q);
        S->in_flight = false;
        const RKCtl &h = *S->ctl_host;
        if (h.phase == PH_DONE || h.phase == PH_IDLE) {
            if (!S->run_store) {  // y_last, which rk_store_kernel did not write
                OutArgs A = out_args(S);
                const unsigned blocks = (unsigned)std::min<int64_t>((S->npad + 255) / 256, 2048);
                hipLaunchKernelGGL(rk_final_state_kernel, dim3(blocks), dim3(256), 0, s, A);
                QA_CHECK_LAUNCH();
            }
            S->run_state = 2;
            *done = 1;
            return QA_OK;
        }
        if (h.phase == PH_ERROR) {
            S->run_state = 3;
            *done = -1;
            return QA_OK;
        }
        if (h.n_attempts == S->last_attempts && h.n_outputs == S->last_outputs) {
            S->run_state = 3;
            qa::set_error("device loop made no progress (phase %d, output %d)", h.phase, h.out_idx);
            return QA_ERR_STATE;
        }
        const long long attempts_before = S->last_attempts;
        S->last_attempts = h.n_attempts;
        S->last_outputs = h.n_outputs;
        const int ndone = std::max(h.out_idx, 1);
        const double per_out = (double)S->run_iters / ndone;
        const long long est = (long long)(per_out * (S->run_ntimes - h.out_idx)) + 1;
        S->chunk = (int)std::max(1LL, std::min(est, (long long)kMaxChunk));
        S->chunk = plan_chunk(S, h, attempts_before, S->last_front, S->chunk, &head);
        S->last_front = h.t_front;
        idle = true;
    }
    S->last_head = head;
    // An engine that has run kBatchAfter iterations captures the batch graph
    // of a kind when a chunk first holds kGraphBatch of it, at a point where
    // its stream is idle (the last chunk has completed); engines that live
    // short lives (a server's requests) seldom pay for the capture.
    if (idle && S->use_graph && S->graph_launches >= kBatchAfter) {
        const int nso = std::min(head, S->chunk);
        if (nso >= kGraphBatch && !S->exec_step_batch) {
            int rc = capture_iteration(S, true, kGraphBatch, &S->graph_step_batch,
                                       &S->exec_step_batch);
            if (rc) return rc;
        }
        if (S->chunk - nso >= kGraphBatch && !S->exec_batch) {
            int rc = capture_iteration(S, false, kGraphBatch, &S->graph_batch, &S->exec_batch);
            if (rc) return rc;
        }
    }
    int rc = launch_iterations(S, true, std::min(head, S->chunk));
    if (rc) return rc;
    rc = launch_iterations(S, false, S->chunk - std::min(head, S->chunk));
    if (rc) return rc;
    S->run_iters += S->chunk;
    S->graph_launches += S->chunk;
    QA_HIP(hipMemcpyAsync(S->ctl_host, S->ctl, sizeof(RKCtl), hipMemcpyDeviceToHost, s));
    QA_HIP(hipEventRecord(S->event, s));
    S->in_flight = true;
    return QA_OK;
}

int qa_rk_finish(qa_rk_solver *S, double *fun_out, int *status, double *t_reached) {
    QA_REQUIRE(S != nullptr, "qa_rk_finish: null solver");
    QA_REQUIRE(S->run_state == 2 || S->run_state == 3, "qa_rk_finish: the run has not ended");
    QA_HIP(hipStreamSynchronize(S->stream));
    const RKCtl &h = *S->ctl_host;
    if (status) *status = h.status;
    if (t_reached) *t_reached = h.t_last;
    const bool failed = S->run_state == 3;
    S->run_state = 0;
    if (fun_out && S->nfun > 0) {
        const int ndone = h.phase == PH_DONE ? S->run_ntimes : h.out_idx;
        QA_HIP(hipMemcpy(fun_out, S->fun_out_dev, sizeof(double2) * (size_t)ndone * S->nfun,
                         hipMemcpyDeviceToHost));
    }
    if (failed || h.phase == PH_ERROR) {
        qa::set_error("integration failed with status %d", h.status);
        return QA_ERR_STATE;
    }
    return QA_OK;
}

int qa_rk_run(qa_rk_solver *S, int ntimes, const double *times, int single_step,
              double *fun_out, void *states_out, int *status, double *t_reached) {
    int rc = qa_rk_begin(S, ntimes, times, single_step, states_out);
    if (rc) return rc;
    int done = 0;
    while (!done) {
        rc = qa_rk_advance(S, &done);
        if (rc) {
            S->run_state = 0;
            return rc;
        }
        if (!done && S->in_flight) QA_HIP(hipEventSynchronize(S->event));
    }
    return qa_rk_finish(S, fun_out, status, t_reached);
}

int qa_rk_get_state(qa_rk_solver *S, double *t, void *y_out) {
    QA_REQUIRE(S != nullptr, "qa_rk_get_state: null solver");
    QA_HIP(hipMemcpyAsync(S->ctl_host, S->ctl, sizeof(RKCtl), hipMemcpyDeviceToHost, S->stream));
    if (y_out)
        QA_HIP(hipMemcpyAsync(y_out, S->y_last, (size_t)S->n * sizeof(double2),
                              hipMemcpyDeviceToDevice, S->stream));
    QA_HIP(hipStreamSynchronize(S->stream));
    if (t) *t = S->ctl_host->t_last;
    return QA_OK;
}

int qa_rk_stats(qa_rk_solver *S, int64_t *stats) {
    QA_REQUIRE(S != nullptr && stats != nullptr, "qa_rk_stats: null argument");
    QA_HIP(hipMemcpyAsync(S->ctl_host, S->ctl, sizeof(RKCtl), hipMemcpyDeviceToHost, S->stream));
    QA_HIP(hipStreamSynchronize(S->stream));
    const RKCtl &h = *S->ctl_host;
    stats[0] = h.n_attempts;
    stats[1] = h.n_accepted;
    stats[2] = h.n_dense;
    stats[3] = (int64_t)h.n_attempts * S->rk_step +
               (int64_t)h.n_dense * (S->interpolate ? S->rk_extra - S->rk_step : 0) +
               S->setup_rhs;
    stats[4] = S->graph_launches;
    stats[5] = h.status;
    stats[6] = h.out_idx;
    stats[7] = h.phase;
    return QA_OK;
}

// Average duration (ms) of each stage kernel, measured with HIP events on
// the engine's stream, `reps` launches per stage.  The stage kernels run on
// the engine's current vectors, so the integration state is clobbered: call
// qa_rk_set_state afterwards.  The control block is saved and restored; while
// profiling, it points at a private far-future output time so the step
// controller of the last stage only touches the control block.
// Profiling runs stage kernels outside a solve: the control block is saved,
// given an output time that is never reached, and restored afterwards.
static int profile_enter(qa_rk_solver *S, RKCtl **backup, double **far_time) {
    hipStream_t s = S->stream;
    QA_HIP(hipStreamSynchronize(s));
    QA_HIP(hipMalloc((void **)backup, sizeof(RKCtl)));
    QA_HIP(hipMalloc((void **)far_time, sizeof(double)));
    QA_HIP(hipMemcpyAsync(*backup, S->ctl, sizeof(RKCtl), hipMemcpyDeviceToDevice, s));
    QA_HIP(hipMemcpyAsync(S->ctl_host, S->ctl, sizeof(RKCtl), hipMemcpyDeviceToHost, s));
    QA_HIP(hipStreamSynchronize(s));
    const double big = 1e300;
    QA_HIP(hipMemcpy(*far_time, &big, sizeof(double), hipMemcpyHostToDevice));
    RKCtl prof = *S->ctl_host;
    prof.times = *far_time;
    prof.ntimes = 1;
    prof.out_idx = 0;
    prof.nsteps_left = 1 << 30;
    prof.nsteps_call = 0;
    QA_HIP(hipMemcpy(S->ctl, &prof, sizeof(RKCtl), hipMemcpyHostToDevice));
    return QA_OK;
}

static int profile_leave(qa_rk_solver *S, RKCtl *backup, double *far_time) {
    hipStream_t s = S->stream;
    QA_HIP(hipMemcpyAsync(S->ctl, backup, sizeof(RKCtl), hipMemcpyDeviceToDevice, s));
    QA_HIP(hipStreamSynchronize(s));
    (void)hipFree(backup);
    (void)hipFree(far_time);
    S->state_set = 0;
    return QA_OK;
}

int qa_rk_profile_stages(qa_rk_solver *S, int reps, double *ms_per_stage) {
    QA_REQUIRE(S != nullptr && reps > 0 && ms_per_stage != nullptr,
               "qa_rk_profile_stages: bad arguments");
    hipStream_t s = S->stream;
    RKCtl *backup = nullptr;
    double *far_time = nullptr;
    {
        const int rc0 = profile_enter(S, &backup, &far_time);
        if (rc0 != QA_OK) return rc0;
    }
    const int nk = S->interpolate ? S->rk_extra : S->rk_step;
    int rc = QA_OK;
    // The stages are launched in solve order (0, 1, ..., nk - 1, 0, 1, ...),
    // so each launch finds the caches as a solve leaves them, and every
    // launch is bracketed by its own pair of events (recorded without host
    // synchronisation; read at the end).
    std::vector<hipEvent_t> ev((size_t)2 * nk * reps, nullptr);
    for (auto &e : ev)
        if (rc == QA_OK && hipEventCreate(&e) != hipSuccess) rc = QA_ERR_HIP;
    for (int r = 0; r < reps && rc == QA_OK; ++r) {
        for (int i = 0; i < nk && rc == QA_OK; ++i) {
            if (i == 0 || i == S->rk_step)
                hipLaunchKernelGGL(rk_set_phase_kernel, dim3(1), dim3(1), 0, s, S->ctl,
                                   i < S->rk_step ? PH_STEP : PH_DENSE);
            const size_t k = (size_t)2 * (r * nk + i);
            if (hipEventRecord(ev[k], s) != hipSuccess) rc = QA_ERR_HIP;
            if (rc == QA_OK) rc = launch_stage(S, i);
            if (rc == QA_OK && hipEventRecord(ev[k + 1], s) != hipSuccess) rc = QA_ERR_HIP;
        }
    }
    if (rc == QA_OK && hipStreamSynchronize(s) != hipSuccess) rc = QA_ERR_HIP;
    for (int i = 0; i < nk && rc == QA_OK; ++i) {
        double tot = 0.0;
        for (int r = 0; r < reps && rc == QA_OK; ++r) {
            const size_t k = (size_t)2 * (r * nk + i);
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) != hipSuccess) rc = QA_ERR_HIP;
            tot += ms;
        }
        ms_per_stage[i] = tot / reps;
    }
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc == QA_ERR_HIP) qa::set_error("qa_rk_profile_stages: event failure");
    const int rc2 = profile_leave(S, backup, far_time);
    return rc != QA_OK ? rc : rc2;
}

int qa_rk_profile_sequence(qa_rk_solver *S, int reps, double *ms_step, double *ms_dense) {
    QA_REQUIRE(S != nullptr && reps > 0 && ms_step != nullptr && ms_dense != nullptr,
               "qa_rk_profile_sequence: bad arguments");
    hipStream_t s = S->stream;
    RKCtl *backup = nullptr;
    double *far_time = nullptr;
    {
        const int rc0 = profile_enter(S, &backup, &far_time);
        if (rc0 != QA_OK) return rc0;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = QA_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = QA_ERR_HIP;
    // part 0: the step stages, part 1: the dense-output stages
    for (int part = 0; part < 2 && rc == QA_OK; ++part) {
        const int i0 = part == 0 ? 0 : S->rk_step;
        const int i1 = part == 0 ? S->rk_step : (S->interpolate ? S->rk_extra : S->rk_step);
        double *out = part == 0 ? ms_step : ms_dense;
        *out = 0.0;
        if (i1 <= i0) continue;
        hipLaunchKernelGGL(rk_set_phase_kernel, dim3(1), dim3(1), 0, s, S->ctl,
                           part == 0 ? PH_STEP : PH_DENSE);
        if (hipEventRecord(e0, s) != hipSuccess) rc = QA_ERR_HIP;
        for (int r = 0; r < reps && rc == QA_OK; ++r)
            for (int i = i0; i < i1 && rc == QA_OK; ++i) rc = launch_stage(S, i);
        float ms = 0.f;
        if (rc == QA_OK && (hipEventRecord(e1, s) != hipSuccess ||
                            hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = QA_ERR_HIP;
        *out = (double)ms / ((double)reps * (i1 - i0));
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc == QA_ERR_HIP) qa::set_error("qa_rk_profile_sequence: event failure");
    const int rc2 = profile_leave(S, backup, far_time);
    return rc != QA_OK ? rc : rc2;
}

// Launch one loop iteration directly (no graph) -- used to time kernels
// without graph replay and by tests of the gating logic.
int qa_rk_iterate(qa_rk_solver *S, int count) {
    QA_REQUIRE(S != nullptr && count >= 0, "qa_rk_iterate: bad arguments");
    for (int k = 0; k < count; ++k) {
        int rc = enqueue_iteration(S, true);
        if (rc) return rc;
    }
    QA_HIP(hipStreamSynchronize(S->stream));
    return QA_OK;
}

}  // extern "C"
