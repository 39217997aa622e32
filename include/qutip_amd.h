/*
 * qutip_amd C ABI -- MI355X (gfx950) kernels for QuTiP's time-evolution hot
 * path: the right-hand-side product and the Runge-Kutta stage updates inside
 * sesolve / mesolve / mcsolve.
 *
 * Conventions
 *   - every function returns an int status: QA_OK (0) or a negative QA_ERR_*;
 *     qa_last_error() returns the message of the last failure on the calling
 *     thread.  No C++ exception crosses the ABI.
 *   - "device pointer" arguments are HBM addresses (from qa_malloc or any
 *     hipMalloc); "host pointer" arguments are ordinary CPU memory.
 *   - complex numbers are interleaved (re, im) doubles (complex128) or floats
 *     (complex64 operator values when vals_fp32 != 0) -- numpy's layout.
 *   - sparse operators: CSR with int32 row pointers and column indices
 *     (QuTiP's idxint, qutip/core/data/base.pxd:3-6), or
 *     SELL-C: rows in slices of C = 64 / L (L = lanes per row, 1/2/4/8);
 *     slice s holds width_s (a multiple of L) entries per row, entry j of
 *     row C s + r at slice_ptr[s] + 64 (j / L) + L r + (j % L), so one
 *     wavefront reads 64 consecutive entries; padding entries have value 0.
 *   - stream: a hipStream_t passed as void*; NULL is the default stream.
 *   - dense matrices are column-major ("Fortran" Dense,
 *     qutip/core/data/dense.pxd:9-33), so vec(rho) is the buffer itself
 *     (column stacking, qutip/core/superoperator.py:291-302).
 */
#ifndef QUTIP_AMD_H
#define QUTIP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QA_VERSION 1

#define QA_OK 0
#define QA_ERR_HIP (-1)   /* a HIP runtime call failed */
#define QA_ERR_ARG (-2)   /* invalid argument (shape, null pointer, range) */
#define QA_ERR_NOMEM (-3) /* device allocation failed */
#define QA_ERR_STATE (-4) /* integrator failure: see the status out-argument */
#define QA_ERR_LIMIT (-5) /* a documented size limit was exceeded */

/* ---- runtime ----------------------------------------------------------- */
const char *qa_last_error(void);
int qa_version(void);
int qa_device_count(int *n);
int qa_set_device(int dev);
int qa_device_info(int dev, char *name, int name_len, int *cu_count, size_t *hbm_bytes);
int qa_mem_info(size_t *free_bytes, size_t *total_bytes);
int qa_malloc(void **ptr, size_t bytes);
int qa_free(void *ptr);
int qa_host_malloc(void **ptr, size_t bytes); /* pinned host memory */
int qa_host_free(void *ptr);
/* h2d / d2h return after the copy completed; d2d is asynchronous */
int qa_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int qa_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int qa_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
/* zero-fill by a kernel (not hipMemset), so it may be recorded into graphs */
int qa_fill_zero(void *dst, size_t bytes, void *stream);
int qa_stream_create(void **stream);
int qa_stream_destroy(void *stream);
int qa_stream_sync(void *stream);
int qa_device_sync(void);
/* events: completion markers on a stream, polled without blocking */
int qa_event_create(void **event);
int qa_event_destroy(void *event);
int qa_event_record(void *event, void *stream);
int qa_event_query(void *event, int *done);

/* ---- data-layer kernels --------------------------------------------------
 * These back the dispatcher specialisations of the device data types.      */

/* out = (accumulate ? out : 0) + scale * A @ x, A CSR nrows x ncols with nnz
 * entries.  Replaces _matmul_csr_vector (qutip/core/data/src/
 * matmul_csr_vector.cpp:107-149) as called by matmul_csr_dense_dense
 * (qutip/core/data/matmul.pyx:226-272) and imatmul_data_dense
 * (matmul.pyx:1207-1215).  FP64 accumulation; lanes per row chosen from
 * nnz / nrows. */
int qa_csr_matvec_z(int nrows, int ncols, const int *rowptr, const int *col, const void *vals,
                    int vals_fp32, int nnz, const void *x, void *out, double scale_re,
                    double scale_im, int accumulate, void *stream);

/* CSR times nvec dense columns (column-major, leading dims ldx, ldo).
 * Replaces the per-column loop of matmul_csr_dense_dense
 * (matmul.pyx:255-265) and _matmul_csr_dense_c_order
 * (qutip/core/data/src/matmul_csr_dense.cpp:20-70). */
int qa_csr_matmat_z(int nrows, int ncols, const int *rowptr, const int *col, const void *vals,
                    int vals_fp32, int nvec, const void *x, int64_t ldx, void *out, int64_t ldo,
                    double scale_re, double scale_im, int accumulate, void *stream);

/* SELL-C version of qa_csr_matvec_z with `lanes` (1, 2, 4, 8) lanes per row
 * and C = 64 / lanes rows per slice (nslices = ceil(nrows / C)). */
int qa_sell_matvec_z(int nrows, int lanes, int nslices, const int *slice_ptr, const int *col,
                     const void *vals, int vals_fp32, const void *x, void *out, double scale_re,
                     double scale_im, int accumulate, void *stream);

/* SELL-C times nvec dense columns (column-major, leading dims ldx >= ncols,
 * ldo >= nrows): every stored entry is read once for each group of 8
 * columns.  The multi-column case of matmul_csr_dense_dense
 * (matmul.pyx:255-265) over the layout of qa_sell_matvec_z. */
int qa_sell_matmat_z(int nrows, int ncols, int lanes, int nslices, const int *slice_ptr,
                     const int *col, const void *vals, int vals_fp32, int nvec, const void *x,
                     int64_t ldx, void *out, int64_t ldo, double scale_re, double scale_im,
                     int accumulate, void *stream);

/* The same product with a coded SELL-C operator: every stored entry is a
 * 16-bit code, low `vbits` bits (6..10) = index into `vtab` (2^vbits
 * complex128 values), high 16 - vbits bits = index into `otab`
 * (2^(16 - vbits) int32 column offsets; column = row + offset, clamped to
 * [0, ncols - 1]).  Four consecutive entries of a lane are one
 * 8-byte word: word slice_ptr[s] + 64 w + lane holds the lane's steps
 * 4w .. 4w + 3 (slice widths are multiples of 4 lanes; slice_ptr counts
 * words).  Same lane -> entry assignment and summation order as
 * qa_sell_matvec_z, hence the same result bits for the same operator.
 * Replaces the CSR product of matmul_csr_dense_dense (matmul.pyx:255-265)
 * for operators whose distinct values and distinct diagonals fit such a
 * split (up to 1024 values with 64 diagonals, 256 with 256, 64 with 1024). */
int qa_sell_coded_matvec_z(int nrows, int ncols, int lanes, int nslices, const int *slice_ptr,
                           const void *codes, const void *vtab, const int *otab, int vbits,
                           const void *x, void *out, double scale_re, double scale_im,
                           int accumulate, void *stream);

/* out = base + sum_j coefs[j] * vecs[j] (real coefficients, nterms <= 16,
 * base may be NULL).  Replaces the copy + zaxpy chain of
 * Explicit_RungeKutta._accumulate (qutip/solver/integrator/explicit_rk.pyx:
 * 432-437 via iadd_dense, qutip/core/data/add.pyx:203-218). */
int qa_lincomb_z(int64_t n, const void *base, int nterms, const void *const *vecs,
                 const double *coefs, void *out, void *stream);

/* out = alpha x + beta y (y may be NULL).  Backs add / iadd / mul / imul
 * (qutip/core/data/add.pyx:324-374, qutip/core/data/mul.pyx:103-139). */
int qa_axpby_z(int64_t n, double are, double aim, const void *x, double bre, double bim,
               const void *y, void *out, void *stream);

/* sqrt(1/n sum (|diff| / (atol + rtol |state|))^2) into *result (host).
 * Replaces wrmn_error_dense (qutip/core/data/ode.pyx:27-52). */
int qa_wrmn_error_z(int64_t n, const void *diff, const void *state, double atol, double rtol,
                    double *result, void *stream);

/* qa_wrmn_error_z without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The per-workgroup
 * partial sums and the result live in a fixed device array of the library
 * (8208 bytes per device: 512 partial sums and 2 result words, kept until
 * the process ends); the launch never uses more than 512 workgroups, so any
 * n > 0 fits.  Same kernels and grid as qa_wrmn_error_z, hence the same
 * bits.  Every call on one device shares that array: calls that overlap on
 * one device (other threads, other streams) race and must be serialised by
 * the caller. */
int qa_wrmn_error_z_noalloc(int64_t n, const void *diff, const void *state, double atol,
                            double rtol, double *result, void *stream);

/* sum |x|^2 into *result (host).  norm.l2 / frobenius are its square root
 * (qutip/core/data/norm.pyx:127-135, 195-216). */
int qa_norm2_z(int64_t n, const void *x, double *result, void *stream);

/* sum conj(x) y into result2[0..1] (host).  Replaces inner_dense
 * (qutip/core/data/inner.pyx:139-164). */
int qa_inner_z(int64_t n, const void *x, const void *y, double *result2, void *stream);

/* trace of a column-major d x d matrix into result2 (host).  Replaces
 * trace_dense / trace_oper_ket_dense (qutip/core/data/trace.pyx:78-86). */
int qa_trace_z(int d, const void *rho, double *result2, void *stream);

/* <psi|O|psi> for O CSR n x n.  Replaces _expect_csr_dense_ket
 * (qutip/core/data/expect.pyx:133-144). */
int qa_expect_csr_ket_z(int n, const int *rowptr, const int *col, const void *vals,
                        int vals_fp32, const void *psi, double *result2, void *stream);

/* tr(O rho) for O CSR d x d and rho column-major d x d.  Replaces
 * _expect_csr_dense_dm (qutip/core/data/expect.pyx:146-158). */
int qa_expect_csr_dm_z(int d, const int *rowptr, const int *col, const void *vals, int vals_fp32,
                       const void *rho, double *result2, void *stream);

/* ---- fused Runge-Kutta engine ---------------------------------------------
 * Replaces Explicit_RungeKutta (qutip/solver/integrator/explicit_rk.pyx) with
 * its derivative QobjEvo.matmul_data (qutip/core/cy/qobjevo.pyx:1103-1116)
 * for a constant generator A in SELL-C form: dy/dt = A y.  The whole step
 * loop (stages, error norm, step control, dense output, output-time
 * interpolation and expectation functionals) runs on the device.
 *
 * The engine keeps the operator pointers it is given (the caller keeps the
 * operator alive) and owns its own stream and work vectors. */
typedef struct qa_rk_solver qa_rk_solver;

/* Operator in SELL-C form with `lanes` lanes per row.
 * Tableau as in explicit_rk.pyx:117-151: a[rk_extra][rk_extra], b[rk_step],
 * c[rk_extra], e[rk_step] (NULL: fixed step), bi[rk_extra][dorder] (NULL: no
 * dense output).  Options as in IntegratorVern7.integrator_options
 * (qutip/solver/integrator/qutip_integrator.py:52-60). */
int qa_rk_create(qa_rk_solver **out, int n, int lanes, int nslices, const int *slice_ptr,
                 const int *col, const void *vals, int vals_fp32, int order, int rk_step,
                 int rk_extra,
                 int dorder, const double *a, const double *b, const double *c,
                 const double *e, const double *bi, double rtol, double atol, int nsteps,
                 double first_step, double min_step, double max_step, int interpolate);
/* The same engine over a coded SELL-C operator (slice_ptr, codes, vtab,
 * otab, vbits as for qa_sell_coded_matvec_z).  The stage products are the terms of the plain
 * SELL-C form summed in the same order, so a solve gives the same bits as
 * qa_rk_create on the same operator, with 2 bytes per stored entry read
 * instead of 20.  Tableau and options as for qa_rk_create
 * (explicit_rk.pyx:117-151, qutip_integrator.py:52-60). */
int qa_rk_create_coded(qa_rk_solver **out, int n, int lanes, int nslices, const int *slice_ptr,
                       const void *codes, const void *vtab, const int *otab, int vbits,
                       int order, int rk_step, int rk_extra, int dorder, const double *a,
                       const double *b, const double *c, const double *e, const double *bi,
                       double rtol, double atol, int nsteps, double first_step,
                       double min_step, double max_step, int interpolate);
/* The matrix-form engine (as qa_rk_create_dense) for banded operators, in
 * FP64: drho/dt = -i H_nh rho + i rho H_nh^dag + sum_c c rho c^dag computed
 * from the diagonals of H_nh and of each c (Dia products:
 * lindblad_matrix_form.pyx:105-203 over matmul_dia_dense_dense,
 * matmul.pyx:429-507).  ndiag diagonals (host arrays): diagonal d belongs to
 * operator diag_op[d] (0 = H_nh, 1 + k = c_k, k < nc), has offset
 * diag_off[d] (|offset| < N) and holds A[r, r + offset] at diag_vals[d][r]
 * (complex128, N entries; entries outside the matrix are ignored).  The
 * products of H_nh's diagonals (2 per diagonal) and of each c's diagonals
 * (pairs) may number at most 96.  rho: N x N complex128 column-major, any
 * N >= 1. */
int qa_rk_create_dia(qa_rk_solver **out, int N, int nc, int ndiag, const int *diag_op,
                     const int *diag_off, const double *diag_vals, int order, int rk_step,
                     int rk_extra, int dorder, const double *a, const double *b, const double *c,
                     const double *e, const double *bi, double rtol, double atol, int nsteps,
                     double first_step, double min_step, double max_step, int interpolate);
/* The same engine for the Lindblad equation in matrix form (mesolve with
 * options {"matrix_form": True}: qutip/core/cy/lindblad_matrix_form.pyx:
 * 105-203): the state is rho, N x N column-major (N a multiple of 16), and
 * drho/dt = -i H_nh rho + i rho H_nh^dag + sum_c c rho c^dag with
 * H_nh = H - i/2 sum c^dag c (host array, N x N complex128 column-major);
 * c_ops: nc such matrices back to back.  The products run on MFMA tiles:
 * precision 0 = FP32 inputs and accumulation, 1 = each FP32 input split
 * into bf16 hi + lo with three products (FP32 accumulation; N % 32 == 0).
 * State, stage combines, error norm and outputs stay complex128 / FP64. */
int qa_rk_create_dense(qa_rk_solver **out, int N, int nc, const double *H_nh, const double *c_ops,
                       int precision, int order, int rk_step, int rk_extra, int dorder,
                       const double *a, const double *b, const double *c, const double *e,
                       const double *bi, double rtol, double atol, int nsteps, double first_step,
                       double min_step, double max_step, int interpolate);
/* Whether the dense engine's states are Hermitian: the
 * assume_hermitian_state switch of LindbladMatrixForm
 * (lindblad_matrix_form.pyx:59-90, set from state.isherm by
 * MESolver._prepare_state, mesolve.py:270-276).  On (and N % 32 == 0),
 * each stage computes the tiles on and above the diagonal and writes each
 * conjugate-transposed below it, so every RHS is exactly Hermitian, as the
 * reference's A + A^dag (:175-189) is; off (the default), the full RHS.
 * No effect on the other engines; get reports what the engine runs. */
int qa_rk_set_hermitian(qa_rk_solver *s, int on);
int qa_rk_get_hermitian(qa_rk_solver *s, int *on);
/* Tile shape of the dense engine's MFMA kernels (the autotuner measures
 * and picks it; no counterpart in the reference, whose zgemm picks its own
 * blocking): tiled 1 = 32 x 32 workgroup tiles staged through LDS
 * (N % 64 == 0), 0 = one 16 x 16 tile per workgroup; nsplit / hsplit =
 * waves per 16 x 16 tile (1, 2, 4, 8) of the full-RHS and of the Hermitian
 * kernels, with N a multiple of 16 nsplit (32 nsplit for the bf16 forms)
 * and of 32 hsplit.  The defaults come from N alone. */
int qa_rk_set_dense_shape(qa_rk_solver *s, int tiled, int nsplit, int hsplit);
int qa_rk_get_dense_shape(qa_rk_solver *s, int *tiled, int *nsplit, int *hsplit);
int qa_rk_destroy(qa_rk_solver *s);

/* out = A x with the engine's generator (device vectors; waits for the
 * result).  For the dense engine this is LindbladMatrixForm.matmul_data. */
int qa_rk_apply(qa_rk_solver *s, const void *x, void *out);
/* The same at time t: out = A(t) x for a time-dependent generator
 * (QobjEvo.matmul_data(t, state), qutip/core/cy/qobjevo.pyx:1103-1116). */
int qa_rk_apply_at(qa_rk_solver *s, double t, const void *x, void *out);

/* ---- time-dependent generators ---------------------------------------------
 * QuTiP's list form A(t) = A_0 + sum_e c_e(t) A_e (QobjEvo elements,
 * qutip/core/cy/_element.pyx:388-446; coefficients, qutip/core/cy/
 * coefficient.pyx:88) with every c_e a program the device evaluates itself:
 * a stack machine over complex128 values, instruction = op | (arg << 8):
 *   0 push consts[arg], 1 push t, 2 add, 3 sub, 4 mul, 5 div, 6 pow, 7 neg,
 *   8 function arg (0 sin, 1 cos, 2 tan, 3 exp, 4 log, 5 log10, 6 sqrt,
 *     7 abs, 8 real, 9 imag, 10 conj, 11 norm = |z|^2, 12 arg, 13 sinh,
 *     14 cosh, 15 tanh, 16 asin, 17 acos, 18 atan, 19 asinh, 20 acosh,
 *     21 atanh, 22 erf of the real part),
 *   9 push the interpolation table at consts[arg] evaluated at t
 *     (InterCoefficient, coefficient.pyx:412-560: consts[arg] = (npts,
 *     rows), consts[arg + 1] = (dt or 0, t0), npts times (real parts), then
 *     `rows` rows of npts polynomial coefficients, the constant term last),
 *   10..15 compare real parts (<, <=, >, >=, ==, !=) to 1 or 0,
 *   16 select (c, a, b -> re c != 0 ? a : b).
 * Program e is code[start[e] .. start[e + 1]) (ncoef <= 8, start[0] = 0,
 * start[ncoef] = ncode); consts: nconst complex128 (host arrays). */

/* Make a coded SELL-C engine (qa_rk_create_coded) time-dependent: entries
 * [vstart[e], vstart[e + 1]) of its value table are element e + 1's values
 * and are multiplied by c_e(t) at every stage time t_prev + c_i dt
 * (explicit_rk.pyx:358-389, 399-410); the other entries are A_0's.  The
 * stage kernels evaluate the programs themselves, so the step loop stays on
 * the device. */
int qa_rk_set_coefficients(qa_rk_solver *s, int ncoef, const int *start, const int *vstart,
                           int ncode, const int *code, int nconst, const double *consts);

/* out_dev[e * ntimes + k] = c_e(times_dev[k]) (complex128, device), the
 * programs as above: the device evaluation of Coefficient.__call__
 * (coefficient.pyx:88-175) at many times. */
int qa_coef_eval(int ncoef, const int *start, int ncode, const int *code, int nconst,
                 const double *consts, int ntimes, const double *times_dev, void *out_dev,
                 void *stream);
void *qa_rk_stream(qa_rk_solver *s);

/* Output functionals evaluated at every output time: functional f has
 * entries p in [fptr[f], fptr[f+1]) with value
 *   kind 0 (linear):    sum_p w_p y[cols_p]            -- tr(O rho), tr(rho)
 *   kind 1 (quadratic): sum_p w_p conj(y[rows_p]) y[cols_p] -- <psi|O|psi>
 * (host arrays; w complex interleaved).  Replaces the per-output e_ops
 * evaluation of Result.add (qutip/solver/result.py:303-340). */
int qa_rk_set_functionals(qa_rk_solver *s, int nfun, const int *kinds, const int *fptr,
                          const int *rows, const int *cols, const double *w);

/* set_initial_value: state y0 (device, n complex128) at t0, first step
 * estimated as in _estimate_first_step (explicit_rk.pyx:204-276). */
int qa_rk_set_state(qa_rk_solver *s, double t0, const void *y0);

/* integrate(t) for each of times[0..ntimes-1] (increasing, host array) --
 * Integrator.run (qutip/solver/integrator/integrator.py:197-212).  With
 * single_step != 0 and ntimes == 1 this is integrate(t, step=True)
 * (mcstep).  fun_out (host, ntimes x nfun complex) receives the
 * functionals; states_out (device, ntimes x npad complex with
 * npad = nslices * 64 / lanes, or NULL)
 * the states.  *status gets the integrator status (explicit_rk.pxd:5-12),
 * *t_reached the time of the last output.  Returns QA_ERR_STATE when the
 * integration failed (status < 0). */
int qa_rk_run(qa_rk_solver *s, int ntimes, const double *times, int single_step,
              double *fun_out, void *states_out, int *status, double *t_reached);

/* The same run without blocking, for many engines in flight at once (the
 * solve server): qa_rk_begin uploads the times and starts the run;
 * qa_rk_advance returns at once -- when the previous chunk of device loop
 * iterations has finished it enqueues the next one -- and sets *done to 1
 * (finished), -1 (failed) or 0; qa_rk_finish waits, copies the functionals
 * and reports the status like qa_rk_run. */
int qa_rk_begin(qa_rk_solver *s, int ntimes, const double *times, int single_step,
                void *states_out);
int qa_rk_advance(qa_rk_solver *s, int *done);
int qa_rk_finish(qa_rk_solver *s, double *fun_out, int *status, double *t_reached);

/* get_state: time and state (device, n complex128) of the last output. */
int qa_rk_get_state(qa_rk_solver *s, double *t, void *y_out);

/* stats[0..7] since the last qa_rk_set_state: step attempts, accepted
 * steps, dense-output preparations, derivative (RHS) evaluations including
 * the two of the first-step estimate, graph launches, status, output index,
 * phase. */
int qa_rk_stats(qa_rk_solver *s, int64_t *stats);

/* Average duration (ms, HIP events on the engine stream, `reps` launches) of
 * each stage kernel into ms_per_stage[0 .. rk_extra-1] (rk_step entries when
 * there is no dense output).  Clobbers the integration state: call
 * qa_rk_set_state afterwards. */
int qa_rk_profile_stages(qa_rk_solver *s, int reps, double *ms_per_stage);

/* Average duration (ms, one pair of HIP events on the engine stream around
 * `reps` back-to-back rounds) of one stage-kernel launch, over the step stages
 * in solve order (*ms_step) and over the dense-output stages (*ms_dense, 0
 * without dense output).  Clobbers the integration state like
 * qa_rk_profile_stages.  No reference counterpart (measurement). */
int qa_rk_profile_sequence(qa_rk_solver *s, int reps, double *ms_step, double *ms_dense);

/* Enqueue `count` loop iterations without the graph and wait (profiling). */
int qa_rk_iterate(qa_rk_solver *s, int count);

/* ---- batched Monte-Carlo trajectories (mcsolve) --------------------------
 * Replaces the per-trajectory loop of MultiTrajSolver.run / map_func
 * (qutip/solver/multitraj.py:169-258, parallel.py:267-337) around
 * MCIntegrator (qutip/solver/mcsolve.py:228-419): all trajectories run in one
 * persistent kernel, one workgroup per trajectory at a time, each with its
 * own step sizes, collapse search and jumps. */

/* Size in bytes of the packed tableau the trajectory kernel reads. */
int qa_mc_tableau_size(void);

/* The LDS plan of a trajectory (small == 0) or whole-solve (small != 0)
 * launch for d rows, without launching: out[0..7] = threads, rows per
 * thread, stages kept in registers, dynamic LDS bytes, static LDS bytes,
 * operator staged in LDS (0/1), bytes for functionals and times, and the
 * dynamic-LDS budget of the shape.  Host arithmetic only (no device call).
 * No reference counterpart (launch configuration). */
int qa_mc_lds_plan(int d, int h_width, int nfun, int ntimes, int fp32, int small,
                   long long *out);

/* Pack a tableau (host arrays as for qa_rk_create) into `out` (host buffer
 * of qa_mc_tableau_size() bytes); the caller copies it to the device. */
int qa_mc_pack_tableau(void *out, int order, int rk_step, int rk_extra, int dorder,
                       const double *a, const double *b, const double *c, const double *e,
                       const double *bi, int interpolate);

/* Run ntraj trajectories of d psi/dt = H_eff psi (H_eff = -iH - 1/2 sum
 * c^dag c, d x d) from psi0 (device, d complex128) over times (device,
 * ntimes).  H_eff is a column-major ELL of width h_width (the largest row
 * length rounded up to a multiple of 4): entry u of row r at h_col / h_vals
 * [u d + r], shorter rows padded with value 0 and any column in [0, d).  c_ops / e_ops: nchan / neops CSR operators concatenated, with
 * absolute row pointers (one block of d + 1 per operator).  Operator values
 * are complex64 when fp32 != 0 (then the state and RHS are complex64 too;
 * norms and expectation values stay FP64).  draws (device, ntraj x
 * max_draws): the uniform numbers each trajectory consumes in the reference
 * order (mcsolve.py:276-279, 388, 405).  Outputs (device): expect_out
 * ntraj x ntimes x neops complex <O> of the normalised state; col_times /
 * col_which ntraj x max_jumps; info ntraj x 4 ints (status, jumps, draws
 * used, RK attempts).  Status 0 is success; -1/-2/-3 the RK failures,
 * -10 collapse time not found, -11 draws exhausted (re-run with more).
 * grid <= 0 launches one workgroup per trajectory.  d <= 1024. */
int qa_mc_run(int d, int ntraj, int ntimes, const double *times_dev, const void *psi0_dev,
              int h_width, const int *h_col, const void *h_vals, int nchan,
              const int *c_rowptr, const int *c_col, const void *c_vals, int neops,
              const int *e_rowptr, const int *e_col, const void *e_vals,
              const void *tableau_dev, const double *draws_dev, int max_draws,
              double rtol, double atol, int nsteps, double first_step, double min_step,
              double max_step, int norm_steps, double norm_t_tol, double norm_tol,
              double norm_min_step, double mc_corr_eps, int fp32, void *expect_out_dev,
              double *col_times_dev, int *col_which_dev, int max_jumps, int *info_dev,
              int grid, void *stream);

/* ---- whole solves of small systems ----------------------------------------
 * Replaces, for a constant generator with n <= 1024, the output loop of
 * Solver.run (qutip/solver/solver_base.py:81-157) around
 * Explicit_RungeKutta.integrate(t) (qutip/solver/integrator/explicit_rk.pyx:
 * 278-331) and the e_ops evaluation of Result.add (result.py:303-340): one
 * workgroup runs the whole integration of one system on the device.  nsys
 * independent systems (initial states y0_dev [nsys][n] complex128, and
 * operators h_stride ELL entries apart, 0 = one shared operator) run in one
 * launch; the batch has no reference counterpart.  The generator A (dy/dt =
 * A y) is a column-major ELL as for qa_mc_run.  Functionals as for
 * qa_rk_set_functionals, on device.  Integrates from t0 through
 * times_dev[0 .. ntimes-1]; fun_out_dev [nsys][ntimes][nfun] complex128;
 * y_final_dev [nsys][n] (or NULL); info_dev [nsys][4]: status (the RK codes),
 * outputs written, dense-output preparations, RK attempts.  The tableau is packed by
 * qa_mc_pack_tableau.  fp32 != 0: complex64 operator, state and stages with
 * FP64 error norm and functionals. */
/* Validate functionals (host arrays, layout of qa_rk_set_functionals) for
 * a state of n entries: fptr[0] == 0 and non-decreasing, every column in
 * [0, n), every row of a quadratic functional in [0, n).  qa_rk_small_run
 * takes its functionals on the device and cannot read them: its callers run
 * this check on the host copies before uploading.  Host only. */
int qa_functionals_check(int n, int nfun, const int *kinds, const int *fptr, const int *rows,
                         const int *cols);

int qa_rk_small_run(int n, int nsys, double t0, int ntimes, const double *times_dev,
                    const void *y0_dev, int h_width, const int *h_col, const void *h_vals,
                    long long h_stride, int nfun, const int *fkind, const int *fptr,
                    const int *frows, const int *fcols, const void *fw, const void *tableau_dev,
                    double rtol, double atol, int nsteps, double first_step, double min_step,
                    double max_step, int fp32, void *fun_out_dev, void *y_final_dev,
                    int *info_dev, int grid, void *stream);

/* ---- batched stochastic trajectories (ssesolve / smesolve) ---------------
 * Replaces the per-trajectory loop of MultiTrajSolver.run around
 * SIntegrator (qutip/solver/stochastic.py:587-900, sode/sode.py:138-260)
 * with the systems of sode/ssystem.pyx:119-250 and the Euler / Platen steps
 * of sode/_sode.pyx:17-120: all ntraj trajectories in one launch, one
 * workgroup per trajectory at a time.
 *   open == 0 (ssesolve): the state is psi (n = d); L = -iH - 1/2 sum c^dag
 *     c and M_i = c_i; drift = L psi + sum_i (e_i / 2 M_i psi - e_i^2 / 8
 *     psi), e_i = <psi|c_i + c_i^dag|psi>; diffusion_i = M_i psi - e_i / 2
 *     psi.
 *   open != 0 (smesolve): the state is vec(rho) (n = d^2, column stacking);
 *     L the Liouvillian of H, sc_ops and c_ops; M_i = spre(c_i) +
 *     spost(c_i^dag); drift = L x, diffusion_i = M_i x - tr(M_i rho) x.
 * method 0 Euler, 1 Platen.  L (width lw) and the M_i (width mw, block i at
 * i mw n entries) are column-major ELL on the device: entry u of row r at
 * u n + r, widths multiples of 4, padding entries value 0 with a column in
 * [0, n).  counts_dev (nout - 1 ints): steps of every output interval (nsteps
 * in total, time step dt); dW_dev [ntraj][nsteps][nsc]: the Wiener
 * increments; x0_dev: the initial state.  e_ops: ne <= 16 CSR operators with
 * absolute row pointers, one block of rows + 1 per operator (rows = n for
 * ssesolve: <psi|O|psi>; rows = d for smesolve: tr(O rho)).  out_dev
 * [ntraj][nout][ne] complex128; info_dev [ntraj]: 0, or -1 when a state
 * stopped being finite.  n <= 1024, 1 <= nsc <= 4; grid <= 0: one workgroup
 * per trajectory. */
int qa_sde_run(int n, int d, int open, int nsc, int method, int ntraj, int nout,
               const int *counts_dev, double dt, long long nsteps, int lw, const int *lcol,
               const void *lval, int mw, const int *mcol, const void *mval, const double *dW_dev,
               const void *x0_dev, int ne, const int *erow, const int *ecol, const void *ew,
               void *out_dev, int *info_dev, int grid, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* QUTIP_AMD_H */
