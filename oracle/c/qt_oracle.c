/*
 * C restatement of QuTiP's CPU time-evolution path -- TEST ORACLE and the
 * timed CPU baseline.  Never linked into the product.
 *
 * What it restates (all paths relative to the reference checkout):
 *   qto_csr_matvec    qutip/core/data/src/matmul_csr_vector.cpp:107-149
 *                     (the portable path; the reference build sets no -m
 *                     flags, meson.build:23-29): out[r] += scale * sum_p
 *                     data[p] * vec[col[p]], real/imag split, row order.
 *   qto_rk_solve      qutip/solver/integrator/explicit_rk.pyx (whole file):
 *                     set_initial_value :204-230, _estimate_first_step
 *                     :232-276, integrate :278-331, _step_in_err :333-356,
 *                     _compute_step :358-389, _error :391-397,
 *                     _prep_dense_out :399-410, _interpolate_step :412-430,
 *                     _accumulate :432-437 with iadd_data's zero skip
 *                     :34-44, _get_timestep :439-449, _recompute_safe_step
 *                     :451-467; the RHS call per stage is the reference's
 *                     zero-fill (imul_data) + accumulate spmv
 *                     (qutip/core/data/matmul.pyx:1207-1215), the stage
 *                     combines are copy + one zaxpy per non-zero coefficient
 *                     (qutip/core/data/add.pyx:203-218).
 *   wrmn_error        qutip/core/data/ode.pyx:27-52.
 *   norms             frobenius (qutip/core/data/norm.pyx, dznrm2): plain
 *                     sqrt(sum |y|^2) here.
 *
 * Output: at every time of tlist, a set of linear functionals
 * f_k(y) = sum_i w_i y[idx_i] of the state (for a vectorised density
 * matrix, tr(O rho) and tr(rho) are such functionals).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cplx;

void qto_csr_matvec(int nrows, const int *row_index, const int *col_index,
                    const double *data_, const double *vec_, double scale_re,
                    double scale_im, double *out_)
{
    const cplx *data = (const cplx *)data_;
    const cplx *vec = (const cplx *)vec_;
    cplx *out = (cplx *)out_;
    for (int row = 0; row < nrows; row++) {
        double dot_re = 0.0, dot_im = 0.0;
        for (int p = row_index[row]; p < row_index[row + 1]; p++) {
            double dr = data[p].re, di = data[p].im;
            double vr = vec[col_index[p]].re, vi = vec[col_index[p]].im;
            dot_re += dr * vr - di * vi;
            dot_im += dr * vi + di * vr;
        }
        out[row].re += scale_re * dot_re - scale_im * dot_im;
        out[row].im += scale_re * dot_im + scale_im * dot_re;
    }
}

enum { AT_FRONT = 2, INTERPOLATED = 1, NORMAL = 0, TOO_MUCH_WORK = -1,
       DT_UNDERFLOW = -2, OUTSIDE_RANGE = -3, NOT_INITIATED = -4 };

typedef struct {
    int n;
    const int *rowptr, *col;
    const double *vals;
    /* a right-hand side other than the CSR product (the banded matrix form) */
    void (*rhs_fn)(void *ctx, const cplx *y, cplx *out);
    void *rhs_ctx;
    /* tableau */
    int order, rk_step, rk_extra, dorder, interpolate, adaptive, fsal;
    const double *a, *b, *c, *e, *bi;
    double rtol, atol, first_step, min_step, max_step;
    int max_numsteps;
    /* state */
    cplx **k, *y, *y_prev, *y_front, *y_temp, *k_fsal;
    double t, t_prev, t_front, dt_safe, dt_int;
    int status;
    long n_rhs, n_attempts;
} rk_t;

static void zero(cplx *v, int n) { memset(v, 0, sizeof(cplx) * (size_t)n); }
static void copy_to(const cplx *src, cplx *dst, int n)
{
    for (int i = 0; i < n; i++) dst[i] = src[i];
}
/* left += factor * right, skipped when factor == 0 (explicit_rk.pyx:34-44) */
static void iadd(cplx *left, const cplx *right, double f, int n)
{
    if (f == 0.0) return;
    for (int i = 0; i < n; i++) {
        left[i].re += f * right[i].re;
        left[i].im += f * right[i].im;
    }
}
static double frob(const cplx *v, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += v[i].re * v[i].re + v[i].im * v[i].im;
    return sqrt(s);
}
static void rhs(rk_t *s, const cplx *y, cplx *out)
{
    s->n_rhs++;
    if (s->rhs_fn) {
        s->rhs_fn(s->rhs_ctx, y, out);
        return;
    }
    zero(out, s->n);
    qto_csr_matvec(s->n, s->rowptr, s->col, s->vals, (const double *)y, 1.0, 0.0,
                   (double *)out);
}
static void accumulate(rk_t *s, cplx *target, const double *factors, double dt, int size)
{
    for (int i = 0; i < size; i++) iadd(target, s->k[i], dt * factors[i], s->n);
}
static double wrmn(const cplx *diff, const cplx *state, double atol, double rtol, int n)
{
    double sum = 0.0;
    for (int i = 0; i < n; i++) {
        double ad = hypot(diff[i].re, diff[i].im);
        double as = hypot(state[i].re, state[i].im);
        double q = ad / (atol + rtol * as);
        sum += q * q;
    }
    return sqrt(sum / n);
}

static double estimate_first_step(rk_t *s, const cplx *y0)
{
    if (!s->adaptive) return 0.0;
    int n = s->n;
    double norm = frob(y0, n);
    double tol = s->atol + norm * s->rtol;
    if (s->fsal) copy_to(s->k_fsal, s->k[0], n);
    else rhs(s, y0, s->k[0]);
    if (norm <= s->atol) norm = 1;
    double tmp_norm = frob(s->k[0], n);
    double factorial = 1.0;
    for (int i = 1; i <= s->order; i++) factorial *= i;
    double dt1, dt2;
    if (tmp_norm >= s->atol * 1e-6)
        dt1 = pow(tol * factorial * pow(norm, s->order), 1.0 / (s->order + 1)) / tmp_norm;
    else
        dt1 = pow(tol * factorial, 1.0 / (s->order + 1)) * norm * 0.5;
    copy_to(y0, s->y_temp, n);
    iadd(s->y_temp, s->k[0], dt1 / 100, n);
    rhs(s, s->y_temp, s->k[1]);
    tmp_norm = frob(s->k[1], n);
    if (tmp_norm >= s->atol * 1e-6)
        dt2 = pow(tol * factorial * pow(norm, s->order), 1.0 / (s->order + 1)) / tmp_norm;
    else
        dt2 = dt1;
    double dt = dt1 < dt2 ? dt1 : dt2;
    if (s->max_step) dt = dt < s->max_step ? dt : s->max_step;
    if (s->min_step) dt = dt > s->min_step ? dt : s->min_step;
    return dt;
}

static double compute_step(rk_t *s, double dt)
{
    int n = s->n, ns = s->rk_extra;
    s->n_attempts++;
    if (s->fsal) copy_to(s->k_fsal, s->k[0], n);
    else rhs(s, s->y_prev, s->k[0]);
    for (int i = 1; i < s->rk_step; i++) {
        copy_to(s->y_prev, s->y_temp, n);
        accumulate(s, s->y_temp, s->a + (size_t)i * ns, dt, i);
        rhs(s, s->y_temp, s->k[i]);
    }
    copy_to(s->y_prev, s->y_front, n);
    accumulate(s, s->y_front, s->b, dt, s->rk_step);
    s->t_front = s->t_prev + dt;
    if (!s->adaptive) return 0.0;
    zero(s->y_temp, n);
    accumulate(s, s->y_temp, s->e, dt, s->rk_step);
    return wrmn(s->y_temp, s->y_front, s->atol, s->rtol, n);
}

static void recompute_safe_step(rk_t *s, double err, double dt)
{
    double factor;
    if (!s->adaptive) return;
    if (err == 0) factor = 10;
    else {
        factor = 0.9 * pow(err, -1.0 / (s->order + 1));
        factor = factor < 10 ? factor : 10;
        factor = factor > 0.2 ? factor : 0.2;
    }
    s->dt_safe = dt * factor;
    if (s->max_step) s->dt_safe = s->dt_safe < s->max_step ? s->dt_safe : s->max_step;
    if (s->min_step) s->dt_safe = s->dt_safe > s->min_step ? s->dt_safe : s->min_step;
}

static double get_timestep(rk_t *s, double t)
{
    double dt_needed = t - s->t_prev;
    if (!s->adaptive) return dt_needed;
    if (s->interpolate) return s->dt_safe;
    if (dt_needed <= s->dt_safe) return dt_needed;
    return dt_needed / ((int)(dt_needed / s->dt_safe) + 1);
}

static int step_in_err(rk_t *s, double t, int max_step)
{
    double error = 1;
    int nsteps = 0;
    while (error >= 1) {
        double dt = get_timestep(s, t);
        error = compute_step(s, dt);
        s->dt_int = dt;
        recompute_safe_step(s, error, dt);
        if (dt == s->min_step && error > 1) { s->status = DT_UNDERFLOW; break; }
        nsteps++;
        if (nsteps > max_step) { s->status = TOO_MUCH_WORK; break; }
    }
    if (s->fsal) copy_to(s->k[s->rk_step - 1], s->k_fsal, s->n);
    return nsteps;
}

static void prep_dense_out(rk_t *s)
{
    int n = s->n, ns = s->rk_extra;
    double dt = s->dt_int;
    for (int i = s->rk_step; i < s->rk_extra; i++) {
        copy_to(s->y_prev, s->y_temp, n);
        accumulate(s, s->y_temp, s->a + (size_t)i * ns, dt, i);
        rhs(s, s->y_temp, s->k[i]);
    }
}

static void interpolate_step(rk_t *s, double t, cplx *out)
{
    double bf[64];
    double tau = (t - s->t_prev) / s->dt_int;
    for (int i = 0; i < s->rk_extra; i++) {
        double v = 0.0;
        for (int j = s->dorder - 1; j >= 0; j--) {
            v += s->bi[(size_t)i * s->dorder + j];
            v *= tau;
        }
        bf[i] = v;
    }
    copy_to(s->y_prev, out, s->n);
    accumulate(s, out, bf, s->dt_int, s->rk_extra);
}

static void integrate(rk_t *s, double t, int step)
{
    int nsteps_left = s->max_numsteps;
    if (t == s->t) return;
    if (t < s->t_prev) { s->status = OUTSIDE_RANGE; return; }
    if (s->interpolate && t < s->t_front) {
        if (s->status != INTERPOLATED) {
            prep_dense_out(s);
            s->status = INTERPOLATED;
        }
        interpolate_step(s, t, s->y);
        s->t = t;
        return;
    }
    s->status = NORMAL;
    if (step && s->t < s->t_front && t > s->t_front) t = s->t_front;
    while (s->t_front < t && s->status >= 0) {
        copy_to(s->y_front, s->y_prev, s->n);
        s->t_prev = s->t_front;
        nsteps_left -= step_in_err(s, t, nsteps_left);
        if (step) break;
    }
    if (s->status < 0) return;
    if (s->t_front > t) {
        prep_dense_out(s);
        s->status = INTERPOLATED;
        s->t = t;
        interpolate_step(s, t, s->y);
    } else {
        s->status = AT_FRONT;
        s->t = s->t_front;
        copy_to(s->y_front, s->y, s->n);
    }
}

static void eval_functionals(const cplx *y, int nfun, const int *fptr, const int *fidx,
                             const cplx *fw, cplx *out)
{
    for (int f = 0; f < nfun; f++) {
        double re = 0, im = 0;
        for (int p = fptr[f]; p < fptr[f + 1]; p++) {
            cplx w = fw[p], v = y[fidx[p]];
            re += w.re * v.re - w.im * v.im;
            im += w.re * v.im + w.im * v.re;
        }
        out[f].re = re;
        out[f].im = im;
    }
}

/* Solve dy/dt = A y over tlist; writes nfun functionals per output time into
 * fun_out[t * nfun + f] (complex interleaved) and the final state into
 * y_final.  stats[0..2] = RHS calls, step attempts, final status.
 * Returns the integrator status (>= 0 on success). */
int qto_rk_solve(int n, const int *rowptr, const int *col, const double *vals,
                 const double *y0, int ntimes, const double *tlist,
                 int order, int rk_step, int rk_extra, int dorder,
                 const double *a, const double *b, const double *c,
                 const double *e, const double *bi,
                 double rtol, double atol, int nsteps, double first_step,
                 double min_step, double max_step, int interpolate,
                 int nfun, const int *fun_ptr, const int *fun_idx, const double *fun_w,
                 double *fun_out, double *y_final, long *stats)
{
    rk_t s;
    memset(&s, 0, sizeof(s));
    s.n = n; s.rowptr = rowptr; s.col = col; s.vals = vals;
    s.order = order; s.rk_step = rk_step; s.rk_extra = rk_extra; s.dorder = dorder;
    s.a = a; s.b = b; s.c = c; s.e = e; s.bi = bi;
    s.adaptive = e != NULL;
    s.interpolate = (bi != NULL) && interpolate;
    s.rtol = rtol; s.atol = atol; s.first_step = first_step;
    s.min_step = min_step ? min_step : 1e-15;
    s.max_step = max_step; s.max_numsteps = nsteps;
    s.dt_safe = atol;
    /* first_same_as_last: np.allclose(a[rk_step-1, :rk_step], b, atol=1e-14) */
    s.fsal = 1;
    for (int j = 0; j < rk_step; j++) {
        double x = a[(size_t)(rk_step - 1) * rk_extra + j], yv = b[j];
        if (fabs(x - yv) > 1e-14 + 1e-5 * fabs(yv)) s.fsal = 0;
    }
    int nk = s.interpolate ? rk_extra : rk_step;
    s.k = (cplx **)malloc(sizeof(cplx *) * (size_t)nk);
    for (int i = 0; i < nk; i++) s.k[i] = (cplx *)calloc((size_t)n, sizeof(cplx));
    s.y = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    s.y_prev = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    s.y_front = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    s.y_temp = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    s.k_fsal = (cplx *)calloc((size_t)n, sizeof(cplx));
    /* set_initial_value */
    copy_to((const cplx *)y0, s.y, n);
    copy_to(s.y, s.y_prev, n);
    copy_to(s.y, s.y_front, n);
    s.t = s.t_prev = s.t_front = tlist[0];
    s.dt_int = 0;
    s.status = NORMAL;
    if (s.fsal) rhs(&s, s.y, s.k_fsal);
    s.dt_safe = s.first_step ? s.first_step : estimate_first_step(&s, s.y);

    cplx *fo = (cplx *)fun_out;
    eval_functionals(s.y, nfun, fun_ptr, fun_idx, (const cplx *)fun_w, fo);
    for (int it = 1; it < ntimes; it++) {
        integrate(&s, tlist[it], 0);
        if (s.status < 0) break;
        eval_functionals(s.y, nfun, fun_ptr, fun_idx, (const cplx *)fun_w,
                         fo + (size_t)it * nfun);
    }
    if (y_final) copy_to(s.y, (cplx *)y_final, n);
    if (stats) { stats[0] = s.n_rhs; stats[1] = s.n_attempts; stats[2] = s.status; }
    int st = s.status;
    for (int i = 0; i < nk; i++) free(s.k[i]);
    free(s.k); free(s.y); free(s.y_prev); free(s.y_front); free(s.y_temp); free(s.k_fsal);
    return st;
}

/* ---- mcsolve trajectories ------------------------------------------------
 * One trajectory of MCIntegrator (qutip/solver/mcsolve.py:245-406) on the
 * RK above in its integrate(t, step=True) form (qutip_integrator.py:84-87):
 * the target norm is the first draw (set_state :276-279); each mcstep
 * that crosses the target is followed by the log-ratio secant search of
 * _find_collapse_time (:321-369) and the jump of _do_collapse (:371-406:
 * the channel by <c^dag c> with one draw when there are several channels,
 * renormalisation, a fresh target draw, and a new RK started at the jump).
 * The draws are numpy's uniform stream of the trajectory's generator,
 * passed in.  Output: <O> of the normalised state at every time of tlist.
 * stats[0..3] = RHS calls, attempts, jumps, draws used.  Returns 0, the RK
 * status when negative, -10 when the collapse time is not found, -11 when
 * the draws run out. */
static void rk_setup(rk_t *s, int n, const int *rowptr, const int *col, const double *vals,
                     int order, int rk_step, int rk_extra, int dorder, const double *a,
                     const double *b, const double *c, const double *e, const double *bi,
                     double rtol, double atol, int nsteps)
{
    memset(s, 0, sizeof(*s));
    s->n = n; s->rowptr = rowptr; s->col = col; s->vals = vals;
    s->order = order; s->rk_step = rk_step; s->rk_extra = rk_extra; s->dorder = dorder;
    s->a = a; s->b = b; s->c = c; s->e = e; s->bi = bi;
    s->adaptive = e != NULL;
    s->interpolate = bi != NULL;
    s->rtol = rtol; s->atol = atol;
    s->min_step = 1e-15;
    s->max_numsteps = nsteps;
    s->fsal = 1;
    for (int j = 0; j < rk_step; j++) {
        double x = a[(size_t)(rk_step - 1) * rk_extra + j], yv = b[j];
        if (fabs(x - yv) > 1e-14 + 1e-5 * fabs(yv)) s->fsal = 0;
    }
    int nk = s->interpolate ? rk_extra : rk_step;
    s->k = (cplx **)malloc(sizeof(cplx *) * (size_t)nk);
    for (int i = 0; i < nk; i++) s->k[i] = (cplx *)calloc((size_t)n, sizeof(cplx));
    s->y = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    s->y_prev = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    s->y_front = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    s->y_temp = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    s->k_fsal = (cplx *)calloc((size_t)n, sizeof(cplx));
}

static void rk_free(rk_t *s)
{
    int nk = s->interpolate ? s->rk_extra : s->rk_step;
    for (int i = 0; i < nk; i++) free(s->k[i]);
    free(s->k); free(s->y); free(s->y_prev); free(s->y_front); free(s->y_temp);
    free(s->k_fsal);
}

/* set_initial_value (explicit_rk.pyx:204-230) */
static void rk_set_state(rk_t *s, double t, const cplx *y0)
{
    int n = s->n;
    copy_to(y0, s->y, n);
    copy_to(y0, s->y_prev, n);
    copy_to(y0, s->y_front, n);
    s->t = s->t_prev = s->t_front = t;
    s->dt_int = 0;
    s->status = NORMAL;
    s->dt_safe = s->atol;
    if (s->fsal) rhs(s, s->y, s->k_fsal);
    s->dt_safe = s->first_step ? s->first_step : estimate_first_step(s, s->y);
}

static double nrm2(const cplx *v, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += v[i].re * v[i].re + v[i].im * v[i].im;
    return s;
}

/* <v|O|v> for a CSR O */
static cplx quad(int n, const int *rp, const int *cl, const cplx *w, const cplx *v)
{
    cplx acc = {0.0, 0.0};
    for (int r = 0; r < n; r++) {
        double re = 0.0, im = 0.0;
        for (int p = rp[r]; p < rp[r + 1]; p++) {
            cplx a = w[p], x = v[cl[p]];
            re += a.re * x.re - a.im * x.im;
            im += a.re * x.im + a.im * x.re;
        }
        acc.re += v[r].re * re + v[r].im * im;
        acc.im += v[r].re * im - v[r].im * re;
    }
    return acc;
}

int qto_mc_traj(int d, const int *h_rowptr, const int *h_col, const double *h_vals,
                int nchan, const int *c_rowptr, const int *c_col, const double *c_vals,
                const int *n_rowptr, const int *n_col, const double *n_vals,
                int neops, const int *e_rowptr, const int *e_col, const double *e_vals,
                const double *psi0, int ntimes, const double *tlist,
                int order, int rk_step, int rk_extra, int dorder,
                const double *a, const double *b, const double *c,
                const double *e, const double *bi, double rtol, double atol, int nsteps,
                int norm_steps, double norm_t_tol, double norm_tol, double norm_min_step,
                double mc_corr_eps, const double *draws, int max_draws,
                double *expect_out, long *stats)
{
    rk_t s;
    rk_setup(&s, d, h_rowptr, h_col, h_vals, order, rk_step, rk_extra, dorder, a, b, c, e, bi,
             rtol, atol, nsteps);
    cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)d);
    cplx *st = (cplx *)malloc(sizeof(cplx) * (size_t)d);
    cplx *out = (cplx *)expect_out;
    int used = 0, jumps = 0, rc = 0;
#define DRAW() (used < max_draws ? draws[used++] : -1.0)
    double target = DRAW();
    rk_set_state(&s, tlist[0], (const cplx *)psi0);
    const cplx *ev = (const cplx *)e_vals;
    double nn0 = nrm2(s.y, d);
    for (int k = 0; k < neops; k++) {
        cplx q = quad(d, e_rowptr + (size_t)k * (d + 1), e_col, ev, s.y);
        out[k].re = q.re / nn0;
        out[k].im = q.im / nn0;
    }
    for (int it = 1; it < ntimes && rc == 0; it++) {
        double t = tlist[it];
        double t_old = s.t;
        copy_to(s.y, st, d);
        double norm_old = nrm2(st, d);
        while (t_old < t) {
            integrate(&s, t, 1);
            if (s.status < 0) { rc = s.status; break; }
            double t_step = s.t;
            double nrm = nrm2(s.y, d);
            if (nrm <= target) {
                /* _find_collapse_time */
                double t_prev = t_old, t_final = t_step, norm = nrm, t_guess = t_final;
                int tries = 0;
                while (tries < norm_steps) {
                    tries++;
                    if (t_final - t_prev < norm_t_tol) { t_guess = t_final; break; }
                    double dt = t_final - t_prev;
                    double ratio = log(norm_old / target) / log(norm_old / norm);
                    if (ratio < norm_min_step) ratio = norm_min_step;
                    if (ratio > 1 - norm_min_step) ratio = 1 - norm_min_step;
                    t_guess = t_prev + dt * ratio;
                    if (t_guess - t_prev < norm_t_tol) t_guess = t_prev + norm_t_tol;
                    integrate(&s, t_guess, 1);
                    if (s.status < 0) { rc = s.status; break; }
                    double ng = nrm2(s.y, d);
                    if (fabs(target - ng) < norm_tol * target) break;
                    else if (ng < target) { t_final = t_guess; norm = ng; }
                    else { t_prev = t_guess; norm_old = ng; }
                }
                if (rc) break;
                if (tries >= norm_steps) { rc = -10; break; }
                /* _do_collapse at t_guess with the state s.y */
                int which = 0;
                if (nchan > 1) {
                    double probs[64], tot = 0.0;
                    for (int ch = 0; ch < nchan; ch++) {
                        cplx q = quad(d, n_rowptr + (size_t)ch * (d + 1), n_col,
                                      (const cplx *)n_vals, s.y);
                        probs[ch] = q.re;
                        tot += q.re;
                    }
                    double u = DRAW();
                    if (u < 0) { rc = -11; break; }
                    double tg = tot * u - probs[0];
                    while (tg > 0 && which <= nchan) {
                        which++;
                        tg -= probs[which];
                    }
                }
                zero(tmp, d);
                qto_csr_matvec(d, c_rowptr + (size_t)which * (d + 1), c_col, c_vals,
                               (const double *)s.y, 1.0, 0.0, (double *)tmp);
                double nj = sqrt(nrm2(tmp, d));
                if (nj < mc_corr_eps) {
                    double ns = sqrt(nrm2(s.y, d));
                    for (int i = 0; i < d; i++) { tmp[i].re = s.y[i].re / ns; tmp[i].im = s.y[i].im / ns; }
                } else {
                    for (int i = 0; i < d; i++) { tmp[i].re /= nj; tmp[i].im /= nj; }
                    jumps++;
                    target = DRAW();
                    if (target < 0) { rc = -11; break; }
                }
                long nr = s.n_rhs, na = s.n_attempts;
                rk_set_state(&s, t_guess, tmp);
                s.n_rhs += nr;
                s.n_attempts += na;
                t_old = s.t;
                norm_old = 1.0;
            } else {
                t_old = t_step;
                norm_old = nrm;
            }
        }
        if (rc) break;
        double nn = nrm2(s.y, d);
        for (int k = 0; k < neops; k++) {
            cplx q = quad(d, e_rowptr + (size_t)k * (d + 1), e_col, ev, s.y);
            out[(size_t)it * neops + k].re = q.re / nn;
            out[(size_t)it * neops + k].im = q.im / nn;
        }
    }
#undef DRAW
    if (stats) { stats[0] = s.n_rhs; stats[1] = s.n_attempts; stats[2] = jumps; stats[3] = used; }
    free(tmp);
    free(st);
    rk_free(&s);
    return rc;
}

/* ---- the matrix form with banded operators --------------------------------
 * LindbladMatrixForm.matmul_data with assume_hermitian_state
 * (qutip/core/cy/lindblad_matrix_form.pyx:170-186) when H_nh and the c_ops
 * are Dia (diagonal storage, the products of matmul_dia_dense_dense,
 * qutip/core/data/matmul.pyx:429-507): A = -i H_nh rho, then for every c
 * T = 1/2 c rho and A += T c^dag, then drho/dt = A + A^dag.  rho is N x N
 * column-major.  A diagonal is (offset o, values v[r] = M[r, r + o]); entries
 * outside the matrix are ignored. */
typedef struct {
    int N, nh, nc;
    const int *h_off;       /* nh offsets */
    const cplx *h_val;      /* nh x N */
    const int *c_ndiag;     /* nc */
    const int *c_off;       /* all c diagonals back to back */
    const cplx *c_val;
    cplx *T;                /* N x N work */
} dia_mf_t;

static void dia_times_dense(int N, int nd, const int *off, const cplx *val, const cplx *x,
                            cplx *out, double sre, double sim)
{
    /* out[r, j] += s sum_d v_d[r] x[r + o_d, j] */
    for (int d = 0; d < nd; d++) {
        int o = off[d];
        const cplx *v = val + (size_t)d * N;
        int r0 = o < 0 ? -o : 0, r1 = o > 0 ? N - o : N;
        for (int j = 0; j < N; j++) {
            const cplx *xc = x + (size_t)j * N;
            cplx *oc = out + (size_t)j * N;
            for (int r = r0; r < r1; r++) {
                double ar = v[r].re * sre - v[r].im * sim, ai = v[r].re * sim + v[r].im * sre;
                cplx xv = xc[r + o];
                oc[r].re += ar * xv.re - ai * xv.im;
                oc[r].im += ar * xv.im + ai * xv.re;
            }
        }
    }
}

static void dense_times_dia_dag(int N, int nd, const int *off, const cplx *val, const cplx *T,
                                cplx *out)
{
    /* out[r, j] += sum_d T[r, j + o_d] conj(c[j, j + o_d]) = (T c^dag)[r, j] */
    for (int d = 0; d < nd; d++) {
        int o = off[d];
        const cplx *v = val + (size_t)d * N;
        int j0 = o < 0 ? -o : 0, j1 = o > 0 ? N - o : N;
        for (int j = j0; j < j1; j++) {
            double cr = v[j].re, ci = -v[j].im;
            const cplx *tc = T + (size_t)(j + o) * N;
            cplx *oc = out + (size_t)j * N;
            for (int r = 0; r < N; r++) {
                oc[r].re += tc[r].re * cr - tc[r].im * ci;
                oc[r].im += tc[r].re * ci + tc[r].im * cr;
            }
        }
    }
}

static void dia_mf_rhs(void *ctx, const cplx *rho, cplx *out)
{
    dia_mf_t *m = (dia_mf_t *)ctx;
    int N = m->N;
    size_t NN = (size_t)N * N;
    memset(out, 0, sizeof(cplx) * NN);
    dia_times_dense(N, m->nh, m->h_off, m->h_val, rho, out, 0.0, -1.0);
    int base = 0;
    for (int k = 0; k < m->nc; k++) {
        int nd = m->c_ndiag[k];
        memset(m->T, 0, sizeof(cplx) * NN);
        dia_times_dense(N, nd, m->c_off + base, m->c_val + (size_t)base * N, rho, m->T, 0.5, 0.0);
        dense_times_dia_dag(N, nd, m->c_off + base, m->c_val + (size_t)base * N, m->T, out);
        base += nd;
    }
    /* out += out^dag (iadd_adjoint_dense) */
    for (int j = 0; j < N; j++)
        for (int r = 0; r <= j; r++) {
            cplx a = out[(size_t)j * N + r], b = out[(size_t)r * N + j];
            cplx v = {a.re + b.re, a.im - b.im};
            out[(size_t)j * N + r] = v;
            out[(size_t)r * N + j].re = v.re;
            out[(size_t)r * N + j].im = -v.im;
        }
}

/* qto_rk_solve on the banded matrix form; t_end: stop after this time (a
 * bounded sample of the window).  stats as qto_rk_solve. */
int qto_rk_solve_dia(int N, int nh, const int *h_off, const double *h_val, int nc,
                     const int *c_ndiag, const int *c_off, const double *c_val,
                     const double *rho0, int ntimes, const double *tlist,
                     int order, int rk_step, int rk_extra, int dorder,
                     const double *a, const double *b, const double *c,
                     const double *e, const double *bi, double rtol, double atol, int nsteps,
                     int nfun, const int *fun_ptr, const int *fun_idx, const double *fun_w,
                     double *fun_out, long *stats)
{
    rk_t s;
    int n = N * N;
    rk_setup(&s, n, NULL, NULL, NULL, order, rk_step, rk_extra, dorder, a, b, c, e, bi, rtol,
             atol, nsteps);
    dia_mf_t m = {N, nh, nc, h_off, (const cplx *)h_val, c_ndiag, c_off, (const cplx *)c_val,
                  (cplx *)malloc(sizeof(cplx) * (size_t)n)};
    s.rhs_fn = dia_mf_rhs;
    s.rhs_ctx = &m;
    rk_set_state(&s, tlist[0], (const cplx *)rho0);
    cplx *fo = (cplx *)fun_out;
    eval_functionals(s.y, nfun, fun_ptr, fun_idx, (const cplx *)fun_w, fo);
    for (int it = 1; it < ntimes; it++) {
        integrate(&s, tlist[it], 0);
        if (s.status < 0) break;
        eval_functionals(s.y, nfun, fun_ptr, fun_idx, (const cplx *)fun_w,
                         fo + (size_t)it * nfun);
    }
    if (stats) { stats[0] = s.n_rhs; stats[1] = s.n_attempts; stats[2] = s.status; }
    int st = s.status;
    free(m.T);
    rk_free(&s);
    return st;
}
