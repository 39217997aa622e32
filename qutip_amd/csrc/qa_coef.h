// Time-dependent coefficients on the device: the scalar functions c_e(t) of
// a generator in QuTiP's list form A(t) = A_0 + sum_e c_e(t) A_e
// (qutip/core/cy/_element.pyx:388-446, coefficient.pyx:88), evaluated by
// the kernels themselves at every stage time, so the step loop never waits
// on the host.
//
// A coefficient is a small program for a stack machine over complex128
// values, compiled on the host from QuTiP's string coefficients
// (qutip/core/coefficient.py:399-453: the same names -- sin, cos, exp,
// sqrt, abs, real, conj, norm, ... -- with the arguments folded in as
// constants) or built for an interpolated array coefficient
// (InterCoefficient, coefficient.pyx:412-560: piecewise polynomials,
// Horner form).  Instruction = op | (argument << 8).
#pragma once
#include <hip/hip_runtime.h>

#include "qa_common.h"

namespace qa {

constexpr int kMaxCoef = 8;     // time-dependent elements per generator
constexpr int kCoefStack = 16;  // stack depth of a program

enum CoefOp {
    CO_CONST = 0,   // push consts[arg]
    CO_T = 1,       // push t
    CO_ADD = 2,
    CO_SUB = 3,
    CO_MUL = 4,
    CO_DIV = 5,
    CO_POW = 6,
    CO_NEG = 7,
    CO_FN = 8,      // replace the top by fn_arg(top)
    CO_INTERP = 9,  // push the interpolation table at consts[arg ...] evaluated at t
    CO_LT = 10,     // comparisons of real parts: 1 or 0
    CO_LE = 11,
    CO_GT = 12,
    CO_GE = 13,
    CO_EQ = 14,
    CO_NE = 15,
    CO_SELECT = 16  // c, a, b -> (re c != 0 ? a : b)
};

enum CoefFn {
    FN_SIN = 0, FN_COS, FN_TAN, FN_EXP, FN_LOG, FN_LOG10, FN_SQRT, FN_ABS, FN_REAL, FN_IMAG,
    FN_CONJ, FN_NORM, FN_ARG, FN_SINH, FN_COSH, FN_TANH, FN_ASIN, FN_ACOS, FN_ATAN, FN_ASINH,
    FN_ACOSH, FN_ATANH, FN_ERF
};

// The programs of one generator: program e (element e + 1 of the list form)
// is code[start[e] .. start[e+1]); entries [vstart[e], vstart[e+1]) of the
// coded operator's value table belong to it.
struct CoefSet {
    int ncoef;
    int start[kMaxCoef + 1];
    int vstart[kMaxCoef + 1];
    const int *code;
    const double2 *consts;
};

__device__ __forceinline__ double2 c_make(double re, double im) { return make_double2(re, im); }

__device__ __forceinline__ double2 c_div(double2 a, double2 b) {
    // Smith's algorithm (scaled by the larger part of b)
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, d = b.x + b.y * r;
        return c_make((a.x + a.y * r) / d, (a.y - a.x * r) / d);
    }
    const double r = b.x / b.y, d = b.x * r + b.y;
    return c_make((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}

__device__ __forceinline__ double2 c_exp(double2 z) {
    const double e = exp(z.x);
    if (z.y == 0.0) return c_make(e, 0.0);
    return c_make(e * cos(z.y), e * sin(z.y));
}

__device__ __forceinline__ double2 c_log(double2 z) {
    return c_make(log(hypot(z.x, z.y)), atan2(z.y, z.x));
}

__device__ __forceinline__ double2 c_sqrt(double2 z) {
    // principal square root, branch cut on the negative real axis
    if (z.x == 0.0 && z.y == 0.0) return c_make(0.0, z.y);
    const double r = hypot(z.x, z.y);
    if (z.x >= 0.0) {
        const double s = sqrt(0.5 * (r + z.x));
        return c_make(s, z.y / (2.0 * s));
    }
    const double s = sqrt(0.5 * (r - z.x));
    return c_make(fabs(z.y) / (2.0 * s), copysign(s, z.y));
}

__device__ __forceinline__ double2 c_sin(double2 z) {
    if (z.y == 0.0) return c_make(sin(z.x), 0.0);
    return c_make(sin(z.x) * cosh(z.y), cos(z.x) * sinh(z.y));
}

__device__ __forceinline__ double2 c_cos(double2 z) {
    if (z.y == 0.0) return c_make(cos(z.x), 0.0);
    return c_make(cos(z.x) * cosh(z.y), -sin(z.x) * sinh(z.y));
}

__device__ __forceinline__ double2 c_sinh(double2 z) {
    if (z.y == 0.0) return c_make(sinh(z.x), 0.0);
    return c_make(sinh(z.x) * cos(z.y), cosh(z.x) * sin(z.y));
}

__device__ __forceinline__ double2 c_cosh(double2 z) {
    if (z.y == 0.0) return c_make(cosh(z.x), 0.0);
    return c_make(cosh(z.x) * cos(z.y), sinh(z.x) * sin(z.y));
}

// z^w: integer exponents by repeated squaring (exact for small powers of
// real bases, as numpy's complex power does for integral exponents), real
// powers of positive reals with pow, the rest as exp(w log z).
__device__ __forceinline__ double2 c_pow(double2 z, double2 w) {
    if (w.y == 0.0 && w.x == rint(w.x) && fabs(w.x) <= 64.0) {
        long long n = (long long)w.x;
        const bool inv = n < 0;
        if (inv) n = -n;
        double2 r = c_make(1.0, 0.0), b = z;
        while (n) {
            if (n & 1) r = cmul(r, b);
            b = cmul(b, b);
            n >>= 1;
        }
        return inv ? c_div(c_make(1.0, 0.0), r) : r;
    }
    if (z.y == 0.0 && w.y == 0.0 && z.x > 0.0) return c_make(pow(z.x, w.x), 0.0);
    if (z.x == 0.0 && z.y == 0.0) return c_make(0.0, 0.0);
    return c_exp(cmul(w, c_log(z)));
}

__device__ __forceinline__ double2 c_fn(int fn, double2 z) {
    const double2 I = c_make(0.0, 1.0), one = c_make(1.0, 0.0);
    switch (fn) {
        case FN_SIN: return c_sin(z);
        case FN_COS: return c_cos(z);
        case FN_TAN: return z.y == 0.0 ? c_make(tan(z.x), 0.0) : c_div(c_sin(z), c_cos(z));
        case FN_EXP: return c_exp(z);
        case FN_LOG: return c_log(z);
        case FN_LOG10: {
            const double2 l = c_log(z);
            return c_make(l.x / M_LN10, l.y / M_LN10);
        }
        case FN_SQRT: return c_sqrt(z);
        case FN_ABS: return c_make(hypot(z.x, z.y), 0.0);
        case FN_REAL: return c_make(z.x, 0.0);
        case FN_IMAG: return c_make(z.y, 0.0);
        case FN_CONJ: return c_make(z.x, -z.y);
        case FN_NORM: return c_make(z.x * z.x + z.y * z.y, 0.0);
        case FN_ARG: return c_make(atan2(z.y, z.x), 0.0);
        case FN_SINH: return c_sinh(z);
        case FN_COSH: return c_cosh(z);
        case FN_TANH: return z.y == 0.0 ? c_make(tanh(z.x), 0.0) : c_div(c_sinh(z), c_cosh(z));
        // inverse functions: the real function inside its real domain, else
        // the principal branch through log and sqrt
        case FN_ASIN:
            if (z.y == 0.0 && fabs(z.x) <= 1.0) return c_make(asin(z.x), 0.0);
            {  // -i log(i z + sqrt(1 - z^2))
                const double2 l = c_log(cadd(cmul(I, z), c_sqrt(csub(one, cmul(z, z)))));
                return c_make(l.y, -l.x);
            }
        case FN_ACOS:
            if (z.y == 0.0 && fabs(z.x) <= 1.0) return c_make(acos(z.x), 0.0);
            {  // pi/2 - asin z
                const double2 l = c_log(cadd(cmul(I, z), c_sqrt(csub(one, cmul(z, z)))));
                return c_make(0.5 * M_PI - l.y, l.x);
            }
        case FN_ATAN:
            if (z.y == 0.0) return c_make(atan(z.x), 0.0);
            {  // i/2 (log(1 - i z) - log(1 + i z))
                const double2 iz = cmul(I, z);
                const double2 d = csub(c_log(csub(one, iz)), c_log(cadd(one, iz)));
                return c_make(-0.5 * d.y, 0.5 * d.x);
            }
        case FN_ASINH:
            if (z.y == 0.0) return c_make(asinh(z.x), 0.0);
            return c_log(cadd(z, c_sqrt(cadd(cmul(z, z), one))));
        case FN_ACOSH:
            if (z.y == 0.0 && z.x >= 1.0) return c_make(acosh(z.x), 0.0);
            return c_log(cadd(z, cmul(c_sqrt(cadd(z, one)), c_sqrt(csub(z, one)))));
        case FN_ATANH:
            if (z.y == 0.0 && fabs(z.x) < 1.0) return c_make(atanh(z.x), 0.0);
            {
                const double2 d = csub(c_log(cadd(one, z)), c_log(csub(one, z)));
                return c_make(0.5 * d.x, 0.5 * d.y);
            }
        case FN_ERF: return c_make(erf(z.x), 0.0);  // real arguments (host-checked)
        default: return c_make(__builtin_nan(""), 0.0);
    }
}

// An interpolation table at consts[o]: (npts, order + 1), (dt, t0), then
// npts times (real parts), then (order + 1) rows of npts complex
// coefficients, the constant term in the last row (InterCoefficient._call,
// coefficient.pyx:520-545).
__device__ __forceinline__ double2 c_interp(const double2 *__restrict__ k, double t) {
    const int npts = (int)k[0].x, nrow = (int)k[0].y;
    const double dt = k[1].x, t0 = k[1].y;
    const double2 *tl = k + 2;
    const double2 *poly = tl + npts;
    if (t <= tl[0].x) return poly[(size_t)(nrow - 1) * npts];
    if (t >= tl[npts - 1].x) return poly[(size_t)(nrow - 1) * npts + npts - 1];
    int idx;
    if (dt != 0.0) {
        idx = (int)((t - t0) / dt);
    } else {
        int lo = 0, hi = npts;
        for (int it = 0; lo + 1 != hi && it < 64; ++it) {
            const int mid = (lo + hi) >> 1;
            if (t < tl[mid].x)
                hi = mid;
            else
                lo = mid;
        }
        idx = lo;
    }
    idx = idx < 0 ? 0 : (idx > npts - 1 ? npts - 1 : idx);
    if (nrow == 1) return poly[idx];
    const double f = t - tl[idx].x;
    double2 out = c_make(0.0, 0.0);
    for (int i = 0; i < nrow; ++i) {
        out = c_make(out.x * f, out.y * f);
        out = cadd(out, poly[(size_t)i * npts + idx]);
    }
    return out;
}

// Program e of the set at time t (one thread).
__device__ __noinline__ double2 coef_eval(const CoefSet &S, int e, double t) {
    double2 st[kCoefStack];
    int sp = 0;
    for (int pc = S.start[e]; pc < S.start[e + 1]; ++pc) {
        const int ins = S.code[pc];
        const int op = ins & 0xff, arg = ins >> 8;
        if (op == CO_CONST) {
            st[sp++] = S.consts[arg];
        } else if (op == CO_T) {
            st[sp++] = c_make(t, 0.0);
        } else if (op == CO_INTERP) {
            st[sp++] = c_interp(S.consts + arg, t);
        } else if (op == CO_NEG) {
            st[sp - 1] = c_make(-st[sp - 1].x, -st[sp - 1].y);
        } else if (op == CO_FN) {
            st[sp - 1] = c_fn(arg, st[sp - 1]);
        } else if (op == CO_SELECT) {
            const double2 b = st[--sp], a = st[--sp], c = st[--sp];
            st[sp++] = c.x != 0.0 ? a : b;
        } else {
            const double2 b = st[--sp], a = st[--sp];
            double2 r;
            switch (op) {
                case CO_ADD: r = cadd(a, b); break;
                case CO_SUB: r = csub(a, b); break;
                case CO_MUL: r = cmul(a, b); break;
                case CO_DIV: r = c_div(a, b); break;
                case CO_POW: r = c_pow(a, b); break;
                case CO_LT: r = c_make(a.x < b.x ? 1.0 : 0.0, 0.0); break;
                case CO_LE: r = c_make(a.x <= b.x ? 1.0 : 0.0, 0.0); break;
                case CO_GT: r = c_make(a.x > b.x ? 1.0 : 0.0, 0.0); break;
                case CO_GE: r = c_make(a.x >= b.x ? 1.0 : 0.0, 0.0); break;
                case CO_EQ: r = c_make(a.x == b.x && a.y == b.y ? 1.0 : 0.0, 0.0); break;
                case CO_NE: r = c_make(a.x != b.x || a.y != b.y ? 1.0 : 0.0, 0.0); break;
                default: r = c_make(__builtin_nan(""), 0.0); break;
            }
            st[sp++] = r;
        }
        if (sp < 0) sp = 0;
        if (sp > kCoefStack - 1) sp = kCoefStack - 1;
    }
    return sp > 0 ? st[sp - 1] : c_make(0.0, 0.0);
}

// Scale the coded operator's value table in LDS for time t: threads
// 0 .. ncoef-1 evaluate the programs into cf, then every entry of element
// e's range is multiplied by c_e(t).  Called by the whole workgroup (two
// barriers).
__device__ __forceinline__ void coef_scale_table(const CoefSet &S, double t, double2 *vt,
                                                 double2 *cf) {
    if ((int)threadIdx.x < S.ncoef) cf[threadIdx.x] = coef_eval(S, threadIdx.x, t);
    __syncthreads();
    const int v0 = S.vstart[0], v1 = S.vstart[S.ncoef];
    for (int i = v0 + (int)threadIdx.x; i < v1; i += (int)blockDim.x) {
        int e = 0;
        while (e + 1 < S.ncoef && i >= S.vstart[e + 1]) ++e;
        vt[i] = cmul(vt[i], cf[e]);
    }
    __syncthreads();
}

}  // namespace qa
