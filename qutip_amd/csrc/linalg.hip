// Data-layer kernels behind the dispatcher specialisations of the device
// types (matmul / add / mul / expect / inner / norm / wrmn_error / trace).
//
// Each entry point cites the reference kernel or dispatcher it replaces.
// Operators: CSR (int32 row pointers and columns, complex128 or complex64
// values) and SELL-C (rows in slices of one wavefront, entries stored
// column-major inside a slice so the 64 lanes of a wave read consecutive
// addresses; padding entries have value 0).
#include <mutex>

#include "qa_api_internal.h"
#include "qa_common.h"

using namespace qa;

// ---------------------------------------------------------------------------
// CSR matvec: G lanes per row (G = 1..64, power of two), FP64 accumulation.
// out[r] = (accumulate ? out[r] : 0) + scale * sum_p A[r, col[p]] x[col[p]]
// ---------------------------------------------------------------------------
template <typename V, int G>
__global__ __launch_bounds__(256) void csr_matvec_kernel(int nrows, const int *__restrict__ rowptr,
                                                         const int *__restrict__ col,
                                                         const V *__restrict__ vals,
                                                         const double2 *__restrict__ x,
                                                         double2 *__restrict__ out, double2 scale,
                                                         int accumulate) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t row = gid / G;
    const int sub = (int)(gid % G);
    double2 acc = make_double2(0.0, 0.0);
    if (row < nrows) {
        const int beg = rowptr[row], end = rowptr[row + 1];
        for (int p = beg + sub; p < end; p += G) cfma(acc, ld(vals, p), x[col[p]]);
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        acc.x += __shfl_xor(acc.x, off, QA_WAVE);
        acc.y += __shfl_xor(acc.y, off, QA_WAVE);
    }
    if (row < nrows && sub == 0) {
        double2 r = cmul(scale, acc);
        if (accumulate) r = cadd(r, out[row]);
        out[row] = r;
    }
}

// Columns of the dense block one thread or wavefront applies each loaded
// operator entry to (the operator is read once per kMatCols columns).
constexpr int kMatCols = 8;

// CSR times a block of nvec dense columns (column-major, leading dimensions
// ldx / ldo): one thread per row and kMatCols columns (grid y), so each
// stored entry is read once for kMatCols columns.  Replaces the per-column
// loop of matmul_csr_dense_dense (qutip/core/data/matmul.pyx:255-265).
template <typename V>
__global__ __launch_bounds__(256) void csr_matmat_kernel(int nrows, const int *__restrict__ rowptr,
                                                         const int *__restrict__ col,
                                                         const V *__restrict__ vals, int nvec,
                                                         const double2 *__restrict__ x, int64_t ldx,
                                                         double2 *__restrict__ out, int64_t ldo,
                                                         double2 scale, int accumulate) {
    const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int v0 = blockIdx.y * kMatCols;
    const int nv = min(kMatCols, nvec - v0);  // uniform over the grid layer
    if (row >= nrows) return;
    const double2 *xb = x + (int64_t)v0 * ldx;
    double2 acc[kMatCols];
#pragma unroll
    for (int v = 0; v < kMatCols; ++v) acc[v] = make_double2(0.0, 0.0);
    for (int p = rowptr[row]; p < rowptr[row + 1]; ++p) {
        const int c = col[p];
        const double2 a = ld(vals, p);
#pragma unroll
        for (int v = 0; v < kMatCols; ++v)
            if (v < nv) cfma(acc[v], a, xb[(int64_t)v * ldx + c]);
    }
#pragma unroll
    for (int v = 0; v < kMatCols; ++v) {
        if (v >= nv) break;
        double2 r = cmul(scale, acc[v]);
        double2 *o = out + (int64_t)(v0 + v) * ldo + row;
        if (accumulate) r = cadd(r, *o);
        *o = r;
    }
}

// SELL-C times a block of dense columns: one wavefront per slice (the lane
// -> entry assignment of sell_row_dot) and kMatCols columns per grid layer.
// Each lane loads its entries (column and value) once, four steps at a
// time, and applies them to every column of the layer; the L partial sums
// of a row are combined with lane shuffles.  Grid: (ceil(nslices 64 / 256),
// ceil(nvec / kMatCols)).
template <typename V, int L>
__global__ __launch_bounds__(256) void sell_matmat_kernel(int nrows, int nslices,
                                                          const int *__restrict__ slice_ptr,
                                                          const int *__restrict__ col,
                                                          const V *__restrict__ vals, int nvec,
                                                          const double2 *__restrict__ x,
                                                          int64_t ldx, double2 *__restrict__ out,
                                                          int64_t ldo, double2 scale,
                                                          int accumulate) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int slice = (int)(t / QA_WAVE), lane = (int)(t % QA_WAVE);
    if (slice >= nslices) return;
    const int v0 = blockIdx.y * kMatCols;
    const int nv = min(kMatCols, nvec - v0);  // uniform over the grid layer
    int base, steps;
    sell_slice(slice_ptr, slice, base, steps);
    const double2 *xb = x + (int64_t)v0 * ldx;
    double2 acc[kMatCols];
#pragma unroll
    for (int v = 0; v < kMatCols; ++v) acc[v] = make_double2(0.0, 0.0);
    const int p0 = base + lane;
    for (int k = 0; k < steps; k += 4) {
        const int m = min(4, steps - k);  // uniform over the slice
        int c[4];
        double2 a[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = p0 + (k + min(u, m - 1)) * QA_WAVE;
            c[u] = col[p];
            a[u] = ld(vals, p);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (u >= m) break;
#pragma unroll
            for (int v = 0; v < kMatCols; ++v)
                if (v < nv) cfma(acc[v], a[u], xb[(int64_t)v * ldx + c[u]]);
        }
    }
    const int64_t r = (int64_t)slice * (QA_WAVE / L) + lane / L;
#pragma unroll
    for (int v = 0; v < kMatCols; ++v) {
#pragma unroll
        for (int off = L / 2; off > 0; off >>= 1) {
            acc[v].x += __shfl_xor(acc[v].x, off, QA_WAVE);
            acc[v].y += __shfl_xor(acc[v].y, off, QA_WAVE);
        }
    }
    if (lane % L != 0 || r >= nrows) return;
#pragma unroll
    for (int v = 0; v < kMatCols; ++v) {
        if (v >= nv) break;
        double2 o = cmul(scale, acc[v]);
        double2 *dst = out + (int64_t)(v0 + v) * ldo + r;
        if (accumulate) o = cadd(o, *dst);
        *dst = o;
    }
}

template <typename V, int L>
__global__ __launch_bounds__(256) void sell_matvec_kernel(int nrows, int nslices,
                                                          const int *__restrict__ slice_ptr,
                                                          const int *__restrict__ col,
                                                          const V *__restrict__ vals,
                                                          const double2 *__restrict__ x,
                                                          double2 *__restrict__ out, double2 scale,
                                                          int accumulate) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int slice = (int)(t / QA_WAVE), lane = (int)(t % QA_WAVE);
    if (slice >= nslices) return;
    int base, steps;
    sell_slice(slice_ptr, slice, base, steps);
    const double2 acc = sell_row_dot<L>(col, vals, x, base, steps, lane);
    const int64_t r = (int64_t)slice * (QA_WAVE / L) + lane / L;
    if (lane % L == 0 && r < nrows) {
        double2 o = cmul(scale, acc);
        if (accumulate) o = cadd(o, out[r]);
        out[r] = o;
    }
}

// The same product over a coded SELL-C operator (value and column-offset
// tables, 16-bit codes; sell_coded_row_dot).  Every workgroup (256 threads)
// stages the two tables in LDS first.
template <int L>
__global__ __launch_bounds__(256) void sell_coded_matvec_kernel(
    int nrows, int ncols, int nslices, const int *__restrict__ slice_ptr,
    const uint2 *__restrict__ code, const double2 *__restrict__ vtab, const int *__restrict__ otab,
    int vbits, const double2 *__restrict__ x, double2 *__restrict__ out, double2 scale,
    int accumulate) {
    __shared__ double2 vt[kCodeTabMax];
    __shared__ int ot[kCodeTabMax];
    code_tables_to_lds(vtab, otab, vbits, vt, ot);
    __syncthreads();
    const CodeSplit cs = {(1u << vbits) - 1u, vbits};
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int slice = (int)(t / QA_WAVE), lane = (int)(t % QA_WAVE);
    if (slice >= nslices) return;
    int base, words;
    sell_slice(slice_ptr, slice, base, words);
    const int64_t r = (int64_t)slice * (QA_WAVE / L) + lane / L;
    const double2 acc =
        sell_coded_row_dot<L>(code, vt, ot, cs, x, base, words, lane, (int)r, ncols - 1);
    if (lane % L == 0 && r < nrows) {
        double2 o = cmul(scale, acc);
        if (accumulate) o = cadd(o, out[r]);
        out[r] = o;
    }
}

// out = base + sum_j coef[j] * vecs[j]   (real coefficients, up to 16 terms)
struct LinComb {
    const double2 *v[16];
    double c[16];
};

__global__ __launch_bounds__(256) void lincomb_kernel(int64_t n, const double2 *__restrict__ base,
                                                      int nterms, LinComb lc,
                                                      double2 *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double2 o = base ? base[i] : make_double2(0.0, 0.0);
        for (int j = 0; j < nterms; ++j) raxpy(o, lc.c[j], lc.v[j][i]);
        out[i] = o;
    }
}

// out = alpha * x + beta * y  (complex scalars; y may be null)
__global__ __launch_bounds__(256) void axpby_kernel(int64_t n, double2 alpha,
                                                    const double2 *__restrict__ x, double2 beta,
                                                    const double2 *__restrict__ y,
                                                    double2 *__restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double2 o = cmul(alpha, x[i]);
        if (y) o = cadd(o, cmul(beta, y[i]));
        out[i] = o;
    }
}

// ---------------------------------------------------------------------------
// Reductions: per-workgroup partial sums, then a one-workgroup kernel adds
// them in a fixed order (deterministic) and writes res[0..1].
// kind 0: wrmn sum (|d|/(atol + rtol |s|))^2 -> sqrt(sum / n)
// kind 1: sum |x|^2
// kind 2: sum conj(x) y (complex)
// kind 3: sum over i of x[i * (d + 1)] (trace of a column-major d x d)
// ---------------------------------------------------------------------------
struct RedScratch {
    double *partials;  // 2 * maxblocks
    double *res;       // 2
};

__global__ __launch_bounds__(256) void reduce_kernel(int kind, int64_t n, const double2 *__restrict__ x,
                                                     const double2 *__restrict__ y, double atol,
                                                     double rtol, int64_t stride, RedScratch S) {
    __shared__ double smem[8];
    double sre = 0.0, sim = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (kind == 0) {
            const double ad = sqrt(cabs2(x[i])), as = sqrt(cabs2(y[i]));
            const double q = ad / (atol + rtol * as);
            sre += q * q;
        } else if (kind == 1) {
            sre += cabs2(x[i]);
        } else if (kind == 2) {
            const double2 a = x[i], b = y[i];
            sre += a.x * b.x + a.y * b.y;
            sim += a.x * b.y - a.y * b.x;
        } else {
            const double2 a = x[i * stride];
            sre += a.x;
            sim += a.y;
        }
    }
    sre = block_sum(sre, smem);
    sim = block_sum(sim, smem);
    if (threadIdx.x == 0) {
        S.partials[2 * blockIdx.x] = sre;
        S.partials[2 * blockIdx.x + 1] = sim;
    }
}

// Adds the per-workgroup partial sums in a fixed order (one workgroup; the
// launch boundary orders it after the partial sums).  mean_sqrt: the result
// is sqrt(sum / n) (the weighted RMS norm).
__global__ __launch_bounds__(256) void reduce_finish_kernel(int nblocks, int mean_sqrt,
                                                            int64_t n, RedScratch S) {
    __shared__ double smem[8];
    double tre = 0.0, tim = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        tre += S.partials[2 * b];
        tim += S.partials[2 * b + 1];
    }
    tre = block_sum(tre, smem);
    tim = block_sum(tim, smem);
    if (threadIdx.x == 0) {
        S.res[0] = mean_sqrt ? sqrt(tre / (double)n) : tre;
        S.res[1] = tim;
    }
}

// <psi|O|psi> for a ket (kind 0) and tr(O rho) for a column-major d x d rho
// (kind 1): one thread per row of O.
template <typename V>
__global__ __launch_bounds__(256) void expect_csr_kernel(int kind, int nrows,
                                                         const int *__restrict__ rowptr,
                                                         const int *__restrict__ col,
                                                         const V *__restrict__ vals,
                                                         const double2 *__restrict__ st,
                                                         RedScratch S) {
    __shared__ double smem[8];
    double sre = 0.0, sim = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows;
         r += (int64_t)gridDim.x * blockDim.x) {
        double2 acc = make_double2(0.0, 0.0);
        if (kind == 0) {
            for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) cfma(acc, ld(vals, p), st[col[p]]);
            const double2 c = st[r];
            sre += c.x * acc.x + c.y * acc.y;
            sim += c.x * acc.y - c.y * acc.x;
        } else {
            // sum_c O[r, c] rho[c, r]; rho[c, r] lives at r * d + c
            for (int p = rowptr[r]; p < rowptr[r + 1]; ++p)
                cfma(acc, ld(vals, p), st[r * (int64_t)nrows + col[p]]);
            sre += acc.x;
            sim += acc.y;
        }
    }
    sre = block_sum(sre, smem);
    sim = block_sum(sim, smem);
    if (threadIdx.x == 0) {
        S.partials[2 * blockIdx.x] = sre;
        S.partials[2 * blockIdx.x + 1] = sim;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {
constexpr int kMaxRedBlocks = 512;
}  // namespace

// Scratch of the allocation-free error norm (qa_wrmn_error_z_noalloc): the
// partial sums of at most kMaxRedBlocks workgroups, then the two result
// words; 8208 bytes per device, held until the process ends.  Host code
// never passes this name to a kernel: it stands for a host-side shadow, not
// the device address, which scratch_address() looks up per device.
static __device__ __align__(256) double scratch[2 * kMaxRedBlocks + 2];

namespace {
std::mutex g_red_mutex;
struct RedState {
    int device = -1;
    void *buf = nullptr;
} g_red[16];

int red_scratch(RedScratch *S) {
    int dev = 0;
    QA_HIP(hipGetDevice(&dev));
    QA_REQUIRE(dev >= 0 && dev < 16, "device index %d out of range", dev);
    RedState &st = g_red[dev];
    if (!st.buf) {
        QA_HIP(hipMalloc(&st.buf, (2 * kMaxRedBlocks + 2) * sizeof(double) + 256));
        QA_HIP(hipMemset(st.buf, 0, (2 * kMaxRedBlocks + 2) * sizeof(double) + 256));
        st.device = dev;
    }
    char *b = (char *)st.buf;
    S->partials = (double *)b;
    S->res = (double *)(b + 2 * kMaxRedBlocks * sizeof(double));
    return QA_OK;
}

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array at its own address).
std::mutex g_scratch_mutex;
double *g_scratch_addr[16] = {};

int scratch_address(double **out) {
    int dev = 0;
    QA_HIP(hipGetDevice(&dev));
    QA_REQUIRE(dev >= 0 && dev < 16, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        QA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        QA_REQUIRE(p != nullptr, "scratch: no device address on device %d", dev);
        g_scratch_addr[dev] = (double *)p;
    }
    *out = g_scratch_addr[dev];
    return QA_OK;
}

unsigned red_blocks(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > kMaxRedBlocks) b = kMaxRedBlocks;
    return (unsigned)b;
}

int finish_reduction(RedScratch &S, unsigned nblocks, int mean_sqrt, int64_t n, double *result,
                     int ncomp, hipStream_t s) {
    QA_CHECK_LAUNCH();
    hipLaunchKernelGGL(reduce_finish_kernel, dim3(1), dim3(256), 0, s, (int)nblocks, mean_sqrt, n,
                       S);
    QA_CHECK_LAUNCH();
    QA_HIP(hipMemcpyAsync(result, S.res, ncomp * sizeof(double), hipMemcpyDeviceToHost, s));
    QA_HIP(hipStreamSynchronize(s));
    return QA_OK;
}

template <typename V>
int launch_csr_matvec(int nrows, const int *rowptr, const int *col, const V *vals,
                      const double2 *x, double2 *out, double2 scale, int accumulate, int G,
                      hipStream_t s) {
    const int64_t threads = (int64_t)nrows * G;
    const unsigned blocks = (unsigned)((threads + 255) / 256);
    if (blocks == 0) return QA_OK;
#define QA_CSR_CASE(g)                                                                   \
    case g:                                                                              \
        hipLaunchKernelGGL((csr_matvec_kernel<V, g>), dim3(blocks), dim3(256), 0, s, nrows, \
                           rowptr, col, vals, x, out, scale, accumulate);                \
        break;
    switch (G) {
        QA_CSR_CASE(1)
        QA_CSR_CASE(2)
        QA_CSR_CASE(4)
        QA_CSR_CASE(8)
        QA_CSR_CASE(16)
        QA_CSR_CASE(32)
        QA_CSR_CASE(64)
        default:
            qa::set_error("csr_matvec: lanes per row must be a power of two <= 64, got %d", G);
            return QA_ERR_ARG;
    }
#undef QA_CSR_CASE
    QA_CHECK_LAUNCH();
    return QA_OK;
}

int lanes_for(int nrows, int nnz) {
    const double avg = nrows > 0 ? (double)nnz / nrows : 0.0;
    int G = 1;
    while (G < 64 && G < avg) G <<= 1;
    return G;
}
}  // namespace

template <typename V>
static int launch_sell(int lanes, unsigned blocks, hipStream_t s, int nrows, int nslices,
                       const int *slice_ptr, const int *col, const V *vals, const double2 *x,
                       double2 *out, double2 sc, int accumulate) {
    switch (lanes) {
        case 1:
            hipLaunchKernelGGL((sell_matvec_kernel<V, 1>), dim3(blocks), dim3(256), 0, s, nrows,
                               nslices, slice_ptr, col, vals, x, out, sc, accumulate);
            break;
        case 2:
            hipLaunchKernelGGL((sell_matvec_kernel<V, 2>), dim3(blocks), dim3(256), 0, s, nrows,
                               nslices, slice_ptr, col, vals, x, out, sc, accumulate);
            break;
        case 4:
            hipLaunchKernelGGL((sell_matvec_kernel<V, 4>), dim3(blocks), dim3(256), 0, s, nrows,
                               nslices, slice_ptr, col, vals, x, out, sc, accumulate);
            break;
        case 8:
            hipLaunchKernelGGL((sell_matvec_kernel<V, 8>), dim3(blocks), dim3(256), 0, s, nrows,
                               nslices, slice_ptr, col, vals, x, out, sc, accumulate);
            break;
        default:
            qa::set_error("SELL lanes per row must be 1, 2, 4 or 8 (got %d)", lanes);
            return QA_ERR_ARG;
    }
    QA_CHECK_LAUNCH();
    return QA_OK;
}

template <typename V>
int launch_sell_matmat(int lanes, dim3 grid, hipStream_t s, int nrows, int nslices,
                       const int *slice_ptr, const int *col, const V *vals, int nvec,
                       const double2 *x, int64_t ldx, double2 *out, int64_t ldo, double2 sc,
                       int accumulate) {
    switch (lanes) {
        case 1:
            hipLaunchKernelGGL((sell_matmat_kernel<V, 1>), grid, dim3(256), 0, s, nrows, nslices,
                               slice_ptr, col, vals, nvec, x, ldx, out, ldo, sc, accumulate);
            break;
        case 2:
            hipLaunchKernelGGL((sell_matmat_kernel<V, 2>), grid, dim3(256), 0, s, nrows, nslices,
                               slice_ptr, col, vals, nvec, x, ldx, out, ldo, sc, accumulate);
            break;
        case 4:
            hipLaunchKernelGGL((sell_matmat_kernel<V, 4>), grid, dim3(256), 0, s, nrows, nslices,
                               slice_ptr, col, vals, nvec, x, ldx, out, ldo, sc, accumulate);
            break;
        default:
            hipLaunchKernelGGL((sell_matmat_kernel<V, 8>), grid, dim3(256), 0, s, nrows, nslices,
                               slice_ptr, col, vals, nvec, x, ldx, out, ldo, sc, accumulate);
            break;
    }
    QA_CHECK_LAUNCH();
    return QA_OK;
}

extern "C" {

int qa_csr_matvec_z(int nrows, int ncols, const int *rowptr, const int *col, const void *vals,
                    int vals_fp32, int nnz, const void *x, void *out, double scale_re,
                    double scale_im, int accumulate, void *stream) {
    QA_REQUIRE(nrows >= 0 && ncols >= 0, "qa_csr_matvec_z: negative shape");
    QA_REQUIRE(nrows == 0 || (rowptr && out), "qa_csr_matvec_z: null pointer");
    if (nrows == 0) return QA_OK;
    const int G = lanes_for(nrows, nnz);
    const double2 sc = make_double2(scale_re, scale_im);
    hipStream_t s = (hipStream_t)stream;
    if (vals_fp32)
        return launch_csr_matvec(nrows, rowptr, col, (const float2 *)vals, (const double2 *)x,
                                 (double2 *)out, sc, accumulate, G, s);
    return launch_csr_matvec(nrows, rowptr, col, (const double2 *)vals, (const double2 *)x,
                             (double2 *)out, sc, accumulate, G, s);
}

int qa_csr_matmat_z(int nrows, int ncols, const int *rowptr, const int *col, const void *vals,
                    int vals_fp32, int nvec, const void *x, int64_t ldx, void *out, int64_t ldo,
                    double scale_re, double scale_im, int accumulate, void *stream) {
    QA_REQUIRE(nrows >= 0 && ncols >= 0 && nvec >= 0, "qa_csr_matmat_z: negative shape");
    QA_REQUIRE(ldx >= ncols && ldo >= nrows, "qa_csr_matmat_z: leading dimension too small");
    QA_REQUIRE(nvec <= 65535, "qa_csr_matmat_z: at most 65535 columns per call");
    if (nrows == 0 || nvec == 0) return QA_OK;
    dim3 grid((unsigned)((nrows + 255) / 256), (unsigned)((nvec + kMatCols - 1) / kMatCols));
    const double2 sc = make_double2(scale_re, scale_im);
    hipStream_t s = (hipStream_t)stream;
    if (vals_fp32)
        hipLaunchKernelGGL(csr_matmat_kernel<float2>, grid, dim3(256), 0, s, nrows, rowptr, col,
                           (const float2 *)vals, nvec, (const double2 *)x, ldx, (double2 *)out,
                           ldo, sc, accumulate);
    else
        hipLaunchKernelGGL(csr_matmat_kernel<double2>, grid, dim3(256), 0, s, nrows, rowptr, col,
                           (const double2 *)vals, nvec, (const double2 *)x, ldx, (double2 *)out,
                           ldo, sc, accumulate);
    QA_CHECK_LAUNCH();
    return QA_OK;
}

int qa_sell_coded_matvec_z(int nrows, int ncols, int lanes, int nslices, const int *slice_ptr,
                           const void *codes, const void *vtab, const int *otab, int vbits,
                           const void *x, void *out, double scale_re, double scale_im,
                           int accumulate, void *stream) {
    QA_REQUIRE(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8,
               "qa_sell_coded_matvec_z: lanes must be 1, 2, 4 or 8");
    const int C = QA_WAVE / lanes;
    QA_REQUIRE(nrows >= 0 && nslices == (nrows + C - 1) / C,
               "qa_sell_coded_matvec_z: nslices must be ceil(nrows / (64 / lanes))");
    QA_REQUIRE(ncols > 0 || nrows == 0, "qa_sell_coded_matvec_z: ncols must be positive");
    QA_REQUIRE(vbits >= 6 && vbits <= 10, "qa_sell_coded_matvec_z: vbits must be 6..10");
    if (nrows == 0) return QA_OK;
    QA_REQUIRE(slice_ptr && codes && vtab && otab && x && out,
               "qa_sell_coded_matvec_z: null pointer");
    const unsigned blocks = (unsigned)(((int64_t)nslices * QA_WAVE + 255) / 256);
    const double2 sc = make_double2(scale_re, scale_im);
    hipStream_t s = (hipStream_t)stream;
    const uint2 *cw = (const uint2 *)codes;
    const double2 *vt = (const double2 *)vtab;
    const double2 *xx = (const double2 *)x;
    double2 *o = (double2 *)out;
    switch (lanes) {
        case 1:
            hipLaunchKernelGGL(sell_coded_matvec_kernel<1>, dim3(blocks), dim3(256), 0, s, nrows,
                               ncols, nslices, slice_ptr, cw, vt, otab, vbits, xx, o, sc, accumulate);
            break;
        case 2:
            hipLaunchKernelGGL(sell_coded_matvec_kernel<2>, dim3(blocks), dim3(256), 0, s, nrows,
                               ncols, nslices, slice_ptr, cw, vt, otab, vbits, xx, o, sc, accumulate);
            break;
        case 4:
            hipLaunchKernelGGL(sell_coded_matvec_kernel<4>, dim3(blocks), dim3(256), 0, s, nrows,
                               ncols, nslices, slice_ptr, cw, vt, otab, vbits, xx, o, sc, accumulate);
            break;
        default:
            hipLaunchKernelGGL(sell_coded_matvec_kernel<8>, dim3(blocks), dim3(256), 0, s, nrows,
                               ncols, nslices, slice_ptr, cw, vt, otab, vbits, xx, o, sc, accumulate);
            break;
    }
    QA_CHECK_LAUNCH();
    return QA_OK;
}

int qa_sell_matvec_z(int nrows, int lanes, int nslices, const int *slice_ptr, const int *col,
                     const void *vals, int vals_fp32, const void *x, void *out, double scale_re,
                     double scale_im, int accumulate, void *stream) {
    QA_REQUIRE(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8,
               "qa_sell_matvec_z: lanes must be 1, 2, 4 or 8");
    const int C = QA_WAVE / lanes;
    QA_REQUIRE(nrows >= 0 && nslices == (nrows + C - 1) / C,
               "qa_sell_matvec_z: nslices must be ceil(nrows / (64 / lanes))");
    if (nrows == 0) return QA_OK;
    const unsigned blocks = (unsigned)(((int64_t)nslices * QA_WAVE + 255) / 256);
    const double2 sc = make_double2(scale_re, scale_im);
    hipStream_t s = (hipStream_t)stream;
    if (vals_fp32)
        return launch_sell(lanes, blocks, s, nrows, nslices, slice_ptr, col, (const float2 *)vals,
                           (const double2 *)x, (double2 *)out, sc, accumulate);
    return launch_sell(lanes, blocks, s, nrows, nslices, slice_ptr, col, (const double2 *)vals,
                       (const double2 *)x, (double2 *)out, sc, accumulate);
}

int qa_sell_matmat_z(int nrows, int ncols, int lanes, int nslices, const int *slice_ptr,
                     const int *col, const void *vals, int vals_fp32, int nvec, const void *x,
                     int64_t ldx, void *out, int64_t ldo, double scale_re, double scale_im,
                     int accumulate, void *stream) {
    QA_REQUIRE(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8,
               "qa_sell_matmat_z: lanes must be 1, 2, 4 or 8");
    const int C = QA_WAVE / lanes;
    QA_REQUIRE(nrows >= 0 && ncols >= 0 && nvec >= 0, "qa_sell_matmat_z: negative shape");
    QA_REQUIRE(nslices == (nrows + C - 1) / C,
               "qa_sell_matmat_z: nslices must be ceil(nrows / (64 / lanes))");
    QA_REQUIRE(ldx >= ncols && ldo >= nrows, "qa_sell_matmat_z: leading dimension too small");
    QA_REQUIRE(nvec <= 65535, "qa_sell_matmat_z: at most 65535 columns per call");
    if (nrows == 0 || nvec == 0) return QA_OK;
    const dim3 grid((unsigned)(((int64_t)nslices * QA_WAVE + 255) / 256),
                    (unsigned)((nvec + kMatCols - 1) / kMatCols));
    const double2 sc = make_double2(scale_re, scale_im);
    hipStream_t s = (hipStream_t)stream;
    if (vals_fp32)
        return launch_sell_matmat(lanes, grid, s, nrows, nslices, slice_ptr, col,
                                  (const float2 *)vals, nvec, (const double2 *)x, ldx,
                                  (double2 *)out, ldo, sc, accumulate);
    return launch_sell_matmat(lanes, grid, s, nrows, nslices, slice_ptr, col,
                              (const double2 *)vals, nvec, (const double2 *)x, ldx, (double2 *)out,
                              ldo, sc, accumulate);
}

int qa_lincomb_z(int64_t n, const void *base, int nterms, const void *const *vecs,
                 const double *coefs, void *out, void *stream) {
    QA_REQUIRE(n >= 0 && nterms >= 0 && nterms <= 16, "qa_lincomb_z: 0 <= nterms <= 16");
    if (n == 0) return QA_OK;
    LinComb lc;
    for (int j = 0; j < 16; ++j) {
        lc.v[j] = j < nterms ? (const double2 *)vecs[j] : nullptr;
        lc.c[j] = j < nterms ? coefs[j] : 0.0;
    }
    unsigned blocks = (unsigned)((n + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(lincomb_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, n,
                       (const double2 *)base, nterms, lc, (double2 *)out);
    QA_CHECK_LAUNCH();
    return QA_OK;
}

int qa_axpby_z(int64_t n, double are, double aim, const void *x, double bre, double bim,
               const void *y, void *out, void *stream) {
    QA_REQUIRE(n >= 0 && (n == 0 || (x && out)), "qa_axpby_z: null pointer");
    if (n == 0) return QA_OK;
    unsigned blocks = (unsigned)((n + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(axpby_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, n,
                       make_double2(are, aim), (const double2 *)x, make_double2(bre, bim),
                       (const double2 *)y, (double2 *)out);
    QA_CHECK_LAUNCH();
    return QA_OK;
}

static int run_reduce(int kind, int64_t n, const void *x, const void *y, double atol, double rtol,
                      int64_t stride, double *result, int ncomp, void *stream) {
    std::lock_guard<std::mutex> lock(g_red_mutex);
    RedScratch S;
    int rc = red_scratch(&S);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const unsigned nb = red_blocks(n);
    hipLaunchKernelGGL(reduce_kernel, dim3(nb), dim3(256), 0, s, kind, n, (const double2 *)x,
                       (const double2 *)y, atol, rtol, stride, S);
    return finish_reduction(S, nb, kind == 0, n, result, ncomp, s);
}

int qa_wrmn_error_z(int64_t n, const void *diff, const void *state, double atol, double rtol,
                    double *result, void *stream) {
    QA_REQUIRE(n > 0 && diff && state && result, "qa_wrmn_error_z: bad arguments");
    return run_reduce(0, n, diff, state, atol, rtol, 1, result, 1, stream);
}

int qa_wrmn_error_z_noalloc(int64_t n, const void *diff, const void *state, double atol,
                            double rtol, double *result, void *stream) {
    QA_REQUIRE(n > 0 && diff && state && result, "qa_wrmn_error_z_noalloc: bad arguments");
    // the grid never exceeds kMaxRedBlocks workgroups, so the partial sums of
    // any n fit the fixed scratch; the kernels and the grid are those of
    // qa_wrmn_error_z, hence the same result bits
    const unsigned nb = red_blocks(n);
    QA_REQUIRE(nb <= (unsigned)kMaxRedBlocks,
               "qa_wrmn_error_z_noalloc: %u partial sums exceed the scratch of %d", nb,
               kMaxRedBlocks);
    double *p = nullptr;
    QA_CHECK(scratch_address(&p));
    RedScratch S;
    S.partials = p;
    S.res = p + 2 * kMaxRedBlocks;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(reduce_kernel, dim3(nb), dim3(256), 0, s, 0, n, (const double2 *)diff,
                       (const double2 *)state, atol, rtol, (int64_t)1, S);
    return finish_reduction(S, nb, 1, n, result, 1, s);
}

int qa_norm2_z(int64_t n, const void *x, double *result, void *stream) {
    QA_REQUIRE(n >= 0 && result, "qa_norm2_z: bad arguments");
    if (n == 0) { *result = 0.0; return QA_OK; }
    return run_reduce(1, n, x, nullptr, 0, 0, 1, result, 1, stream);
}

int qa_inner_z(int64_t n, const void *x, const void *y, double *result2, void *stream) {
    QA_REQUIRE(n >= 0 && result2, "qa_inner_z: bad arguments");
    if (n == 0) { result2[0] = result2[1] = 0.0; return QA_OK; }
    return run_reduce(2, n, x, y, 0, 0, 1, result2, 2, stream);
}

int qa_trace_z(int d, const void *rho, double *result2, void *stream) {
    QA_REQUIRE(d >= 0 && result2, "qa_trace_z: bad arguments");
    if (d == 0) { result2[0] = result2[1] = 0.0; return QA_OK; }
    return run_reduce(3, d, rho, nullptr, 0, 0, (int64_t)d + 1, result2, 2, stream);
}

static int run_expect(int kind, int nrows, const int *rowptr, const int *col, const void *vals,
                      int vals_fp32, const void *st, double *result2, void *stream) {
    std::lock_guard<std::mutex> lock(g_red_mutex);
    RedScratch S;
    int rc = red_scratch(&S);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const unsigned nb = red_blocks(nrows);
    if (vals_fp32)
        hipLaunchKernelGGL(expect_csr_kernel<float2>, dim3(nb), dim3(256), 0, s, kind, nrows,
                           rowptr, col, (const float2 *)vals, (const double2 *)st, S);
    else
        hipLaunchKernelGGL(expect_csr_kernel<double2>, dim3(nb), dim3(256), 0, s, kind, nrows,
                           rowptr, col, (const double2 *)vals, (const double2 *)st, S);
    return finish_reduction(S, nb, 0, nrows, result2, 2, s);
}

int qa_expect_csr_ket_z(int n, const int *rowptr, const int *col, const void *vals,
                        int vals_fp32, const void *psi, double *result2, void *stream) {
    QA_REQUIRE(n > 0 && rowptr && psi && result2, "qa_expect_csr_ket_z: bad arguments");
    return run_expect(0, n, rowptr, col, vals, vals_fp32, psi, result2, stream);
}

int qa_expect_csr_dm_z(int d, const int *rowptr, const int *col, const void *vals, int vals_fp32,
                       const void *rho, double *result2, void *stream) {
    QA_REQUIRE(d > 0 && rowptr && rho && result2, "qa_expect_csr_dm_z: bad arguments");
    return run_expect(1, d, rowptr, col, vals, vals_fp32, rho, result2, stream);
}

}  // extern "C"
