// Shared device helpers for the gfx950 kernels of qutip_amd.
//
// Complex numbers are stored interleaved (re, im) as double2 (complex128) or
// float2 (complex64) -- the layout of numpy complex arrays and of QuTiP's
// Dense/CSR data (qutip/core/data/dense.pxd:9-33, csr.pxd:17-30), so host
// buffers copy to HBM without repacking.  Reductions are done over the
// 64-lane CDNA wavefront with cross-lane shuffles, then across the waves of a
// workgroup through LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define QA_WAVE 64

namespace qa {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
    return make_double2(a.x + b.x, a.y + b.y);
}
// acc += a * b (complex, four fused multiply-adds)
__device__ __forceinline__ void cfma(double2 &acc, double2 a, double2 b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}
// acc += s * v (real scale)
__device__ __forceinline__ void raxpy(double2 &acc, double s, double2 v) {
    acc.x = fma(s, v.x, acc.x);
    acc.y = fma(s, v.y, acc.y);
}
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }

// Value loads: operator values may be stored in complex64 (float2) and are
// widened to FP64 for accumulation.
__device__ __forceinline__ double2 ld(const double2 *p, int64_t i) { return p[i]; }
__device__ __forceinline__ double2 ld(const float2 *p, int64_t i) {
    float2 v = p[i];
    return make_double2((double)v.x, (double)v.y);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = QA_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, QA_WAVE);
    return v;
}

// Sum over the workgroup; every thread gets the result.  `smem` must hold
// blockDim.x / 64 doubles.  Deterministic: fixed tree order.
__device__ __forceinline__ double block_sum(double v, double *smem) {
    const int lane = threadIdx.x & (QA_WAVE - 1);
    const int wid = threadIdx.x / QA_WAVE;
    const int nw = (blockDim.x + QA_WAVE - 1) / QA_WAVE;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) smem[wid] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < nw; ++w) t += smem[w];
    __syncthreads();
    return t;
}

__device__ __forceinline__ double2 block_sum2(double2 v, double *smem) {
    double re = block_sum(v.x, smem);
    double im = block_sum(v.y, smem);
    return make_double2(re, im);
}

// SELL-C row dot product with L lanes per row (C = 64 / L rows per slice;
// one wavefront per slice).  Entry j of row C s + r is stored at
// slice_ptr[s] + 64 (j / L) + L r + (j % L), so each load instruction of the
// wave reads 64 consecutive values and columns; lane l = L r + q handles
// entries q, q + L, ... of row r and the L partial sums are combined with
// cross-lane shuffles.  The result is valid in every lane of the row.
//
// The slice bounds come from sell_slice(), which a kernel can call before
// other loads (the control-block read of the RK stage kernels) so that the
// two round trips overlap.
__device__ __forceinline__ void sell_slice(const int *__restrict__ slice_ptr, int slice,
                                           int &base, int &steps) {
    base = slice_ptr[slice];
    steps = (slice_ptr[slice + 1] - base) / QA_WAVE;
}

// M consecutive entries (steps k .. k + M - 1) of this lane: all column and
// value loads are issued, then all gathers, then the multiply-adds, so the
// chunk costs one column round trip and one gather round trip.
template <int M, typename V>
__device__ __forceinline__ void sell_chunk(const int *__restrict__ col, const V *__restrict__ vals,
                                           const double2 *__restrict__ x, int p, double2 &a0,
                                           double2 &a1) {
    int c[M];
    double2 v[M], xv[M];
#pragma unroll
    for (int u = 0; u < M; ++u) c[u] = col[p + u * QA_WAVE];
#pragma unroll
    for (int u = 0; u < M; ++u) v[u] = ld(vals, p + u * QA_WAVE);
#pragma unroll
    for (int u = 0; u < M; ++u) xv[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < M; ++u) cfma((u & 1) ? a1 : a0, v[u], xv[u]);
}

template <int L, typename V>
__device__ __forceinline__ double2 sell_row_dot(const int *__restrict__ col,
                                                const V *__restrict__ vals,
                                                const double2 *__restrict__ x, int base,
                                                int steps, int lane) {
    // Every lane of a slice makes the same number of steps (the slice width
    // divided by L), so the loop below is not divergent.  Entries go in
    // chunks of eight; the last chunk (1..8 entries, a wave-uniform count)
    // is dispatched once to a fully unrolled body of exactly that length, so
    // no load is predicated and none waits for another.
    const int p0 = base + lane;
    double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
    int k = 0;
    for (; k + 8 < steps; k += 8) sell_chunk<8>(col, vals, x, p0 + k * QA_WAVE, a0, a1);
    const int p = p0 + k * QA_WAVE;
    switch (steps - k) {
        case 8: sell_chunk<8>(col, vals, x, p, a0, a1); break;
        case 7: sell_chunk<7>(col, vals, x, p, a0, a1); break;
        case 6: sell_chunk<6>(col, vals, x, p, a0, a1); break;
        case 5: sell_chunk<5>(col, vals, x, p, a0, a1); break;
        case 4: sell_chunk<4>(col, vals, x, p, a0, a1); break;
        case 3: sell_chunk<3>(col, vals, x, p, a0, a1); break;
        case 2: sell_chunk<2>(col, vals, x, p, a0, a1); break;
        case 1: sell_chunk<1>(col, vals, x, p, a0, a1); break;
        default: break;
    }
    double2 acc = cadd(a0, a1);
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) {
        acc.x += __shfl_xor(acc.x, off, QA_WAVE);
        acc.y += __shfl_xor(acc.y, off, QA_WAVE);
    }
    return acc;
}

// Sum of `n` per-workgroup partials by the whole workgroup, in a fixed
// order (strided per thread, then the block tree): deterministic.
__device__ __forceinline__ double sum_partials(const double *partials, int n, double *smem) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += partials[i];
    return block_sum(v, smem);
}

}  // namespace qa
