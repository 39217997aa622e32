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

// "Last workgroup done" handshake: every workgroup publishes its partial
// results, then increments `counter`; the workgroup that observes
// gridDim.x - 1 finishes the reduction (classic threadfence reduction).  The
// device-scope fences write the partials back past the per-XCD L2 and make
// the last workgroup re-read them from memory.  The counter is reset by the
// last workgroup, so the next kernel that uses it starts from zero.
__device__ __forceinline__ bool last_block_done(unsigned int *counter) {
    __shared__ int is_last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int prev = atomicAdd(counter, 1u);
        is_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (is_last) {
        __threadfence();
        if (threadIdx.x == 0) atomicExch(counter, 0u);
    }
    return is_last != 0;
}

// SELL-C row dot product with L lanes per row (C = 64 / L rows per slice;
// one wavefront per slice).  Entry j of row C s + r is stored at
// slice_ptr[s] + 64 (j / L) + L r + (j % L), so each load instruction of the
// wave reads 64 consecutive values and columns; lane l = L r + q handles
// entries q, q + L, ... of row r and the L partial sums are combined with
// cross-lane shuffles.  Two accumulators keep four loads in flight.  The
// result is valid in every lane of the row.
template <int L, typename V>
__device__ __forceinline__ double2 sell_row_dot(const int *__restrict__ slice_ptr,
                                                const int *__restrict__ col,
                                                const V *__restrict__ vals,
                                                const double2 *__restrict__ x, int slice, int lane) {
    // Every lane of a slice makes the same number of steps (the slice width
    // divided by L), so the loop below is not divergent.  Eight entries are
    // loaded before any is used: their column loads, value loads and x
    // gathers overlap instead of forming a chain of dependent memory round
    // trips.
    const int base = slice_ptr[slice];
    const int steps = (slice_ptr[slice + 1] - base) / QA_WAVE;
    const int p0 = base + lane;
    double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
    int k = 0;
    for (; k + 8 <= steps; k += 8) {
        int c[8];
        double2 v[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) c[u] = col[p0 + (k + u) * QA_WAVE];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = ld(vals, p0 + (k + u) * QA_WAVE);
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
            cfma(a0, v[u], xv[u]);
            cfma(a1, v[u + 1], xv[u + 1]);
        }
    }
    if (k + 4 <= steps) {
        int c[4];
        double2 v[4], xv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u] = col[p0 + (k + u) * QA_WAVE];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = ld(vals, p0 + (k + u) * QA_WAVE);
#pragma unroll
        for (int u = 0; u < 4; ++u) xv[u] = x[c[u]];
        cfma(a0, v[0], xv[0]);
        cfma(a1, v[1], xv[1]);
        cfma(a0, v[2], xv[2]);
        cfma(a1, v[3], xv[3]);
        k += 4;
    }
    if (k + 2 <= steps) {
        const int c0 = col[p0 + k * QA_WAVE], c1 = col[p0 + (k + 1) * QA_WAVE];
        const double2 v0 = ld(vals, p0 + k * QA_WAVE), v1 = ld(vals, p0 + (k + 1) * QA_WAVE);
        const double2 x0 = x[c0], x1 = x[c1];
        cfma(a0, v0, x0);
        cfma(a1, v1, x1);
        k += 2;
    }
    if (k < steps) cfma(a0, ld(vals, p0 + k * QA_WAVE), x[col[p0 + k * QA_WAVE]]);
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
