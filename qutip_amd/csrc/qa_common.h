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
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
    return make_double2(a.x - b.x, a.y - b.y);
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

// A double moved across lanes by the DPP pattern C (both 32-bit halves).
template <int C>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), C, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), C, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_value(double v, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Sum over the 64 lanes of a wavefront (every lane active), the result in
// every lane.  Within each row of 16 lanes, DPP moves (lane xor 1, xor 2,
// the half-row mirror, the row mirror) leave the row's sum in all of its
// lanes (each addition is commutative, so every lane holds the same bits);
// the four row sums are then read from lanes 0, 16, 32 and 48 and added,
// (r0 + r1) + (r2 + r3).  A fixed order, so deterministic; register moves
// instead of the six LDS-crossbar exchanges per 32-bit half of a shuffle
// butterfly.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move<0xB1>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_move<0x4E>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_move<0x141>(v);  // row_half_mirror
    v += dpp_move<0x140>(v);  // row_mirror
    return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
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

constexpr int kRowChunk = 24;  // entries of a lane per chunk (sell_row_dot)
static_assert(kRowChunk % 2 == 0, "chunks must start at even steps");

template <int L, typename V>
__device__ __forceinline__ double2 sell_row_dot(const int *__restrict__ col,
                                                const V *__restrict__ vals,
                                                const double2 *__restrict__ x, int base,
                                                int steps, int lane) {
    // Every lane of a slice makes the same number of steps (the slice width
    // divided by L), so the code below is not divergent.  A lane's entries
    // go in chunks of up to kRowChunk; the last chunk (a wave-uniform count)
    // is dispatched once to a fully unrolled body of exactly that length, so
    // no load is predicated and none waits for another.  A row that fits one
    // chunk costs two dependent round trips (columns and values, then the
    // gathers), whatever its length.  Chunks start at even steps, so step
    // parity (a0 / a1) is the same for every chunk size.
    const int p0 = base + lane;
    double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
    int k = 0;
    for (; k + kRowChunk < steps; k += kRowChunk)
        sell_chunk<kRowChunk>(col, vals, x, p0 + k * QA_WAVE, a0, a1);
    const int p = p0 + k * QA_WAVE;
    switch (steps - k) {
#define QA_SELL_CASE(m) \
        case m: sell_chunk<m>(col, vals, x, p, a0, a1); break;
        QA_SELL_CASE(24) QA_SELL_CASE(23) QA_SELL_CASE(22) QA_SELL_CASE(21) QA_SELL_CASE(20)
        QA_SELL_CASE(19) QA_SELL_CASE(18) QA_SELL_CASE(17) QA_SELL_CASE(16) QA_SELL_CASE(15)
        QA_SELL_CASE(14) QA_SELL_CASE(13) QA_SELL_CASE(12) QA_SELL_CASE(11) QA_SELL_CASE(10)
        QA_SELL_CASE(9) QA_SELL_CASE(8) QA_SELL_CASE(7) QA_SELL_CASE(6) QA_SELL_CASE(5)
        QA_SELL_CASE(4) QA_SELL_CASE(3) QA_SELL_CASE(2) QA_SELL_CASE(1)
#undef QA_SELL_CASE
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

// Coded SELL-C: the slices and the lane -> entry assignment of sell_row_dot,
// with every stored entry replaced by a 16-bit code.  The low `vbits` bits
// (6..10) index a table of 2^vbits complex values, the high 16 - vbits bits
// a table of 2^(16 - vbits) column offsets (column = row + offset, clamped
// to [0, cmax], so no code can send a gather out of range).  Operators
// assembled from tensor products hold few distinct values and few distinct
// diagonals, so the 20 bytes of a value and a column become 2 (index and
// value compression in the manner of CSR-DU / CSR-VI: Kourtis, Goumas and
// Koziris, 2008).  Four consecutive entries of a lane (its steps 4w ..
// 4w + 3) are packed into one 8-byte word, word slice_ptr[s] + 64 w + lane
// of the slice, so each load of the wavefront reads 512 consecutive bytes;
// slice widths are multiples of 4 L and slice_ptr counts words.  The
// products are the plain path's terms summed in the plain path's order (even
// steps into a0, odd into a1, then the lane shuffles), so both forms give
// the same bits.  `vt`, `ot` point to the tables in LDS.
constexpr int kCodeTabMax = 1024;  // entries of either table at most (vbits 6..10)
static_assert(kRowChunk == 24, "the coded dispatch below has cases 1..6 words");

struct CodeSplit {
    unsigned vmask;  // 2^vbits - 1
    int vbits;
};

template <int M>
__device__ __forceinline__ void sell_coded_chunk(const uint2 *__restrict__ code, const double2 *vt,
                                                 const int *ot, CodeSplit cs,
                                                 const double2 *__restrict__ x, int w, int row,
                                                 int cmax, double2 &a0, double2 &a1) {
    uint2 cw[M];
#pragma unroll
    for (int u = 0; u < M; ++u) cw[u] = code[w + u * QA_WAVE];
    unsigned h[4 * M];
#pragma unroll
    for (int u = 0; u < M; ++u) {
        h[4 * u + 0] = cw[u].x & 0xffffu;
        h[4 * u + 1] = cw[u].x >> 16;
        h[4 * u + 2] = cw[u].y & 0xffffu;
        h[4 * u + 3] = cw[u].y >> 16;
    }
    double2 xv[4 * M];
#pragma unroll
    for (int e = 0; e < 4 * M; ++e) xv[e] = x[min(max(row + ot[h[e] >> cs.vbits], 0), cmax)];
#pragma unroll
    for (int e = 0; e < 4 * M; ++e) cfma((e & 1) ? a1 : a0, vt[h[e] & cs.vmask], xv[e]);
}

template <int L>
__device__ __forceinline__ double2 sell_coded_row_dot(const uint2 *__restrict__ code,
                                                      const double2 *vt, const int *ot,
                                                      CodeSplit cs, const double2 *__restrict__ x,
                                                      int base, int words, int lane, int row,
                                                      int cmax) {
    // Words go in chunks of up to kRowChunk / 4 (the entries of a chunk of
    // sell_row_dot), the last one dispatched to a body of exactly its
    // length; the word count is uniform over the slice.
    constexpr int MW = kRowChunk / 4;
    const int w0 = base + lane;
    double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
    int k = 0;
    for (; k + MW < words; k += MW)
        sell_coded_chunk<MW>(code, vt, ot, cs, x, w0 + k * QA_WAVE, row, cmax, a0, a1);
    const int w = w0 + k * QA_WAVE;
    switch (words - k) {
        case 6: sell_coded_chunk<6>(code, vt, ot, cs, x, w, row, cmax, a0, a1); break;
        case 5: sell_coded_chunk<5>(code, vt, ot, cs, x, w, row, cmax, a0, a1); break;
        case 4: sell_coded_chunk<4>(code, vt, ot, cs, x, w, row, cmax, a0, a1); break;
        case 3: sell_coded_chunk<3>(code, vt, ot, cs, x, w, row, cmax, a0, a1); break;
        case 2: sell_coded_chunk<2>(code, vt, ot, cs, x, w, row, cmax, a0, a1); break;
        case 1: sell_coded_chunk<1>(code, vt, ot, cs, x, w, row, cmax, a0, a1); break;
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

// Coded rows of uniform width (every slice `uwords` words per lane, word w
// of slice s at s uwords 64 + 64 w + lane): the words do not depend on a
// slice pointer, so a kernel can request them with its first loads.  `cw`
// holds the lane's words (kUWordsMax, the ones past uwords unused); the
// terms are summed as by sell_coded_row_dot (the same bits).
constexpr int kUWordsMax = kRowChunk / 4;

template <int M>
__device__ __forceinline__ void sell_coded_words(const uint2 (&cw)[kUWordsMax], const double2 *vt,
                                                 const int *ot, CodeSplit cs,
                                                 const double2 *__restrict__ x, int row, int cmax,
                                                 double2 &a0, double2 &a1) {
    unsigned h[4 * M];
#pragma unroll
    for (int u = 0; u < M; ++u) {
        h[4 * u + 0] = cw[u].x & 0xffffu;
        h[4 * u + 1] = cw[u].x >> 16;
        h[4 * u + 2] = cw[u].y & 0xffffu;
        h[4 * u + 3] = cw[u].y >> 16;
    }
    double2 xv[4 * M];
#pragma unroll
    for (int e = 0; e < 4 * M; ++e) xv[e] = x[min(max(row + ot[h[e] >> cs.vbits], 0), cmax)];
#pragma unroll
    for (int e = 0; e < 4 * M; ++e) cfma((e & 1) ? a1 : a0, vt[h[e] & cs.vmask], xv[e]);
}

template <int L>
__device__ __forceinline__ double2 sell_coded_row_dot_uniform(const uint2 (&cw)[kUWordsMax],
                                                              int uwords, const double2 *vt,
                                                              const int *ot, CodeSplit cs,
                                                              const double2 *__restrict__ x,
                                                              int row, int cmax) {
    double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
    switch (uwords) {
        case 6: sell_coded_words<6>(cw, vt, ot, cs, x, row, cmax, a0, a1); break;
        case 5: sell_coded_words<5>(cw, vt, ot, cs, x, row, cmax, a0, a1); break;
        case 4: sell_coded_words<4>(cw, vt, ot, cs, x, row, cmax, a0, a1); break;
        case 3: sell_coded_words<3>(cw, vt, ot, cs, x, row, cmax, a0, a1); break;
        case 2: sell_coded_words<2>(cw, vt, ot, cs, x, row, cmax, a0, a1); break;
        default: sell_coded_words<1>(cw, vt, ot, cs, x, row, cmax, a0, a1); break;
    }
    double2 acc = cadd(a0, a1);
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) {
        acc.x += __shfl_xor(acc.x, off, QA_WAVE);
        acc.y += __shfl_xor(acc.y, off, QA_WAVE);
    }
    return acc;
}

// The tables of a coded operator into LDS (`vt`: kCodeTabMax double2,
// `ot`: kCodeTabMax int) by a workgroup of exactly 256 threads, with direct
// global -> LDS loads: lane l of a wave writes entry i0 + l, i0 the wave's
// uniform base, so no table value passes through registers and every load
// is issued before anything waits.  Indices are clamped to the table sizes
// (2^vbits, 2^(16 - vbits)): all loads are in range; slots past a table's
// size receive copies of its last entry, which no code reaches.  The caller
// synchronises the workgroup (__syncthreads waits for the loads) before
// reading the tables and before any wave of the workgroup exits.
typedef __attribute__((address_space(1))) const void qa_global_cvoid;
typedef __attribute__((address_space(3))) void qa_lds_void;

__device__ __forceinline__ void code_tables_to_lds(const double2 *__restrict__ vtab,
                                                   const int *__restrict__ otab, int vbits,
                                                   double2 *vt, int *ot) {
    const int nv = 1 << vbits, no = 1 << (16 - vbits);
    const int wave = (int)threadIdx.x / QA_WAVE, lane = (int)threadIdx.x % QA_WAVE;
#pragma unroll
    for (int q = 0; q < kCodeTabMax / 256; ++q) {
        const int i0 = 256 * q + QA_WAVE * wave;
        __builtin_amdgcn_global_load_lds((qa_global_cvoid *)(vtab + min(i0 + lane, nv - 1)),
                                         (qa_lds_void *)(vt + i0), 16, 0, 0);
        __builtin_amdgcn_global_load_lds((qa_global_cvoid *)(otab + min(i0 + lane, no - 1)),
                                         (qa_lds_void *)(ot + i0), 4, 0, 0);
    }
}

// Sum of `n` per-workgroup partials by the whole workgroup, in a fixed
// order (strided per thread, then the block tree): deterministic.
__device__ __forceinline__ double sum_partials(const double *partials, int n, double *smem) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += partials[i];
    return block_sum(v, smem);
}

}  // namespace qa
