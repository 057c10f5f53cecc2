// HIP kernels of the GeMSLR-FGMRES solve path (gfx950, wave64).
//
// Vector blocks are row-interleaved: entry (row i, column c) of an n x nrhs
// block is at [i * nrhs + c]; nrhs is a power of two <= 64 so that a wave
// holds whole rows of columns.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels_trsv.hpp"

namespace pgs {
namespace dev {

constexpr int kWave = 64;

// lane shuffle of a real or complex value
template <typename V>
__device__ __forceinline__ V shfl_xor_v(V v, int m, int width = kWave) { return __shfl_xor(v, m, width); }
template <typename R>
__device__ __forceinline__ cplx<R> shfl_xor_v(cplx<R> v, int m, int width = kWave) {
  return cplx<R>(__shfl_xor(v.re, m, width), __shfl_xor(v.im, m, width));
}

template <typename To, typename From>
__device__ __forceinline__ To cvt(From v) { return (To)v; }

// Streaming read of a basis value (non-temporal: each sweep reads a basis
// vector once, and the basis is far larger than the caches).  At 128^3 x 64
// the delayed-CGS2 dots went from 0.69 to 0.76 of the HBM peak and the update
// from 0.58 to 0.59; the same loads of q and w (reused across the dot chunks)
// and of the preconditioned basis in the solution update gained nothing.
__device__ __forceinline__ double ld_stream(const double* p) { return __builtin_nontemporal_load(p); }
template <typename R>
__device__ __forceinline__ cplx<R> ld_stream(const cplx<R>* p) {
  return cplx<R>(__builtin_nontemporal_load(&p->re), __builtin_nontemporal_load(&p->im));
}

// Row and column of element t of a row-interleaved block of nrhs columns.
// Every entry point admits only power-of-two column counts, so a shift and a
// mask replace the 64-bit division (tens of instructions per element in the
// per-element kernels); other counts take the division.
__device__ __forceinline__ int64_t row_of(int64_t t, int nrhs) {
  return (nrhs & (nrhs - 1)) == 0 ? t >> __builtin_ctz((unsigned)nrhs) : t / nrhs;
}
__device__ __forceinline__ int col_of(int64_t t, int nrhs) {
  return (nrhs & (nrhs - 1)) == 0 ? (int)(t & (nrhs - 1)) : (int)(t % nrhs);
}

// ---------------------------------------------------------------------------
// permutations (IntVectorClass::GatherPerm / ScatterRperm,
// ParGeMSLR/SRC/vectors/int_vector.cpp:629-696) fused with a precision change
// ---------------------------------------------------------------------------
// cscale (may be null): a per-column factor applied to the input first (the
// pending normalization of a Krylov vector)
template <typename VI, typename VO>
__global__ void k_gather(int64_t n, int nrhs, const int* __restrict__ perm, const VI* __restrict__ in, VO* __restrict__ out,
                         const double* __restrict__ cscale) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nrhs) return;
  const int64_t i = row_of(t, nrhs);
  const int c = col_of(t, nrhs);
  VI a = in[(int64_t)perm[i] * nrhs + c];
  if (cscale) a *= VI(cscale[c]);
  out[t] = cvt<VO>(a);
}

// Four consecutive columns of a row per thread (nrhs a multiple of 4, both
// buffers 16-byte aligned): 16-byte and wider accesses instead of one value
// per lane.
template <typename V, int W>
struct alignas(W * sizeof(V) < 16 ? W * sizeof(V) : 16) ColPack {
  V v[W];
};

template <typename VI, typename VO>
__global__ void k_gather4(int64_t n, int nrhs, const int* __restrict__ perm, const VI* __restrict__ in,
                          VO* __restrict__ out, const double* __restrict__ cscale) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t e = 4 * t;
  if (e >= n * nrhs) return;
  const int64_t i = row_of(e, nrhs);
  const int c = col_of(e, nrhs);
  ColPack<VI, 4> a = *reinterpret_cast<const ColPack<VI, 4>*>(in + (int64_t)perm[i] * nrhs + c);
  if (cscale) {
#pragma unroll
    for (int u = 0; u < 4; ++u) a.v[u] *= VI(cscale[c + u]);
  }
  ColPack<VO, 4> b;
#pragma unroll
  for (int u = 0; u < 4; ++u) b.v[u] = cvt<VO>(a.v[u]);
  *reinterpret_cast<ColPack<VO, 4>*>(out + e) = b;
}

template <typename VI, typename VO>
__global__ void k_scatter4(int64_t n, int nrhs, const int* __restrict__ perm, const VI* __restrict__ in,
                           VO* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t e = 4 * t;
  if (e >= n * nrhs) return;
  const int64_t i = row_of(e, nrhs);
  const int c = col_of(e, nrhs);
  const ColPack<VI, 4> a = *reinterpret_cast<const ColPack<VI, 4>*>(in + e);
  ColPack<VO, 4> b;
#pragma unroll
  for (int u = 0; u < 4; ++u) b.v[u] = cvt<VO>(a.v[u]);
  *reinterpret_cast<ColPack<VO, 4>*>(out + (int64_t)perm[i] * nrhs + c) = b;
}

template <typename VI, typename VO>
__global__ void k_scatter(int64_t n, int nrhs, const int* __restrict__ perm, const VI* __restrict__ in, VO* __restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nrhs) return;
  const int64_t i = row_of(t, nrhs);
  const int c = col_of(t, nrhs);
  out[(int64_t)perm[i] * nrhs + c] = cvt<VO>(in[t]);
}


// ---------------------------------------------------------------------------
// CSR SpMV (CsrMatrixPMatVecHostTemplate, ParGeMSLR/SRC/matrices/
// matrixops.cpp:5139-5330).  mode 0: y = alpha*A*x + beta*y (beta == 0 does
// not read y); mode 1: y = r - A*x.  Accumulation in VACC.
// Single RHS: LPR lanes per row, entries strided over the lanes, DPP/shuffle
// reduction.  Multi RHS: one thread per (row, column) over the row's entries.
// ---------------------------------------------------------------------------
// Workgroups are dealt to the 8 XCDs round-robin; the SpMV kernels renumber
// them so that XCD x takes the contiguous range [x * per, (x + 1) * per) of
// logical workgroups (per = gridDim.x / 8, the grid padded to a multiple of
// 8): the rows of one XCD are contiguous, and the x values their neighbours
// read stay in that XCD's L2.
constexpr int kXcds = 8;

__device__ __forceinline__ int64_t xcd_block() {
  const int64_t per = gridDim.x / kXcds;
  return (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

template <typename VA, typename VX, typename VY, typename VACC, int LPR>
__global__ void k_spmv1(int nrows, const int* __restrict__ ia, const int* __restrict__ ja, const VA* __restrict__ a,
                        const VX* __restrict__ x, VY* __restrict__ y, const VY* __restrict__ r, VACC alpha, VACC beta,
                        int mode) {
  int64_t g = xcd_block() * blockDim.x + threadIdx.x;
  int row = (int)(g / LPR);
  int q = (int)(g % LPR);
  VACC acc = 0;
  if (row < nrows) {
    int k1 = ia[row + 1];
    for (int k = ia[row] + q; k < k1; k += LPR) acc += (VACC)a[k] * (VACC)x[ja[k]];
  }
#pragma unroll
  for (int m = LPR / 2; m > 0; m >>= 1) acc += shfl_xor_v(acc, m, LPR);
  if (row < nrows && q == 0) {
    if (mode == 1) y[row] = cvt<VY>((VACC)r[row] - acc);
    else if (beta == (VACC)0) y[row] = cvt<VY>(alpha * acc);
    else y[row] = cvt<VY>(alpha * acc + beta * (VACC)y[row]);
  }
}

template <typename VA, typename VX, typename VY, typename VACC>
__global__ void k_spmvN(int nrows, int nrhs, const int* __restrict__ ia, const int* __restrict__ ja,
                        const VA* __restrict__ a, const VX* __restrict__ x, VY* __restrict__ y,
                        const VY* __restrict__ r, VACC alpha, VACC beta, int mode) {
  int64_t t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nrows * nrhs) return;
  const int row = (int)row_of(t, nrhs);
  const int c = col_of(t, nrhs);
  // the element this thread updates is loaded first, in flight with the row's entries
  const VACC y0 = mode == 1 ? (VACC)r[t] : (beta == (VACC)0 ? VACC(0) : (VACC)y[t]);
  VACC acc = 0;
  const int k1 = ia[row + 1];
  // eight entries at a time: their indices and values, then their x gathers,
  // are all in flight before the first multiply-add (summed in entry order)
  for (int k = ia[row]; k < k1; k += 8) {
    int jj[8];
    VACC av[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const bool in = k + u < k1;
      jj[u] = in ? ja[k + u] : 0;
      av[u] = in ? (VACC)a[k + u] : VACC(0);
    }
    VACC xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] = k + u < k1 ? (VACC)x[(int64_t)jj[u] * nrhs + c] : VACC(0);
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (k + u < k1) acc += av[u] * xv[u];
  }
  if (mode == 1) y[t] = cvt<VY>(y0 - acc);
  else if (beta == (VACC)0) y[t] = cvt<VY>(alpha * acc);
  else y[t] = cvt<VY>(alpha * acc + beta * y0);
}

// CW (4 or 8) consecutive columns per thread (nrhs a multiple of CW, x, y
// and r 16-byte aligned): the row's indices and values are loaded once for
// CW columns and every x gather and y access is 16 bytes or wider; each
// column's sum runs over the entries in the same order as k_spmvN.
template <typename VA, typename VX, typename VY, typename VACC, int CW = 4>
__global__ void k_spmvN4(int nrows, int nrhs, const int* __restrict__ ia, const int* __restrict__ ja,
                         const VA* __restrict__ a, const VX* __restrict__ x, VY* __restrict__ y,
                         const VY* __restrict__ r, VACC alpha, VACC beta, int mode) {
  const int64_t e = CW * (xcd_block() * blockDim.x + threadIdx.x);
  if (e >= (int64_t)nrows * nrhs) return;
  const int row = (int)row_of(e, nrhs);
  const int c = col_of(e, nrhs);
  using PY = ColPack<VY, CW>;
  using PX = ColPack<VX, CW>;
  PY y0;
  if (mode == 1) y0 = *reinterpret_cast<const PY*>(r + e);
  else if (!(beta == (VACC)0)) y0 = *reinterpret_cast<const PY*>(y + e);
  VACC acc[CW];
#pragma unroll
  for (int q = 0; q < CW; ++q) acc[q] = VACC(0);
  const int k1 = ia[row + 1];
  for (int k = ia[row]; k < k1; k += 8) {
    int jj[8];
    VACC av[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const bool in = k + u < k1;
      jj[u] = in ? ja[k + u] : 0;
      av[u] = in ? (VACC)a[k + u] : VACC(0);
    }
    PX xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (k + u < k1) xv[u] = *reinterpret_cast<const PX*>(x + (int64_t)jj[u] * nrhs + c);
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (k + u < k1) {
#pragma unroll
        for (int q = 0; q < CW; ++q) acc[q] += av[u] * (VACC)xv[u].v[q];
      }
  }
  PY out;
#pragma unroll
  for (int q = 0; q < CW; ++q) {
    if (mode == 1) out.v[q] = cvt<VY>((VACC)y0.v[q] - acc[q]);
    else if (beta == (VACC)0) out.v[q] = cvt<VY>(alpha * acc[q]);
    else out.v[q] = cvt<VY>(alpha * acc[q] + beta * (VACC)y0.v[q]);
  }
  *reinterpret_cast<PY*>(y + e) = out;
}

// (block-diagonal ILU solves: kernels_trsv.hpp)

// ---------------------------------------------------------------------------
// Low-rank correction z += WH (W^H z) (GemslrClass::SolveApplyLowRankLevel,
// ParGeMSLR/SRC/preconditioners/gemslr.cpp:3681-3719).  W, WH row-major m x k.
// Stage 1: per-workgroup partial sums of W^H z for every (j, column);
// stage 2: sum of the partials (fixed order -> deterministic);
// stage 3: z[i, c] += sum_j WH[i, j] t[j, c].
// ---------------------------------------------------------------------------
template <typename V, int NT>
__global__ __launch_bounds__(NT) void k_lr_dot(int m, int k, int nrhs, const V* __restrict__ W,
                                               const V* __restrict__ z, V* __restrict__ part, int rows_per_blk) {
  // thread layout: c = tid % nrhs, row lane = tid / nrhs
  __shared__ V sh[NT];
  const int tid = threadIdx.x;
  const int c = tid % nrhs;
  const int lanes = NT / nrhs;
  const int rl = tid / nrhs;
  const int i0 = blockIdx.x * rows_per_blk;
  const int i1 = min(m, i0 + rows_per_blk);
  for (int j = 0; j < k; ++j) {
    V acc = 0;
    for (int i = i0 + rl; i < i1; i += lanes) acc += sconj(W[(int64_t)i * k + j]) * z[(int64_t)i * nrhs + c];
    sh[tid] = acc;
    __syncthreads();
    for (int s = NT / 2; s >= nrhs; s >>= 1) {
      if (tid < s) sh[tid] += sh[tid + s];
      __syncthreads();
    }
    if (tid < nrhs) part[((int64_t)blockIdx.x * k + j) * nrhs + tid] = sh[tid];
    __syncthreads();
  }
}

// Stage 1 with the k partial sums of an element in registers (KM >= k,
// loops unrolled to KM with the uniform guard j < k): every lane walks the
// elements g = (i, c) of z with a grid stride and adds conj(W[i, j]) * z[i, c] to
// sum j; the sums are then reduced over the lanes of a column and over the
// waves into part[blk * k * nrhs + j * nrhs + c].  Dynamic LDS:
// (NT / 64) * k * nrhs values of V.
template <typename V, int NT, int KM>
__global__ __launch_bounds__(NT) void k_lr_dot_reg(int m, int k, int nrhs, const V* __restrict__ W,
                                                   const V* __restrict__ z, V* __restrict__ part) {
  extern __shared__ unsigned char lr_smem[];
  V* red = reinterpret_cast<V*>(lr_smem);
  const int tid = threadIdx.x;
  V acc[KM];
#pragma unroll
  for (int j = 0; j < KM; ++j) acc[j] = 0;
  const int64_t total = (int64_t)m * nrhs;
  for (int64_t g = (int64_t)blockIdx.x * NT + tid; g < total; g += (int64_t)gridDim.x * NT) {
    const V zv = z[g];
    const V* wr = W + row_of(g, nrhs) * k;
#pragma unroll
    for (int j = 0; j < KM; ++j)
      if (j < k) acc[j] += sconj(wr[j]) * zv;
  }
  const int wid = tid / 64, lane = tid % 64;
#pragma unroll
  for (int j = 0; j < KM; ++j) {
    if (j < k) {
      V s = acc[j];
      for (int o = 32; o >= nrhs; o >>= 1) s += shfl_xor_v(s, o);
      if (lane < nrhs) red[(wid * k + j) * nrhs + lane] = s;
    }
  }
  __syncthreads();
  for (int t = tid; t < k * nrhs; t += NT) {
    V s = 0;
    for (int q = 0; q < NT / 64; ++q) s += red[q * k * nrhs + t];
    part[(int64_t)blockIdx.x * k * nrhs + t] = s;
  }
}

template <typename V>
__global__ void k_lr_final(int nblk, int k, int nrhs, const V* __restrict__ part, V* __restrict__ t) {
  // one wave per output, four per workgroup of 256; fixed summation order
  const int g = blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (g >= k * nrhs) return;  // whole wave exits together
  V s = 0;
  for (int b = lane; b < nblk; b += 64) s += part[(int64_t)b * k * nrhs + g];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += shfl_xor_v(s, o);
  if (lane == 0) t[g] = s;
}

template <typename V>
__global__ void k_lr_update(int m, int k, int nrhs, const V* __restrict__ WH, const V* __restrict__ t, V* __restrict__ z) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)m * nrhs) return;
  const int64_t i = row_of(g, nrhs);
  const int c = col_of(g, nrhs);
  V acc = 0;
  for (int j = 0; j < k; ++j) acc += WH[i * k + j] * t[j * nrhs + c];
  z[g] += acc;
}

// ---------------------------------------------------------------------------
// The same low-rank term on the matrix cores for a block of FP32 columns
// (nrhs = 16 * CT with CT = 1, 2, 4; k <= 16 * KT with KT = 1, 2):
// v_mfma_f32_16x16x4_f32 tiles.  Lane layout of one 16x16x4 product
// D = A B + C: A[i][q] comes from lane i + 16 q, B[q][j] from lane j + 16 q,
// and lane j + 16 r holds D[4 r + e][j] in element e.
// ---------------------------------------------------------------------------
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Stage 1: each wave multiplies W^T (k x its rows) by z (its rows x nrhs)
// four rows per product, with the rank index as the product's rows and the
// columns as its columns; the four waves' tiles are summed in LDS and written
// as the workgroup's partial part[blk * k * nrhs + j * nrhs + c] (summed over
// the workgroups by k_lr_final).  rows_per_wave is a multiple of 4.
template <int KT, int CT>
__global__ __launch_bounds__(256) void k_lr_dot_mfma(int m, int k, int nrhs, const float* __restrict__ W,
                                                     const float* __restrict__ z, float* __restrict__ part,
                                                     int rows_per_wave) {
  __shared__ float red[4][KT * 16][CT * 16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int jj = lane & 15, q = lane >> 4;
  const int64_t gw = (int64_t)blockIdx.x * 4 + wv;
  const int64_t i_begin = gw * rows_per_wave;
  const int64_t i_end = min((int64_t)m, i_begin + rows_per_wave);
  f32x4 acc[KT][CT];
#pragma unroll
  for (int h = 0; h < KT; ++h)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) acc[h][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int nit = i_end > i_begin ? (int)((i_end - i_begin + 3) / 4) : 0;
  for (int it = 0; it < nit; ++it) {
    const int64_t i = i_begin + 4 * it + q;
    const bool ok = i < i_end;
    float a[KT], b[CT];
#pragma unroll
    for (int h = 0; h < KT; ++h) {
      const int j = 16 * h + jj;
      a[h] = (ok && j < k) ? W[i * k + j] : 0.f;
    }
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) b[ct] = ok ? z[i * nrhs + 16 * ct + jj] : 0.f;
#pragma unroll
    for (int h = 0; h < KT; ++h)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) acc[h][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[h], b[ct], acc[h][ct], 0, 0, 0);
  }
#pragma unroll
  for (int h = 0; h < KT; ++h)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int e = 0; e < 4; ++e) red[wv][16 * h + 4 * q + e][16 * ct + jj] = acc[h][ct][e];
  __syncthreads();
  for (int t = threadIdx.x; t < k * nrhs; t += 256) {
    const int j = t / nrhs, c = t % nrhs;
    part[(int64_t)blockIdx.x * k * nrhs + t] = (red[0][j][c] + red[1][j][c]) + (red[2][j][c] + red[3][j][c]);
  }
}

// Stage 3: z[i, c] += sum_j WH[i, j] t[j, c], one 16-row tile per wave.  The
// rank index of product q and operand row p is j = 16 h + 4 p + q, so a lane
// loads WH[i, 16 h + 4 p .. + 3] as one 16-byte vector (k % 4 == 0) and
// feeds its four values to four products; t's operands use the same j.
template <int KT, int CT>
__global__ __launch_bounds__(256) void k_lr_update_mfma(int m, int k, int nrhs, const float* __restrict__ WH,
                                                        const float* __restrict__ t, float* __restrict__ z) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int jj = lane & 15, p = lane >> 4;
  const int64_t i0 = ((int64_t)blockIdx.x * 4 + wv) * 16;
  if (i0 >= m) return;  // whole wave
  float bt[KT][4][CT];
#pragma unroll
  for (int h = 0; h < KT; ++h)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = 16 * h + 4 * p + q;
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) bt[h][q][ct] = j < k ? t[j * nrhs + 16 * ct + jj] : 0.f;
    }
  const int64_t i = i0 + jj;
  float a[KT][4];
#pragma unroll
  for (int h = 0; h < KT; ++h) {
    const int j0 = 16 * h + 4 * p;
    if ((k & 3) == 0 && i < m && j0 + 3 < k) {
      const float4 v = *reinterpret_cast<const float4*>(WH + i * k + j0);
      a[h][0] = v.x;
      a[h][1] = v.y;
      a[h][2] = v.z;
      a[h][3] = v.w;
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) a[h][q] = (i < m && j0 + q < k) ? WH[i * k + j0 + q] : 0.f;
    }
  }
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    f32x4 c;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t r = i0 + 4 * p + e;
      c[e] = r < m ? z[r * nrhs + 16 * ct + jj] : 0.f;
    }
#pragma unroll
    for (int h = 0; h < KT; ++h)
#pragma unroll
      for (int q = 0; q < 4; ++q) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[h][q], bt[h][q][ct], c, 0, 0, 0);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t r = i0 + 4 * p + e;
      if (r < m) z[r * nrhs + 16 * ct + jj] = c[e];
    }
  }
}

// ---------------------------------------------------------------------------
// Krylov kernels (FP64).  A block of p vectors V_j (each n x nrhs) at
// V + j*ldv.  State per column in KState.
// ---------------------------------------------------------------------------
// Delayed CGS2: a tentative basis vector whose squared distance from the
// span of the final ones, q.q - |s|^2, is at most this fraction of q.q is
// taken as dependent (it lies in the span to working precision).
constexpr double kDcgs2DependentTol = 1e-10;

struct KState {
  double normb, tolr, normr, scale;
  int active;   // running in the current cycle
  int status;   // 0 running, 1 converged, 2 breakdown, 3 max iterations, 4 zero rhs/residual,
                // 5 non-finite value (NaN or Inf in b, x0 or the iteration): retired at once
  int iters;    // total inner iterations (FlexGmresClass::_iter)
  int icur;     // inner steps of the current cycle
};

// partial dot products d[j, c] = sum_rows conj(V_j[r, c]) * w[r, c] for j < p
// (w may be one of the V_j), accumulated in S.  One partial per workgroup.
template <typename VV, typename VW, typename S, int NT>
__global__ __launch_bounds__(NT) void k_multidot(int64_t n, int nrhs, int p, const VV* __restrict__ V, int64_t ldv,
                                                 const VW* __restrict__ w, S* __restrict__ part,
                                                 int64_t rows_per_blk) {
  __shared__ S sh[NT];
  const int tid = threadIdx.x;
  const int c = tid % nrhs;
  const int lanes = NT / nrhs;
  const int rl = tid / nrhs;
  const int64_t i0 = (int64_t)blockIdx.x * rows_per_blk;
  const int64_t i1 = min(n, i0 + rows_per_blk);
  for (int j = 0; j < p; ++j) {
    const VV* vj = V + j * ldv;
    S acc = 0;
    for (int64_t i = i0 + rl; i < i1; i += lanes) acc += sconj((S)vj[i * nrhs + c]) * (S)w[i * nrhs + c];
    // reduce across lanes that share column c
    for (int m = kWave / 2; m >= nrhs; m >>= 1) acc += shfl_xor_v(acc, m);
    const int wid = tid / kWave, lane = tid % kWave;
    if (lane < nrhs) sh[wid * nrhs + lane] = acc;
    __syncthreads();
    if (tid < nrhs) {
      S s = 0;
      for (int q = 0; q < NT / kWave; ++q) s += sh[q * nrhs + tid];
      part[((int64_t)blockIdx.x * p + j) * nrhs + tid] = s;
    }
    __syncthreads();
  }
}

// per-workgroup partial sums of |x[r, c]|^2
template <typename S, int NT>
__global__ __launch_bounds__(NT) void k_colnorm2(int64_t n, int nrhs, const S* __restrict__ x,
                                                 double* __restrict__ part, int64_t rows_per_blk) {
  __shared__ double sh[NT];
  const int tid = threadIdx.x;
  const int c = tid % nrhs;
  const int lanes = NT / nrhs;
  const int64_t i0 = (int64_t)blockIdx.x * rows_per_blk;
  const int64_t i1 = min(n, i0 + rows_per_blk);
  double acc = 0;
  for (int64_t i = i0 + tid / nrhs; i < i1; i += lanes) acc += (double)sabs2(x[i * nrhs + c]);
  for (int m = kWave / 2; m >= nrhs; m >>= 1) acc += __shfl_xor(acc, m);
  const int wid = tid / kWave, lane = tid % kWave;
  if (lane < nrhs) sh[wid * nrhs + lane] = acc;
  __syncthreads();
  if (tid < nrhs) {
    double s = 0;
    for (int q = 0; q < NT / kWave; ++q) s += sh[q * nrhs + tid];
    part[(int64_t)blockIdx.x * nrhs + tid] = s;
  }
}

// out[g] = sum over the nblk partials of output g (one wave per output; the
// partials are summed in a fixed order, so the result does not depend on
// timing).  Grid: ceil(p*nrhs / 4) workgroups of 256 threads.
__global__ __launch_bounds__(256) void k_dot_final_wave(int nblk, int p, int nrhs, const double* __restrict__ part,
                                                        double* __restrict__ out, int accumulate) {
  const int g = blockIdx.x * 4 + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int total = p * nrhs;
  if (g >= total) return;  // whole wave exits together
  double s = 0;
  for (int b = lane; b < nblk; b += kWave) s += part[(int64_t)b * total + g];
#pragma unroll
  for (int m = kWave / 2; m > 0; m >>= 1) s += __shfl_xor(s, m);
  if (lane == 0) out[g] = accumulate ? out[g] + s : s;
}

// w[r, c] -= sum_{j<p} V_j[r, c] * h[j, c]  and, when part != nullptr, the
// per-workgroup partial sums of |w[r, c]|^2 (part[blk * nrhs + c]).  V and w
// hold VV, h and the arithmetic S (VV = S, or float vectors with double
// coefficients).  Dynamic shared memory: p*nrhs values of S, then
// (NT/64)*nrhs doubles.
template <typename VV, typename S, int NT>
__global__ __launch_bounds__(NT) void k_multi_axpy_norm(int64_t n, int nrhs, int p, const VV* __restrict__ V,
                                                        int64_t ldv, const S* __restrict__ h,
                                                        VV* __restrict__ w, double* __restrict__ part) {
  extern __shared__ unsigned char ma_smem[];
  S* hs = reinterpret_cast<S*>(ma_smem);
  double* red = reinterpret_cast<double*>(hs + p * nrhs);
  for (int t = threadIdx.x; t < p * nrhs; t += NT) hs[t] = h[t];
  __syncthreads();
  const int tid = threadIdx.x;
  const int c = tid % nrhs;
  const int64_t per_blk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * per_blk;
  const int64_t i1 = min(n, i0 + per_blk);
  const int lanes = NT / nrhs;
  double sq = 0;
  for (int64_t i = i0 + tid / nrhs; i < i1; i += lanes) {
    const int64_t g = i * nrhs + c;
    S acc = (S)w[g];
    for (int j = 0; j < p; ++j) acc -= (S)V[j * ldv + g] * hs[j * nrhs + c];
    const VV out = (VV)acc;
    w[g] = out;
    sq += (double)sabs2(out);
  }
  if (part == nullptr) return;  // uniform
  for (int m = kWave / 2; m >= nrhs; m >>= 1) sq += __shfl_xor(sq, m);
  const int wid = tid / kWave, lane = tid % kWave;
  if (lane < nrhs) red[wid * nrhs + lane] = sq;
  __syncthreads();
  if (tid < nrhs) {
    double s = 0;
    for (int q = 0; q < NT / kWave; ++q) s += red[q * nrhs + tid];
    part[(int64_t)blockIdx.x * nrhs + tid] = s;
  }
}

// One classical Gram-Schmidt sweep over the basis V_0..V_{p-1} with an
// element's basis values and the partial sums held in registers.  LPE lanes
// share an element (consecutive lanes); lane `sub` of an element holds V_j for
// j = u * LPE + sub, u < pu = ceil(p / LPE) <= PM (compile-time bound; loops
// are unrolled to PM with the wave-uniform guard u < pu, and all loads of an
// element are issued before any is used).  With LPE = 2 and p odd, lane 1
// reads V_p (which must be readable) and gives it weight zero.
//   MODE 0: d[j, c] = sum_rows V_j[r, c] * w[r, c]
//   MODE 1: w -= sum_j V_j h[j, :], then d as in mode 0 with the updated w
//           from the same registers (V is read once for both)
//   MODE 2: w -= sum_j V_j h[j, :], then d[0, c] = sum_rows w[r, c]^2
// The update sums V_j h[j, :] in increasing j within a lane (with LPE = 2 the
// two lanes' sums are added) and subtracts the sum.  Every lane walks the
// elements (blockIdx.x * NT + tid) / LPE + k * gridDim.x * NT / LPE; at the
// end the sums are reduced over the lanes of a column, then over the waves,
// into part[blk * q * nrhs + j * nrhs + c], q = p (modes 0, 1) or 1 (mode 2),
// in a fixed order (deterministic).  Requires nrhs * LPE <= 64.
// Dynamic LDS: (MODE ? p * nrhs : 0) + (NT / 64) * q * nrhs doubles.
template <int NT, int PM, int MODE, int LPE>
__global__ __launch_bounds__(NT) void k_basis_reg(int64_t n, int nrhs, int p, const double* __restrict__ V,
                                                  int64_t ldv, const double* __restrict__ h, double* __restrict__ w,
                                                  double* __restrict__ part) {
  static_assert(MODE != 2 || LPE == 1, "the norm sweep uses one lane per element");
  constexpr int NA = MODE == 2 ? 1 : PM;  // partial sums per lane
  const int q = MODE == 2 ? 1 : p;        // outputs per column
  extern __shared__ double sm[];
  double* hs = sm;
  double* red = sm + (MODE ? p * nrhs : 0);
  const int tid = threadIdx.x;
  if (MODE)
    for (int t = tid; t < p * nrhs; t += NT) hs[t] = h[t];
  __syncthreads();
  const int sub = tid % LPE;
  const int c = (tid / LPE) % nrhs;
  const int pu = (p + LPE - 1) / LPE;
  const double* Vl = V + (int64_t)sub * ldv;
  double acc[NA];
#pragma unroll
  for (int u = 0; u < NA; ++u) acc[u] = 0;
  const int64_t total = n * nrhs;
  const int64_t stride = (int64_t)gridDim.x * (NT / LPE);
  for (int64_t g = ((int64_t)blockIdx.x * NT + tid) / LPE; g < total; g += stride) {
    double wv = w[g];
    double v[PM];
#pragma unroll
    for (int u = 0; u < PM; ++u)
      if (u < pu) v[u] = ld_stream(Vl + (int64_t)(u * LPE) * ldv + g);
    if constexpr (MODE != 0) {
      double sv = 0;
#pragma unroll
      for (int u = 0; u < PM; ++u)
        if (u < pu) {
          const int j = u * LPE + sub;
          sv += v[u] * (j < p ? hs[j * nrhs + c] : 0.0);
        }
      if constexpr (LPE == 2) sv += __shfl_xor(sv, 1);
      wv -= sv;
      if (sub == 0) w[g] = wv;
    }
    if constexpr (MODE == 2) {
      acc[0] += wv * wv;
    } else {
#pragma unroll
      for (int u = 0; u < PM; ++u)
        if (u < pu) acc[u] += v[u] * wv;
    }
  }
  const int wid = tid / kWave, lane = tid % kWave;
#pragma unroll
  for (int u = 0; u < NA; ++u) {
    if (u < (MODE == 2 ? 1 : pu)) {
      const int j = MODE == 2 ? 0 : u * LPE + sub;
      double s = acc[u];
      for (int m = kWave / 2; m >= nrhs * LPE; m >>= 1) s += __shfl_xor(s, m);
      if (lane < nrhs * LPE && j < q) red[(wid * q + j) * nrhs + lane / LPE] = s;
    }
  }
  __syncthreads();
  for (int t = tid; t < q * nrhs; t += NT) {
    double s = 0;
    for (int k = 0; k < NT / kWave; ++k) s += red[k * q * nrhs + t];
    part[(int64_t)blockIdx.x * q * nrhs + t] = s;
  }
}

// ---------------------------------------------------------------------------
// Delayed CGS2 (PGS_ORTH_DCGS2, real bases): the re-orthogonalization of the
// newest basis vector q is done in the next step's sweep, together with the
// first projection of the next vector w, so a step reads the basis twice
// instead of three times.  Basis slots 0..p-1 are final, q is tentative.
// ---------------------------------------------------------------------------

// Dots of the basis chunk V_{j0..j0+cnt-1} with q and with w (w may be null):
// per workgroup partials part[blk * nout * nrhs + o * nrhs + c] for
//   o < cnt:             V_{j0+o} . q
//   cnt <= o < 2 cnt:    V_{j0+o-cnt} . w          (w != null)
//   o = 2 cnt, 2 cnt+1:  q . q and q . w           (first chunk, w != null)
//   o = cnt:             q . q                     (first chunk, w == null)
// nout = 2 cnt + 2 (w != null, first), 2 cnt (w != null), cnt + 1 / cnt
// (w == null, first / not).  LPE lanes share an element, lane `sub` holds
// V_{j0 + u LPE + sub}, u < ceil(cnt / LPE) <= PM; every lane walks the
// elements with a grid stride; reductions over lanes, then over waves in a
// fixed order (one LDS table of nout * nrhs doubles, the waves add in turn).
// With nrhs * LPE > 64 a wave covers only part of a row's columns, and each
// of its lanes adds its own sums (no lane reduction).
template <typename S, int NT, int PM, int LPE>
__global__ __launch_bounds__(NT) void k_dcgs2_dots(int64_t n, int nrhs, int j0, int cnt, const S* __restrict__ V,
                                                   int64_t ldv, const S* __restrict__ q, const S* __restrict__ w,
                                                   int first, S* __restrict__ part) {
  extern __shared__ unsigned char dd_raw[];
  S* dd_sm = reinterpret_cast<S*>(dd_raw);
  const bool hw = w != nullptr;
  const int nout = (hw ? 2 * cnt : cnt) + (first ? (hw ? 2 : 1) : 0);
  const int tid = threadIdx.x;
  const int sub = tid % LPE;
  const int cth = (tid / LPE) % nrhs;  // the lane's column (the grid stride keeps it)
  const int pu = (cnt + LPE - 1) / LPE;
  const S* Vl = V + (int64_t)(j0 + sub) * ldv;
  S aq[PM], aw[PM];
  S aa = S(0.0), ab = S(0.0);
#pragma unroll
  for (int u = 0; u < PM; ++u) {
    aq[u] = S(0.0);
    aw[u] = S(0.0);
  }
  const int64_t total = n * nrhs;
  const int64_t stride = (int64_t)gridDim.x * (NT / LPE);
  for (int64_t g = ((int64_t)blockIdx.x * NT + tid) / LPE; g < total; g += stride) {
    const S qv = q[g];
    const S wv = hw ? w[g] : S(0.0);
    S v[PM];
#pragma unroll
    for (int u = 0; u < PM; ++u)
      if (u < pu && u * LPE + sub < cnt) v[u] = ld_stream(Vl + (int64_t)(u * LPE) * ldv + g);
      else v[u] = S(0.0);
    // dots conjugate the basis (x^H y)
#pragma unroll
    for (int u = 0; u < PM; ++u)
      if (u < pu) {
        const S vc = sconj(v[u]);
        aq[u] += vc * qv;
        aw[u] += vc * wv;
      }
    if (sub == 0) {
      aa += sconj(qv) * qv;
      ab += sconj(qv) * wv;
    }
  }
  // reduce over the lanes of a column (lanes c LPE + sub keep the sums)
#pragma unroll
  for (int u = 0; u < PM; ++u) {
    if (u < pu) {
      for (int mm = kWave / 2; mm >= nrhs * LPE; mm >>= 1) {
        aq[u] += shfl_xor_v(aq[u], mm);
        aw[u] += shfl_xor_v(aw[u], mm);
      }
    }
  }
  for (int mm = kWave / 2; mm >= nrhs * LPE; mm >>= 1) {
    aa += shfl_xor_v(aa, mm);
    ab += shfl_xor_v(ab, mm);
  }
  for (int t = tid; t < nout * nrhs; t += NT) dd_sm[t] = S(0.0);
  __syncthreads();
  // the waves add their sums to the table in turn (a fixed order)
  const int wid = tid / kWave, lane = tid % kWave;
  for (int k = 0; k < NT / kWave; ++k) {
    if (wid == k && lane < nrhs * LPE) {
#pragma unroll
      for (int u = 0; u < PM; ++u) {
        const int j = u * LPE + sub;
        if (u < pu && j < cnt) {
          dd_sm[j * nrhs + cth] += aq[u];
          if (hw) dd_sm[(cnt + j) * nrhs + cth] += aw[u];
        }
      }
      if (first && lane % LPE == 0) {
        const int o = hw ? 2 * cnt : cnt;
        dd_sm[o * nrhs + cth] += aa;
        if (hw) dd_sm[(o + 1) * nrhs + cth] += ab;
      }
    }
    __syncthreads();
  }
  for (int t = tid; t < nout * nrhs; t += NT) part[(int64_t)blockIdx.x * nout * nrhs + t] = dd_sm[t];
}

// w <- w - sum_{j<p} V_j e_j - e_p q: the projection of w on the final
// basis v_0..v_p, written in terms of the raw basis slots 0..p (q = slot p,
// which stays as it is), with per-workgroup partials of |w|^2 in
// part[blk * nrhs + c].  coef = [e_0..e_{p-1} (p x nrhs) | e_p (nrhs)],
// copied to LDS.  Element layout as k_basis_reg (LPE lanes per element, lane
// sub holds V_{u LPE + sub}).
template <typename S, int NT, int PM, int LPE>
__global__ __launch_bounds__(NT) void k_dcgs2_update(int64_t n, int nrhs, int p, const S* __restrict__ V,
                                                     int64_t ldv, const S* __restrict__ q, S* __restrict__ w,
                                                     const S* __restrict__ coef, double* __restrict__ part) {
  extern __shared__ unsigned char du_raw[];
  const int nc = (p + 1) * nrhs;
  S* cs = reinterpret_cast<S*>(du_raw);
  double* red = reinterpret_cast<double*>(cs + nc);
  const int tid = threadIdx.x;
  for (int t = tid; t < nc; t += NT) cs[t] = coef[t];
  // a wave covers 64 / LPE elements: all columns when nrhs * LPE <= 64, else
  // half of them, and the other half of its row of the table stays zero
  for (int t = tid; t < (NT / kWave) * nrhs; t += NT) red[t] = 0.0;
  __syncthreads();
  const S* gc = cs;
  const S* gam = cs + p * nrhs;
  const int sub = tid % LPE;
  const int c = (tid / LPE) % nrhs;
  const int pu = (p + LPE - 1) / LPE;
  const S* Vl = V + (int64_t)sub * ldv;
  double acc = 0.0;
  const int64_t total = n * nrhs;
  const int64_t stride = (int64_t)gridDim.x * (NT / LPE);
  for (int64_t g = ((int64_t)blockIdx.x * NT + tid) / LPE; g < total; g += stride) {
    S v[PM];
#pragma unroll
    for (int u = 0; u < PM; ++u)
      if (u < pu) v[u] = (u * LPE + sub < p) ? ld_stream(Vl + (int64_t)(u * LPE) * ldv + g) : S(0.0);
    S gv = S(0.0);
#pragma unroll
    for (int u = 0; u < PM; ++u)
      if (u < pu) {
        const int j = u * LPE + sub;
        gv += v[u] * (j < p ? gc[j * nrhs + c] : S(0.0));
      }
    if constexpr (LPE == 2) gv += shfl_xor_v(gv, 1);
    if (sub == 0) {
      const S un = w[g] - gv - gam[c] * q[g];
      w[g] = un;
      acc += (double)sabs2(un);
    }
  }
  const int wid = tid / kWave, lane = tid % kWave;
  for (int mm = kWave / 2; mm >= nrhs * LPE; mm >>= 1) acc += __shfl_xor(acc, mm);
  if (lane < nrhs * LPE && lane % LPE == 0) red[wid * nrhs + c] = acc;
  __syncthreads();
  for (int t = tid; t < nrhs; t += NT) {
    double s2 = 0.0;
    for (int k = 0; k < NT / kWave; ++k) s2 += red[k * nrhs + t];
    part[(int64_t)blockIdx.x * nrhs + t] = s2;
  }
}

// Per column (one workgroup of 64 lanes, blockIdx.x = column), after the
// dots of step p.  The basis slots hold raw vectors: slot j < p is step j's
// first projection as k_dcgs2_update left it, and the final (orthonormal)
// basis is v_j = sum_{i<=j} slot_i R[i, j], R upper triangular (row-major,
// (m+1) x (m+1) per column); slot p = q still waits for its scale sg
// (vscale).  From the raw dots d = [slot_j^H q (p) | slot_j^H w (p) | q.q |
// q.w] (final pass: [slot_j^H q | q.q]):
//   s_j = v_j^H (sg q) = sg sum_{i<=j} conj(R[i, j]) d_i,  a_j = v_j^H w alike;
// then q's norm nu = sqrt(sg^2 q.q - |s|^2), the correction of the tentative
// Hessenberg column p-1 (H[0..p-1, p-1] += H[p, p-1] s; H[p, p-1] *= nu) and
// its Givens update as k_givens (convergence test, cycle end); then (not
// final) H[0..p-1, p] = a, H[p, p] = c_p = (sg q.w - s^H a) / nu.  The final
// v_p = (sg q - sum_j v_j s_j) / nu is R's column p: R[i, p] = -(R s)_i / nu,
// R[p, p] = sg / nu; and w - sum_{j<p} v_j a_j - c_p v_p = w - sum_{i<=p}
// slot_i e_i with e_i = (R (a - gm s))_i for i < p, e_p = sg gm, gm = c_p / nu:
// coef = [e_0..e_p] (rows of nrhs) for k_dcgs2_update, zero for columns that
// are not active.  p = 0: q is the exact v0 (nu = 1, no column to finalize).
// Dynamic LDS: 2 p values of S.
template <typename S>
__global__ __launch_bounds__(64) void k_dcgs2_fix(int nrhs, int m, int p, int maxits, int final_pass,
                                                  const S* __restrict__ d, S* __restrict__ H, S* __restrict__ cs,
                                                  S* __restrict__ sn, S* __restrict__ rs, KState* __restrict__ st,
                                                  S* __restrict__ coef, const double* __restrict__ vscale,
                                                  S* __restrict__ R) {
  extern __shared__ unsigned char fx_raw[];
  S* sv = reinterpret_cast<S*>(fx_raw);  // s_j
  S* av = sv + p;                        // a_j (not final)
  __shared__ int go_sh;
  __shared__ double in_sh;
  __shared__ S gm_sh;
  const int c = blockIdx.x;
  const int lane = threadIdx.x;
  const int64_t ldr = m + 1;
  S* Rc = R + (int64_t)c * ldr * ldr;
  const double sg = vscale ? vscale[c] : 1.0;
  const bool was_active = st[c].active != 0;
  // s and a on the final basis (lanes over j, rows of R read along j)
  if (was_active) {
    for (int j = lane; j < p; j += kWave) {
      S ts = S(0.0), ta = S(0.0);
      for (int i = 0; i <= j; ++i) {
        const S r = sconj(Rc[i * ldr + j]);
        ts += r * d[i * nrhs + c];
        if (!final_pass) ta += r * d[(p + i) * nrhs + c];
      }
      sv[j] = S(sg) * ts;
      if (!final_pass) av[j] = ta;
    }
  }
  __syncthreads();
  if (lane == 0) {
    // the column's scalar part; returns whether w gets coefficients
    auto scalar_part = [&]() -> bool {
      KState s = st[c];
      if (!s.active) return false;
      const double eps = 2.220446049250313e-16;
      // d rows: slot_j^H q (j < p), then slot_j^H w (j < p), then q.q, q.w
      // (final: slot_j^H q, q.q)
      const double qq = sg * sg * (double)sreal(final_pass ? d[p * nrhs + c] : d[2 * p * nrhs + c]);
      const S qw = final_pass ? S(0.0) : S(sg) * d[(2 * p + 1) * nrhs + c];
      double nu = 1.0;
      S* Hc = H + (int64_t)c * (m + 1) * m;
      auto retire = [&](int status) {
        s.status = status;
        s.active = 0;
        s.scale = 0.0;
        st[c] = s;
      };
      if (!isfinite(qq) || !isfinite(sabs_d(qw))) {
        retire(5);
        return false;
      }
      // the tentative v_p lies in the span of v_0..v_{p-1} to working
      // precision when q.q - |s|^2 cancels (|s| = 1 - 5e-11 and closer): nu
      // and the final v_p = (q - V s) / nu are then rounding noise, so the
      // column's cycle ends with Hessenberg column p - 1 (whose subdiagonal
      // nu * h is negligible there, a lucky breakdown) and restarts from its
      // true residual instead of continuing on a basis vector that is not
      // orthogonal
      bool dependent = false;
      if (p > 0) {
        double ss = 0.0;
        for (int j = 0; j < p; ++j) ss += (double)sabs2(sv[j]);
        if (!isfinite(ss)) {
          retire(5);
          return false;
        }
        dependent = !(qq - ss > kDcgs2DependentTol * qq);
        nu = sqrt(fmax(qq - ss, 0.0));
        const int col = p - 1;
        const S hq = Hc[p + col * (m + 1)];
        for (int j = 0; j < p; ++j) Hc[j + col * (m + 1)] += hq * sv[j];
        Hc[p + col * (m + 1)] = hq * S(nu);
        // Givens update of column col, as k_givens (fgmres.hpp:428-461)
        S* csc = cs + (int64_t)c * m;
        S* snc = sn + (int64_t)c * m;
        S* rsc = rs + (int64_t)c * (m + 1);
        s.iters += 1;
        s.icur = p;
        for (int j = 1; j < p; ++j) {
          const S hii = Hc[(j - 1) + col * (m + 1)];
          const S hj = Hc[j + col * (m + 1)];
          Hc[(j - 1) + col * (m + 1)] = sconj(csc[j - 1]) * hii + snc[j - 1] * hj;
          Hc[j + col * (m + 1)] = -snc[j - 1] * hii + csc[j - 1] * hj;
        }
        const S hii = Hc[(p - 1) + col * (m + 1)];
        const S hii1 = Hc[p + col * (m + 1)];
        const double g = sqrt((double)sabs2(hii) + (double)sabs2(hii1));
        if (!isfinite(g)) {
          retire(5);
          return false;
        }
        if (fabs(g) < eps) {  // breakdown: the reference leaves x untouched (fgmres.hpp:440-443)
          retire(2);
          return false;
        }
        csc[p - 1] = hii / S(g);
        snc[p - 1] = hii1 / S(g);
        rsc[p] = -snc[p - 1] * rsc[p - 1];
        rsc[p - 1] = sconj(csc[p - 1]) * rsc[p - 1];
        Hc[(p - 1) + col * (m + 1)] = sconj(csc[p - 1]) * hii + snc[p - 1] * hii1;
        s.normr = sabs_d(rsc[p]);
        if (s.normr <= s.tolr) {
          s.status = 1;
          s.active = 0;
        } else if (p >= m || s.iters >= maxits || dependent) {
          s.active = 0;  // cycle ends for this column
        }
        st[c] = s;
        if (!s.active) return false;
      }
      if (final_pass) return false;
      const double in = nu > 0.0 ? 1.0 / nu : 0.0;
      // coefficient of w on the final v_p: (q - V s)^H w / nu = (q.w - s^H a) / nu
      S sa = S(0.0);
      for (int j = 0; j < p; ++j) sa += sconj(sv[j]) * av[j];
      const S cp = (qw - sa) * S(in);
      for (int j = 0; j < p; ++j) Hc[j + p * (m + 1)] = av[j];
      Hc[p + p * (m + 1)] = cp;
      in_sh = in;
      gm_sh = cp * S(in);
      return true;
    };
    go_sh = scalar_part() ? 1 : 0;
  }
  __syncthreads();
  if (final_pass) return;
  if (!go_sh) {
    for (int j = lane; j <= p; j += kWave) coef[(int64_t)j * nrhs + c] = S(0.0);
    return;
  }
  const double in = in_sh;
  const S gm = gm_sh;
  // R's column p and the update's coefficients (lanes over rows of R)
  for (int i = lane; i < p; i += kWave) {
    S rsv = S(0.0), rg = S(0.0);
    for (int j = i; j < p; ++j) {
      const S r = Rc[i * ldr + j];
      rsv += r * sv[j];
      rg += r * (av[j] - gm * sv[j]);
    }
    Rc[i * ldr + p] = -rsv * S(in);
    coef[(int64_t)i * nrhs + c] = rg;
  }
  if (lane == 0) {
    Rc[p * ldr + p] = S(sg * in);
    coef[(int64_t)p * nrhs + c] = S(sg) * gm;
  }
}

// Sums the nblk partials of one k_dcgs2_dots launch into the rows of the
// step's dot table d (rows: s_j, then a_j, then q.q and q.w; or s_j, q.q
// without w), one wave per output, in a fixed order.
template <typename S>
__global__ __launch_bounds__(256) void k_dcgs2_final(int nblk, int nout, int cnt, int nrhs, int hw, int p, int j0,
                                                     const S* __restrict__ part, S* __restrict__ d) {
  const int g = blockIdx.x * 4 + threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int total = nout * nrhs;
  if (g >= total) return;  // whole wave exits together
  S s = S(0.0);
  for (int b = lane; b < nblk; b += kWave) s += part[(int64_t)b * total + g];
#pragma unroll
  for (int mm = kWave / 2; mm > 0; mm >>= 1) s += shfl_xor_v(s, mm);
  if (lane != 0) return;
  const int o = g / nrhs, c = g % nrhs;
  int row;
  if (o < cnt) row = j0 + o;
  else if (hw && o < 2 * cnt) row = p + j0 + (o - cnt);
  else row = (hw ? 2 * p : p) + (o - (hw ? 2 * cnt : cnt));
  d[row * nrhs + c] = s;
}

// After k_dcgs2_update of step p: the tentative H[p+1, p] = ||w|| and the
// scale that normalizes w (zero for columns that are not active).
template <typename S>
__global__ void k_dcgs2_norm(int nrhs, int m, int p, const double* __restrict__ t2, S* __restrict__ H,
                             KState* __restrict__ st, double* __restrict__ vscale) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  KState s = st[c];
  if (!s.active) {
    s.scale = 0.0;
  } else if (!isfinite(t2[c])) {
    s.status = 5;
    s.active = 0;
    s.scale = 0.0;
  } else {
    const double t = sqrt(t2[c]);
    H[(int64_t)c * (m + 1) * m + (p + 1) + p * (m + 1)] = S(t);
    s.scale = t > 0.0 ? 1.0 / t : 0.0;
  }
  st[c] = s;
  if (vscale) vscale[c] = s.scale;  // w stays as it is; its readers apply the scale
}

// v[r, c] *= s[c]   (s from KState.scale)
template <typename S>
__global__ void k_scale_cols(int64_t n, int nrhs, const KState* __restrict__ st, S* __restrict__ v) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n * nrhs) return;
  v[g] *= S(st[col_of(g, nrhs)].scale);
}

// FGMRES start (fgmres.hpp:318-359): normb, normr from squared norms.
__global__ void k_fgmres_init(int nrhs, const double* __restrict__ nb2, const double* __restrict__ nr2, double tol,
                              int absolute, KState* __restrict__ st, double* __restrict__ x_zero_flag) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  const double eps = 2.220446049250313e-16;
  KState s;
  s.normb = sqrt(nb2[c]);
  s.normr = sqrt(nr2[c]);
  s.tolr = absolute ? tol : tol * s.normb;
  s.scale = 0.0;
  s.iters = 0;
  s.icur = 0;
  s.active = 0;
  s.status = 0;
  x_zero_flag[c] = 0.0;
  if (!isfinite(s.normb) || !isfinite(s.normr)) {  // NaN or Inf in b or x0: x stays as given
    s.status = 5;
  } else if (s.normb < eps) {  // zero right-hand side: x = 0 (fgmres.hpp:321-328)
    s.status = 4;
    x_zero_flag[c] = 1.0;
  } else if (s.normr < eps) {  // initial guess already exact (fgmres.hpp:341-348)
    s.status = 4;
  } else {
    s.active = 1;
  }
  st[c] = s;
}

// x[r, c] = 0 where flag[c] != 0
template <typename S>
__global__ void k_zero_cols(int64_t n, int nrhs, const double* __restrict__ flag, S* __restrict__ x) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n * nrhs) return;
  if (flag[col_of(g, nrhs)] != 0.0) x[g] = S(0.0);
}

// Start of a restart cycle (fgmres.hpp:382-386): rs[0] = normr, v0 /= normr.
template <typename S>
__global__ void k_cycle_start(int nrhs, int m, KState* __restrict__ st, S* __restrict__ rs,
                              double* __restrict__ vscale = nullptr) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  if (vscale) vscale[c] = 1.0;  // v0 is normalized in place
  KState s = st[c];
  if (s.status == 0) {
    s.active = 1;
    s.icur = 0;
    rs[(int64_t)c * (m + 1)] = S(s.normr);
    s.scale = 1.0 / s.normr;
  } else {
    s.active = 0;
    s.scale = 0.0;
  }
  st[c] = s;
}

// Hessenberg column + Givens update for inner step i (1-based), fgmres.hpp:424-461
// (complex: c = h_ii / gamma and s = h_i+1,i / gamma, rotations applied with
// conj(c) as the reference does).  hcol[j*nrhs + c], j <= i-1 : orthogonalization
// coefficients; t2[c] = ||w||^2.
template <typename S>
__global__ void k_givens(int nrhs, int m, int i, int maxits, const S* __restrict__ hcol,
                         const double* __restrict__ t2, S* __restrict__ H, S* __restrict__ cs,
                         S* __restrict__ sn, S* __restrict__ rs, KState* __restrict__ st) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  KState s = st[c];
  if (!s.active) {
    s.scale = 0.0;
    st[c] = s;
    return;
  }
  const double eps = 2.220446049250313e-16;
  S* Hc = H + (int64_t)c * (m + 1) * m;
  S* csc = cs + (int64_t)c * m;
  S* snc = sn + (int64_t)c * m;
  S* rsc = rs + (int64_t)c * (m + 1);
  const int col = i - 1;
  for (int j = 0; j < i; ++j) Hc[j + col * (m + 1)] = hcol[j * nrhs + c];
  const double t = sqrt(t2[c]);
  Hc[i + col * (m + 1)] = S(t);
  s.iters += 1;
  s.icur = i;
  for (int j = 1; j < i; ++j) {
    const S hii = Hc[(j - 1) + col * (m + 1)];
    const S hj = Hc[j + col * (m + 1)];
    Hc[(j - 1) + col * (m + 1)] = sconj(csc[j - 1]) * hii + snc[j - 1] * hj;
    Hc[j + col * (m + 1)] = -snc[j - 1] * hii + csc[j - 1] * hj;
  }
  const S hii = Hc[(i - 1) + col * (m + 1)];
  const S hii1 = Hc[i + col * (m + 1)];
  const double gam = sqrt((double)sabs2(hii) + (double)sabs2(hii1));
  if (!isfinite(gam) || !isfinite(t)) {  // NaN or Inf from the apply or the matrix product
    s.status = 5;
    s.active = 0;
    s.scale = 0.0;
    st[c] = s;
    return;
  }
  if (fabs(gam) < eps) {  // breakdown: the reference leaves x untouched (fgmres.hpp:440-443)
    s.status = 2;
    s.active = 0;
    s.scale = 0.0;
    st[c] = s;
    return;
  }
  csc[i - 1] = hii / S(gam);
  snc[i - 1] = hii1 / S(gam);
  rsc[i] = -snc[i - 1] * rsc[i - 1];
  rsc[i - 1] = sconj(csc[i - 1]) * rsc[i - 1];
  Hc[(i - 1) + col * (m + 1)] = sconj(csc[i - 1]) * hii + snc[i - 1] * hii1;
  s.normr = sabs_d(rsc[i]);
  s.scale = (t > 0.0) ? 1.0 / t : 0.0;
  if (s.normr <= s.tolr) {
    s.status = 1;
    s.active = 0;
  } else if (i >= m || s.iters >= maxits) {
    s.active = 0;  // cycle ends for this column
  }
  st[c] = s;
}

// Back substitution on the triangularized H (fgmres.hpp:471-480); coef[j*nrhs+c].
template <typename S>
__global__ void k_backsolve(int nrhs, int m, const S* __restrict__ H, S* __restrict__ rs,
                            const KState* __restrict__ st, S* __restrict__ coef, int* __restrict__ upd) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  KState s = st[c];
  for (int j = 0; j < m; ++j) coef[j * nrhs + c] = S(0.0);
  upd[c] = 0;
  if (s.status == 2 || s.status == 4 || s.status == 5 || s.icur == 0) return;
  const S* Hc = H + (int64_t)c * (m + 1) * m;
  S* rsc = rs + (int64_t)c * (m + 1);
  const int i = s.icur;
  rsc[i - 1] /= Hc[(i - 1) + (i - 1) * (m + 1)];
  for (int k = i - 2; k >= 0; --k) {
    for (int j = k + 1; j < i; ++j) rsc[k] -= Hc[k + j * (m + 1)] * rsc[j];
    rsc[k] /= Hc[k + k * (m + 1)];
  }
  for (int j = 0; j < i; ++j) coef[j * nrhs + c] = rsc[j];
  upd[c] = 1;
}

// x[r, c] += sum_j coef[j, c] * Z_j[r, c] for columns with upd[c], summed
// in S, x stored as VX (dynamic shared memory: p*nrhs values of S)
template <typename VZ, typename S, typename VX = S>
__global__ void k_update_x(int64_t n, int nrhs, int p, const VZ* __restrict__ Z, int64_t ldz,
                           const S* __restrict__ coef, const int* __restrict__ upd, VX* __restrict__ x) {
  extern __shared__ unsigned char ux_smem[];
  S* cs = reinterpret_cast<S*>(ux_smem);
  for (int t = threadIdx.x; t < p * nrhs; t += blockDim.x) cs[t] = coef[t];
  __syncthreads();
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n * nrhs) return;
  const int c = col_of(g, nrhs);
  if (!upd[c]) return;
  S acc = (S)x[g];
  for (int j = 0; j < p; ++j) acc += cs[j * nrhs + c] * (S)Z[j * ldz + g];
  x[g] = (VX)acc;
}

// x[t] = 0 for t < m (a kernel rather than a memset: launches may be
// captured into graphs)
template <typename V>
__global__ void k_fill_zero(int64_t m, V* __restrict__ x) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) x[t] = V(0.0);
}

// Restart bookkeeping: any column still running?  flag[0] = number of running columns.
__global__ void k_count_running(int nrhs, const KState* __restrict__ st, int maxits, int* __restrict__ flag) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int cnt = 0;
  for (int c = 0; c < nrhs; ++c)
    if (st[c].status == 0 && st[c].iters < maxits) ++cnt;
  flag[0] = cnt;
}

__global__ void k_count_active(int nrhs, const KState* __restrict__ st, int* __restrict__ flag) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int cnt = 0;
  for (int c = 0; c < nrhs; ++c) cnt += st[c].active ? 1 : 0;
  flag[0] = cnt;
}

// copy v (double) into z (VZ) : identity preconditioner
template <typename S, typename VZ>
__global__ void k_copy_conv(int64_t m, const S* __restrict__ in, VZ* __restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) out[t] = (VZ)in[t];
}

// z[r, c] = v[r, c] * cscale[c] : identity preconditioner on a Krylov vector
// whose normalization is pending
template <typename S, typename VZ>
__global__ void k_copy_scale_cols(int64_t n, int nrhs, const double* __restrict__ cscale, const S* __restrict__ in,
                                  VZ* __restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n * nrhs) out[t] = (VZ)(in[t] * S(cscale[col_of(t, nrhs)]));
}

// ---------------------------------------------------------------------------
// single-vector operations of SequentialVectorClass on the device
// (Axpy / Scale / Fill / Dot, ParGeMSLR/SRC/vectors/sequential_vector.cpp:
// 706-732, 1037-1172; vectorops.cpp:93-146, 633-680)
// ---------------------------------------------------------------------------
template <typename S>
__global__ void k_vec_axpy(int64_t n, S alpha, const S* __restrict__ x, S* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += alpha * x[i];
}

template <typename S>
__global__ void k_vec_scale(int64_t n, S alpha, S* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= alpha;
}

template <typename S>
__global__ void k_vec_fill(int64_t n, S v, S* __restrict__ x) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = v;
}

// Per-workgroup partials of sum_i conj(x_i) y_i: part[blk * 2] (real part),
// part[blk * 2 + 1] (imaginary part, 0 for real values); summed in a fixed
// order by k_dot_final_wave with p = 2.  256 threads.
template <typename S>
__global__ __launch_bounds__(256) void k_vec_dot(int64_t n, const S* __restrict__ x, const S* __restrict__ y,
                                                 double* __restrict__ part) {
  __shared__ double red[2][4];
  double re = 0.0, im = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const S t = sconj(x[i]) * y[i];
    if constexpr (scalar_traits<S>::is_complex) {
      re += t.re;
      im += t.im;
    } else {
      re += t;
    }
  }
#pragma unroll
  for (int m = kWave / 2; m > 0; m >>= 1) {
    re += __shfl_xor(re, m);
    im += __shfl_xor(im, m);
  }
  const int wv = threadIdx.x / kWave;
  if (threadIdx.x % kWave == 0) {
    red[0][wv] = re;
    red[1][wv] = im;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    const double* r = red[threadIdx.x];
    part[(int64_t)blockIdx.x * 2 + threadIdx.x] = (r[0] + r[1]) + (r[2] + r[3]);
  }
}

}  // namespace dev
}  // namespace pgs
