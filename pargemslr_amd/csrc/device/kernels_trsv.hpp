// Block-diagonal ILU solve x = U^{-1} L^{-1} rhs for one GeMSLR level
// (IluClass::SolveHost, ParGeMSLR/SRC/preconditioners/ilu.cpp:3752-3783):
// L unit lower (strict part stored), U strictly upper, Dinv = 1/pivots.
//
// Layout (built on the host by trsv_pack): per diagonal block a sequence of
// steps (level sets, L sweep then U sweep); step s is
//   desc[s] = {row_off, ent_off, nrows, width}   (block-relative offsets)
//   ent[ent_off + t * width + j] = (column, value) of slot j of row t
//   ri[row_off + t]              = (row, 1/pivot)   (1/pivot used on U steps)
// Padded slots hold (0, 0).  x is first loaded with rhs and both sweeps run in
// place, so padded slots only ever read finite values.  The pipelined kernels
// let a lane read slot q + e * LPR of its row unclamped: past the row's width
// that is a padded or a neighbouring entry (a valid column) or, at the end of
// the block, a zero from an out-of-range buffer load, and its weight is
// replaced by zero.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pgs {
namespace dev {

// Complex value of the device path: (re, im), arithmetic as in std::complex.
template <typename R>
struct alignas(2 * sizeof(R)) cplx {
  R re, im;
  cplx() = default;
  __host__ __device__ constexpr cplx(R r) : re(r), im(0) {}
  __host__ __device__ constexpr cplx(R r, R i) : re(r), im(i) {}
  template <typename Q>
  __host__ __device__ explicit constexpr cplx(const cplx<Q>& o) : re((R)o.re), im((R)o.im) {}
  __host__ __device__ cplx& operator+=(const cplx& o) { re += o.re; im += o.im; return *this; }
  __host__ __device__ cplx& operator-=(const cplx& o) { re -= o.re; im -= o.im; return *this; }
  __host__ __device__ cplx& operator*=(const cplx& o) {
    const R r = re * o.re - im * o.im;
    im = re * o.im + im * o.re;
    re = r;
    return *this;
  }
  __host__ __device__ cplx& operator/=(const cplx& o) {
    const R d = o.re * o.re + o.im * o.im;
    const R r = (re * o.re + im * o.im) / d;
    im = (im * o.re - re * o.im) / d;
    re = r;
    return *this;
  }
};
template <typename R> __host__ __device__ inline cplx<R> operator+(cplx<R> a, const cplx<R>& b) { return a += b; }
template <typename R> __host__ __device__ inline cplx<R> operator-(cplx<R> a, const cplx<R>& b) { return a -= b; }
template <typename R> __host__ __device__ inline cplx<R> operator*(cplx<R> a, const cplx<R>& b) { return a *= b; }
template <typename R> __host__ __device__ inline cplx<R> operator/(cplx<R> a, const cplx<R>& b) { return a /= b; }
template <typename R> __host__ __device__ inline cplx<R> operator-(const cplx<R>& a) { return cplx<R>(-a.re, -a.im); }
template <typename R> __host__ __device__ inline bool operator==(const cplx<R>& a, const cplx<R>& b) {
  return a.re == b.re && a.im == b.im;
}

// Scalar traits: real type, conjugate, squared magnitude.
template <typename V> struct scalar_traits { using real = V; static constexpr bool is_complex = false; };
template <typename R> struct scalar_traits<cplx<R>> { using real = R; static constexpr bool is_complex = true; };
template <typename V> __host__ __device__ inline V sconj(const V& v) { return v; }
template <typename R> __host__ __device__ inline cplx<R> sconj(const cplx<R>& v) { return cplx<R>(v.re, -v.im); }
template <typename V> __host__ __device__ inline V sabs2(const V& v) { return v * v; }
template <typename R> __host__ __device__ inline R sabs2(const cplx<R>& v) { return v.re * v.re + v.im * v.im; }
template <typename V> __host__ __device__ inline double sabs_d(const V& v) { return fabs((double)v); }
template <typename R> __host__ __device__ inline double sabs_d(const cplx<R>& v) {
  return sqrt((double)v.re * v.re + (double)v.im * v.im);
}
template <typename V> __host__ __device__ inline V sreal(const V& v) { return v; }
template <typename R> __host__ __device__ inline R sreal(const cplx<R>& v) { return v.re; }

template <typename V>
struct alignas(8) TPair;
template <>
struct alignas(8) TPair<float> {
  int col;
  float val;
};
template <>
struct alignas(16) TPair<double> {  // value first: a 12-byte load takes (value, column)
  double val;
  int col;
  int pad;
};
template <>
struct alignas(16) TPair<cplx<float>> {  // one 12-byte load takes (value, column)
  cplx<float> val;
  int col;
  int pad;
};
template <>
struct alignas(16) TPair<cplx<double>> {  // a 16-byte load takes the value, a 4-byte load the column
  cplx<double> val;
  int col;
  int pad[3];
};

constexpr int kBufferRsrcWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit dwords, no swizzle

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, kBufferRsrcWord3);
}

// Buffer loads: reads outside [0, bytes) of the resource return zeros.
__device__ __forceinline__ int4 bload_i4(__amdgpu_buffer_rsrc_t r, int off) {
  auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  return make_int4((int)v[0], (int)v[1], (int)v[2], (int)v[3]);
}
__device__ __forceinline__ TPair<float> bload_pair(__amdgpu_buffer_rsrc_t r, int off, float) {
  auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
  TPair<float> p;
  p.col = (int)v[0];
  p.val = __uint_as_float(v[1]);
  return p;
}
__device__ __forceinline__ TPair<double> bload_pair(__amdgpu_buffer_rsrc_t r, int off, double) {
  auto v = __builtin_amdgcn_raw_buffer_load_b96(r, off, 0, 0);
  TPair<double> p;
  p.val = __hiloint2double((int)v[1], (int)v[0]);
  p.col = (int)v[2];
  p.pad = 0;
  return p;
}
__device__ __forceinline__ TPair<cplx<float>> bload_pair(__amdgpu_buffer_rsrc_t r, int off, cplx<float>) {
  auto v = __builtin_amdgcn_raw_buffer_load_b96(r, off, 0, 0);
  TPair<cplx<float>> p;
  p.val = cplx<float>(__uint_as_float(v[0]), __uint_as_float(v[1]));
  p.col = (int)v[2];
  p.pad = 0;
  return p;
}
__device__ __forceinline__ TPair<cplx<double>> bload_pair(__amdgpu_buffer_rsrc_t r, int off, cplx<double>) {
  auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  TPair<cplx<double>> p;
  p.val = cplx<double>(__hiloint2double((int)v[1], (int)v[0]), __hiloint2double((int)v[3], (int)v[2]));
  p.col = (int)__builtin_amdgcn_raw_buffer_load_b32(r, off + 16, 0, 0);
  return p;
}
// The first three fields of a step descriptor (row offset, entry offset, rows).
__device__ __forceinline__ int3 bload_i3(__amdgpu_buffer_rsrc_t r, int off) {
  auto v = __builtin_amdgcn_raw_buffer_load_b96(r, off, 0, 0);
  return make_int3((int)v[0], (int)v[1], (int)v[2]);
}

// Sum over the 4 lanes of a quad with DPP quad permutations (xor 1, xor 2).
__device__ __forceinline__ float quad_sum(float v) {
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
  return v;
}
__device__ __forceinline__ double quad_perm_f64(double v, int ctrl_sel) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  if (ctrl_sel == 1) {
    lo = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xF, 0xF, false);
  } else {
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x4E, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x4E, 0xF, 0xF, false);
  }
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double quad_sum(double v) {
  v += quad_perm_f64(v, 1);
  v += quad_perm_f64(v, 2);
  return v;
}
// Sum over lane pairs (xor 1).
__device__ __forceinline__ float pair_sum(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ double pair_sum(double v) { return v + quad_perm_f64(v, 1); }
// The value of the lane q ^ 1 / q ^ 2 of the same quad.
__device__ __forceinline__ float dpp_xor1(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_xor2(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
}
__device__ __forceinline__ double dpp_xor1(double v) { return quad_perm_f64(v, 1); }
__device__ __forceinline__ double dpp_xor2(double v) { return quad_perm_f64(v, 2); }
template <typename R>
__device__ __forceinline__ cplx<R> dpp_xor1(cplx<R> v) { return cplx<R>(dpp_xor1(v.re), dpp_xor1(v.im)); }
template <typename R>
__device__ __forceinline__ cplx<R> dpp_xor2(cplx<R> v) { return cplx<R>(dpp_xor2(v.re), dpp_xor2(v.im)); }
template <typename R>
__device__ __forceinline__ cplx<R> quad_sum(cplx<R> v) { return cplx<R>(quad_sum(v.re), quad_sum(v.im)); }
template <typename R>
__device__ __forceinline__ cplx<R> pair_sum(cplx<R> v) { return cplx<R>(pair_sum(v.re), pair_sum(v.im)); }

// Direct kernel: any width, nrhs columns (row-interleaved); one thread per
// (row of the step, column), plain loads.  Used for levels whose rows are too
// wide for the pipelined kernel.
template <typename V, int NT>
__global__ __launch_bounds__(NT) void k_trsv_direct(const int* __restrict__ blk_row0, const int* __restrict__ blk_step,
                                                    const int* __restrict__ blk_stepU,
                                                    const int64_t* __restrict__ blk_ent0,
                                                    const int64_t* __restrict__ blk_ri0, const int4* __restrict__ desc,
                                                    const TPair<V>* __restrict__ ent, const TPair<V>* __restrict__ ri,
                                                    const V* __restrict__ rhs, V* __restrict__ x, int nrhs) {
  const int b = blockIdx.x;
  const int r0 = blk_row0[b], nb = blk_row0[b + 1] - r0;
  const int tid = threadIdx.x;
  const V* rb = rhs + (int64_t)r0 * nrhs;
  V* xb = x + (int64_t)r0 * nrhs;
  const TPair<V>* eb = ent + blk_ent0[b];
  const TPair<V>* rib = ri + blk_ri0[b];
  for (int i = tid; i < nb * nrhs; i += NT) xb[i] = rb[i];
  __syncthreads();
  const int s0 = blk_step[b], sU = blk_stepU[b], s1 = blk_step[b + 1];
  for (int s = s0; s < s1; ++s) {
    const int4 d = desc[s];
    const bool upper = (s >= sU);
    const int items = d.z * nrhs;
    for (int it = tid; it < items; it += NT) {
      const int t = it / nrhs;
      const int c = it - t * nrhs;
      const TPair<V> rr = rib[d.x + t];
      V acc = xb[(int64_t)rr.col * nrhs + c];
      for (int j = 0; j < d.w; ++j) {
        const TPair<V> p = eb[d.y + t * d.w + j];
        acc -= p.val * xb[(int64_t)p.col * nrhs + c];
      }
      if (upper) acc *= rr.val;
      xb[(int64_t)rr.col * nrhs + c] = acc;
    }
    __syncthreads();
  }
}

// Pipelined kernel: one workgroup per (diagonal block, RHS column); grid
// index = column * nblocks + block, so all columns of a block are dispatched
// to the same XCD (round-robin over 8 XCDs when nblocks % 8 == 0) and share
// its L2 copy of the factors.  A step is solved by LPR lanes per row with EPL
// slots per lane (steps are at most NT/LPR rows and LPR*EPL slots wide).  The
// slots, row pairs and descriptors of the next D steps are in flight while
// the current step runs; every load is unconditional (buffer loads, indices
// clamped into the step) so the hardware wait counters advance in order and
// only the data of the current step is waited for.
// XLDS: the block's column of x lives in LDS (else in global memory).
template <typename V, bool XLDS, int NT, int LPR, int EPL, int D>
__global__ __launch_bounds__(NT) void k_trsv_pipe(int nblocks, const int* __restrict__ blk_row0,
                                                  const int* __restrict__ blk_step, const int* __restrict__ blk_stepU,
                                                  const int64_t* __restrict__ blk_ent0,
                                                  const int64_t* __restrict__ blk_ri0, const int4* __restrict__ desc,
                                                  const TPair<V>* __restrict__ ent, const TPair<V>* __restrict__ ri,
                                                  const V* __restrict__ rhs, V* __restrict__ x, int nrhs) {
  extern __shared__ unsigned char smem_raw[];
  const int b = blockIdx.x % nblocks;
  const int c = blockIdx.x / nblocks;
  const int r0 = blk_row0[b], nb = blk_row0[b + 1] - r0;
  const int sbeg = blk_step[b], send = blk_step[b + 1], sU = blk_stepU[b];
  const int64_t e0 = blk_ent0[b], e1 = blk_ent0[b + 1];
  const int64_t q0 = blk_ri0[b], q1 = blk_ri0[b + 1];
  const __amdgpu_buffer_rsrc_t rs_ent = make_rsrc(ent + e0, (e1 - e0) * (int64_t)sizeof(TPair<V>));
  const __amdgpu_buffer_rsrc_t rs_ri = make_rsrc(ri + q0, (q1 - q0) * (int64_t)sizeof(TPair<V>));
  const __amdgpu_buffer_rsrc_t rs_desc = make_rsrc(desc + sbeg, (int64_t)(send - sbeg) * 16);
  const int tid = threadIdx.x;
  const int q = tid % LPR;
  const int t0 = tid / LPR;
  V* xg = x + (int64_t)r0 * nrhs + c;
  V* xs = XLDS ? reinterpret_cast<V*>(smem_raw) : nullptr;

  TPair<V> pe[D][EPL];
  TPair<V> pr[D];
  int4 dd[D];  // descriptor of the step held by stage k
  int4 dn[D];  // descriptor of the step stage k loads next
  constexpr int PS = (int)sizeof(TPair<V>);

#define PGS_LOAD_STAGE(K, DESC)                                                 \
  do {                                                                          \
    const int4 d_ = (DESC);                                                     \
    dd[K] = d_;                                                                 \
    const int tt_ = min(t0, max(d_.z - 1, 0));                                  \
    pr[K] = bload_pair(rs_ri, (d_.x + tt_) * PS, V(0));                         \
    const int eb_ = (d_.y + tt_ * d_.w + q) * PS;                               \
    _Pragma("unroll") for (int e_ = 0; e_ < EPL; ++e_)                          \
      pe[K][e_] = bload_pair(rs_ent, eb_ + e_ * LPR * PS, V(0));                \
  } while (0)

#pragma unroll
  for (int k = 0; k < D; ++k) {
    PGS_LOAD_STAGE(k, bload_i4(rs_desc, k * 16));
    dn[k] = bload_i4(rs_desc, (k + D) * 16);
  }
  // block column <- rhs column
  const V* rb = rhs + (int64_t)r0 * nrhs + c;
  if (XLDS) {
    for (int i = tid; i < nb; i += NT) xs[i] = rb[(int64_t)i * nrhs];
  } else {
    for (int i = tid; i < nb; i += NT) xg[(int64_t)i * nrhs] = rb[(int64_t)i * nrhs];
  }
  __syncthreads();
  for (int base = sbeg; base < send; base += D) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const int s = base + k;
      const int4 d = dd[k];
      // gather first (all slots, columns are always in range), then mask the values
      const int row = pr[k].col;
      const V xrow = XLDS ? xs[row] : xg[(int64_t)row * nrhs];
      V xv[EPL];
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        const int cj = pe[k][e].col;
        xv[e] = XLDS ? xs[cj] : xg[(int64_t)cj * nrhs];
      }
      V acc = 0;
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        const V w = (q + e * LPR < d.w) ? pe[k][e].val : (V)0;
        acc += w * xv[e];
      }
      if constexpr (LPR == 4) {
        acc = quad_sum(acc);
      } else if constexpr (LPR == 2) {
        acc = pair_sum(acc);
      } else {
#pragma unroll
        for (int m = LPR / 2; m > 0; m >>= 1) acc += __shfl_xor(acc, m, LPR);
      }
      const V dv = pr[k].val;
      if (s < send && q == 0 && t0 < d.z) {
        V v = xrow - acc;
        if (s >= sU) v *= dv;
        if (XLDS) xs[row] = v;
        else xg[(int64_t)row * nrhs] = v;
      }
      if constexpr (XLDS) {
        // the stage registers are free again: start the loads of step s + D
        // before waiting at the barrier (the barrier only waits for LDS)
        PGS_LOAD_STAGE(k, dn[k]);
        dn[k] = bload_i4(rs_desc, (s - sbeg + 2 * D) * 16);
        __syncthreads();
      } else {
        // x lives in global memory: the barrier's release fence waits for all
        // outstanding vector memory operations, so issue the loads after it
        __syncthreads();
        PGS_LOAD_STAGE(k, dn[k]);
        dn[k] = bload_i4(rs_desc, (s - sbeg + 2 * D) * 16);
      }
    }
  }
#undef PGS_LOAD_STAGE
  if (XLDS) {
    for (int i = tid; i < nb; i += NT) xg[(int64_t)i * nrhs] = xs[i];
  }
}

// Ring kernel: one workgroup per (diagonal block, group of CG right-hand-side
// columns); grid index = group * nblocks + block (all groups of a block on one
// XCD, as above).  The packed columns are ring slots: the value of the row
// stored at pair index p lives in LDS slot p & wmask (CG values per slot)
// from its step until its last reader, so the LDS footprint is the sweep
// window, not the block.  The L sweep reads rhs and writes y to x in global
// memory; the U sweep reads y back from x and overwrites it with the solution
// (a row's y is read by the lanes that later write that row).  Rows are
// prefetched D steps ahead (descriptor, row pair, slots); the right-hand side
// of a step, whose address depends on the row pair, is loaded D/2 steps ahead.
// Lane q of a row's LPR lanes finalizes columns q, q + LPR, ...  Lanes past
// a step's rows gather and write nothing.  With 8- and 16-byte values they
// re-read the step's last row (cache lines already fetched); with 4-byte
// values they read the following rows' pairs and slabs (valid ring slots, or
// zeros past the block from the bounded buffers), which saves the clamp's
// instructions in a loop bound by instruction issue.
template <typename V, int CG>
struct alignas(CG * sizeof(V) < 16 ? CG * sizeof(V) : 16) RingRow {
  V v[CG];
};

// Ring-packed entries hold their ring slot times ring_unit<V>(): the byte
// offset of the slot in a ring of 4-column rows, so that 4-column groups (the
// default) gather without scaling the slot and the others scale it by one
// shift.
template <typename V>
constexpr int ring_unit() {
  return 4 * (int)sizeof(V);
}
template <typename V, int CG>
__device__ __forceinline__ unsigned ring_byte(int c) {
  constexpr int RB = CG * (int)sizeof(V), U = ring_unit<V>();
  if constexpr (RB >= U) return (unsigned)c * (RB / U);
  else return (unsigned)c / (unsigned)(U / RB);
}

// A ring row read through its LDS byte address.  The ring kernel has no
// static LDS, so its dynamic LDS starts at address 0 (checked on the host
// when the kernels are set up): the address is the slot's byte offset plus
// the group's ring offset, and with one group the gather takes the entry's
// column as its address unchanged.
template <typename T>
__device__ __forceinline__ T lds_read(unsigned addr) {
  constexpr int N = (int)sizeof(T) / 4;
  if constexpr (N == 1) {
    return __builtin_bit_cast(T, *reinterpret_cast<const __attribute__((address_space(3))) unsigned*>((size_t)addr));
  } else {
    using W = unsigned __attribute__((ext_vector_type(N)));
    return __builtin_bit_cast(T, *reinterpret_cast<const __attribute__((address_space(3))) W*>((size_t)addr));
  }
}

__device__ __forceinline__ float bload_val(__amdgpu_buffer_rsrc_t r, int off, float) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ double bload_val(__amdgpu_buffer_rsrc_t r, int off, double) {
  auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
  return __hiloint2double((int)v[1], (int)v[0]);
}
// Buffer stores (vector memory): writes outside the resource are dropped.
__device__ __forceinline__ void bstore_val(__amdgpu_buffer_rsrc_t r, int off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, off, 0, 0);
}
__device__ __forceinline__ void bstore_val(__amdgpu_buffer_rsrc_t r, int off, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r, off, 0, 0);
}
__device__ __forceinline__ cplx<float> bload_val(__amdgpu_buffer_rsrc_t r, int off, cplx<float>) {
  auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
  return cplx<float>(__uint_as_float(v[0]), __uint_as_float(v[1]));
}
__device__ __forceinline__ cplx<double> bload_val(__amdgpu_buffer_rsrc_t r, int off, cplx<double>) {
  auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  return cplx<double>(__hiloint2double((int)v[1], (int)v[0]), __hiloint2double((int)v[3], (int)v[2]));
}
__device__ __forceinline__ void bstore_val(__amdgpu_buffer_rsrc_t r, int off, cplx<float> v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r, off, 0, 0);
}
__device__ __forceinline__ void bstore_val(__amdgpu_buffer_rsrc_t r, int off, cplx<double> v) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v), r, off, 0, 0);
}

// Two consecutive values in one buffer access (8-byte and smaller values).
template <typename V>
__device__ __forceinline__ RingRow<V, 2> bload_val2(__amdgpu_buffer_rsrc_t r, int off) {
  static_assert(sizeof(V) <= 8, "two values in one access");
  RingRow<V, 2> o;
  if constexpr (sizeof(V) == 4) {
    auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    o = __builtin_bit_cast(RingRow<V, 2>, v);
  } else {
    auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    o = __builtin_bit_cast(RingRow<V, 2>, v);
  }
  return o;
}
template <typename V>
__device__ __forceinline__ void bstore_val2(__amdgpu_buffer_rsrc_t r, int off, const RingRow<V, 2>& o) {
  static_assert(sizeof(V) <= 8, "two values in one access");
  if constexpr (sizeof(V) == 4)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, o), r, off, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, o), r, off, 0, 0);
}

// An empty asm that takes a value as read and rewritten: the value must have
// arrived, and it counts as produced here rather than by its load.
__device__ __forceinline__ void pin_reg(int& v) { __asm__ __volatile__("" : "+v"(v)); }
__device__ __forceinline__ void pin_reg(float& v) { __asm__ __volatile__("" : "+v"(v)); }
__device__ __forceinline__ void pin_reg(double& v) { __asm__ __volatile__("" : "+v"(v)); }
template <typename R>
__device__ __forceinline__ void pin_reg(cplx<R>& v) {
  pin_reg(v.re);
  pin_reg(v.im);
}
template <typename V>
__device__ __forceinline__ void pin_reg(TPair<V>& p) {
  pin_reg(p.col);
  pin_reg(p.val);
}
__device__ __forceinline__ void pin_reg(int3& d) {
  pin_reg(d.x);
  pin_reg(d.y);
  pin_reg(d.z);
}

template <typename V, int NT, int LPR, int EPL, int D, int CG, bool UPPER, int G>
__device__ __forceinline__ void ring_sweep(int s_begin, int s_end, int sbeg, int wmask, int nrhs, int c0,
                                           __amdgpu_buffer_rsrc_t rs_desc, __amdgpu_buffer_rsrc_t rs_ri,
                                           __amdgpu_buffer_rsrc_t rs_ent, __amdgpu_buffer_rsrc_t rs_src, V* xs,
                                           unsigned xsb,
                                           __amdgpu_buffer_rsrc_t rs_dst) {
  constexpr int PS = (int)sizeof(TPair<V>);
  constexpr int VS = (int)sizeof(V);
  constexpr int RW = LPR * EPL;  // slab width of every ring-packed step
  static_assert(D % 2 == 0, "stage depth must be even");
  static_assert((LPR == 4 && (CG == 1 || CG == 2 || CG == 4 || CG == 8)) ||
                    (LPR == 2 && (CG == 1 || CG == 2 || CG == 4)),
                "quad lanes with 1, 2, 4 or 8 columns, or lane pairs with 1, 2 or 4 columns");
  constexpr int NCL = CG == 2 * LPR ? 2 : 1;  // columns finalized per lane
  // a lane's two columns are adjacent: one access for both (8-byte and
  // smaller values)
  constexpr bool kPairIO = NCL == 2 && sizeof(V) <= 8;
  const int tid = threadIdx.x % NT;  // thread within this column group
  const int q = tid % LPR;
  const int t0 = tid / LPR;
  // after the reduction lane q holds column q % CG (lanes q < CG write it), or
  // with two columns per lane (8 columns on quads, 4 on lane pairs) columns
  // 2q and 2q + 1
  int cl[NCL];
  if constexpr (NCL == 2) {
    cl[0] = 2 * q;
    cl[1] = 2 * q + 1;
  } else {
    cl[0] = q % CG;
  }
  const bool writer = q < CG;
  const int dummy = (wmask + 1) * CG;  // LDS slot written by lanes with nothing to write
  // row * nrhs as a shift: the launch admits power-of-two column counts only
  // (a 32-bit multiply is a quarter-rate instruction, two per step otherwise)
  const int csh = __builtin_ctz((unsigned)nrhs);
  constexpr bool kClampRows = sizeof(V) > 4;
  // Gathers through LDS byte addresses and the prologue's register pinning
  // (below) make every geometry faster (level 0 at 128^3: 1 to 64 RHS 3-12 %
  // less time) except one: the one-wave FP32 kernel with 4-column groups
  // (the 32-RHS launch) runs 12 % slower with them (1.47 against 1.31 ms on
  // one device), with or without the other, and at ring depths 8 and 10 as
  // well.  That kernel keeps the generic gather and the plain prologue.
  constexpr bool kLdsPath = !(NT == 64 && LPR == 4 && CG == 4 && sizeof(V) == 4);

  TPair<V> pe[D][EPL];
  TPair<V> pr[D];
  V xr[D][NCL];
  int sw[D];   // ring slot of this lane's row in the step held by stage k, -1 past the step's rows
  int3 dn[D];  // descriptor of the step stage k loads next

#define PGS_RING_PAIR(K, DESC)                                                  \
  do {                                                                          \
    const int3 d_ = (DESC);                                                     \
    const int tt_ = kClampRows ? min(t0, max(d_.z - 1, 0)) : t0;                \
    const int r_ = d_.x + tt_;                                                  \
    sw[K] = t0 < d_.z ? (r_ & wmask) : -1;                                      \
    pr[K] = bload_pair(rs_ri, r_ * PS, V(0));                                   \
  } while (0)
#define PGS_RING_ENT(K, DESC)                                                   \
  do {                                                                          \
    const int3 d_ = (DESC);                                                     \
    const int tt_ = kClampRows ? min(t0, max(d_.z - 1, 0)) : t0;                \
    const int eb_ = (d_.y + tt_ * RW + q) * PS;                                 \
    _Pragma("unroll") for (int e_ = 0; e_ < EPL; ++e_)                          \
      pe[K][e_] = bload_pair(rs_ent, eb_ + e_ * LPR * PS, V(0));                \
  } while (0)
#define PGS_RING_STAGE(K, DESC) \
  do {                          \
    const int3 ds_ = (DESC);    \
    PGS_RING_PAIR(K, ds_);      \
    PGS_RING_ENT(K, ds_);       \
  } while (0)
#define PGS_RING_RHS(K)                                                           \
  do {                                                                            \
    if constexpr (kPairIO) {                                                      \
      const RingRow<V, 2> t_ = bload_val2<V>(rs_src, ((pr[K].col << csh) + c0 + cl[0]) * VS); \
      xr[K][0] = t_.v[0];                                                         \
      xr[K][1] = t_.v[1];                                                         \
    } else {                                                                      \
      _Pragma("unroll") for (int j_ = 0; j_ < NCL; ++j_) xr[K][j_] =              \
          bload_val(rs_src, ((pr[K].col << csh) + c0 + cl[j_]) * VS, V(0));       \
    }                                                                             \
  } while (0)

  const int rel = s_begin - sbeg;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    PGS_RING_STAGE(k, bload_i3(rs_desc, (rel + k) * 16));
    dn[k] = bload_i3(rs_desc, (rel + k + D) * 16);
  }
#pragma unroll
  for (int k = 0; k < D / 2; ++k) PGS_RING_RHS(k);
  // The compiler gives the loop's first steps the load waits of the worst
  // path into them, and the path from here is the worst: the prologue's last
  // loads are right-hand sides and slabs of the first steps, and every
  // iteration would wait for nearly all loads in flight once per 2 D steps.
  // Taking the prologue's registers as redefined here (one wait for its
  // loads, once per sweep) leaves the loop the waits of its own back edge.
#pragma unroll
  for (int k = 0; k < (kLdsPath ? D : 0); ++k) {
    if (k < D / 2) {
#pragma unroll
      for (int j = 0; j < NCL; ++j) pin_reg(xr[k][j]);
    }
    pin_reg(pr[k]);
    pin_reg(dn[k]);
#pragma unroll
    for (int e = 0; e < EPL; ++e) pin_reg(pe[k][e]);
  }

  for (int base = s_begin; base < s_end; base += 2 * D) {
#pragma unroll
    for (int kk = 0; kk < 2 * D; ++kk) {
      const int k = kk % D;
      const int s = base + kk;
      // gather this step's slots from the ring and take what the arithmetic
      // needs out of the stage registers
      RingRow<V, CG> xv[EPL];
      V wgt[EPL];
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        if constexpr (kLdsPath)
          xv[e] = lds_read<RingRow<V, CG>>(xsb + ring_byte<V, CG>(pe[k][e].col));
        else
          xv[e] = *reinterpret_cast<const RingRow<V, CG>*>(reinterpret_cast<const unsigned char*>(xs) +
                                                            ring_byte<V, CG>(pe[k][e].col));
        wgt[e] = pe[k][e].val;  // padded slots weigh zero
      }
      const int slot = sw[k];
      const int row = pr[k].col;
      const V dv = pr[k].val;
      const bool wr = s < s_end && slot >= 0 && writer;
      V xin[NCL];
#pragma unroll
      for (int j = 0; j < NCL; ++j) xin[j] = xr[k][j];
      V acc[CG];
#pragma unroll
      for (int c = 0; c < CG; ++c) acc[c] = 0;
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
#pragma unroll
        for (int c = 0; c < CG; ++c) acc[c] += wgt[e] * xv[e].v[c];
      }
      // sum over the row's lanes, scattered so that lane q ends with column q % CG
      // (with 8 columns: columns 2q and 2q + 1)
      V a[NCL];
      if constexpr (CG == 8) {
        // xor 2: keep columns 0-3 (q bit 1 clear) or 4-7, send the other half
        const bool b0 = q & 1, b1 = q & 2;
        V h[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) h[c] = (b1 ? acc[c + 4] : acc[c]) + dpp_xor2(b1 ? acc[c] : acc[c + 4]);
        // xor 1: keep columns 0-1 of the half (q bit 0 clear) or 2-3
#pragma unroll
        for (int c = 0; c < 2; ++c) a[c] = (b0 ? h[c + 2] : h[c]) + dpp_xor1(b0 ? h[c] : h[c + 2]);
      } else if constexpr (LPR == 2 && CG == 4) {
        // xor 1: keep columns 0-1 (q = 0) or 2-3 (q = 1), send the other pair
        const bool b0 = q & 1;
#pragma unroll
        for (int c = 0; c < 2; ++c) a[c] = (b0 ? acc[c + 2] : acc[c]) + dpp_xor1(b0 ? acc[c] : acc[c + 2]);
      } else if constexpr (LPR == 2) {
        if constexpr (CG == 2) {
          const bool b0 = q & 1;
          a[0] = (b0 ? acc[1] : acc[0]) + dpp_xor1(b0 ? acc[0] : acc[1]);
        } else {
          a[0] = pair_sum(acc[0]);
        }
      } else if constexpr (CG == 4) {
        // xor 1: keep columns with bit 0 = q bit 0, send the other two
        const bool b0 = q & 1, b1 = q & 2;
        const V k0 = (b0 ? acc[1] : acc[0]) + dpp_xor1(b0 ? acc[0] : acc[1]);
        const V k1 = (b0 ? acc[3] : acc[2]) + dpp_xor1(b0 ? acc[2] : acc[3]);
        // xor 2: keep the pair with bit 1 = q bit 1
        a[0] = (b1 ? k1 : k0) + dpp_xor2(b1 ? k0 : k1);
      } else if constexpr (CG == 2) {
        const bool b0 = q & 1;
        const V k0 = (b0 ? acc[1] : acc[0]) + dpp_xor1(b0 ? acc[0] : acc[1]);
        a[0] = k0 + dpp_xor2(k0);
      } else {
        a[0] = quad_sum(acc[0]);
      }
      // lanes with nothing to write store to the dummy LDS slot and out of
      // the buffer's range (dropped): no branch
      if constexpr (kPairIO) {
        RingRow<V, 2> o;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          o.v[j] = xin[j] - a[j];
          if (UPPER) o.v[j] *= dv;
        }
        *reinterpret_cast<RingRow<V, 2>*>(xs + (wr ? slot * CG + cl[0] : dummy)) = o;
        bstore_val2<V>(rs_dst, wr ? ((row << csh) + c0 + cl[0]) * VS : 0x7ffffff0, o);
      } else {
#pragma unroll
        for (int j = 0; j < NCL; ++j) {
          V v = xin[j] - a[j];
          if (UPPER) v *= dv;
          xs[wr ? slot * CG + cl[j] : dummy] = v;
          bstore_val(rs_dst, wr ? ((row << csh) + c0 + cl[j]) * VS : 0x7ffffff0, v);
        }
      }
      // refill this stage with step s + D and load the right-hand side of
      // step s + D/2, whose row pair arrived D/2 steps ago.  The refill
      // follows the last use of the stage's registers, so the old and the
      // new contents are never live together: the loop needs no register
      // copies at its back edge, and no copy waits for a load in flight.
      __builtin_amdgcn_sched_barrier(0);
      PGS_RING_STAGE(k, dn[k]);
      dn[k] = bload_i3(rs_desc, (s - sbeg + 2 * D) * 16);
      PGS_RING_RHS((k + D / 2) % D);
      if constexpr (NT == 64 && G == 1) {
        // one wave: its LDS operations complete in issue order, so the next
        // step's gathers see this step's writes without a wait or barrier;
        // only the compiler must keep the order
        __builtin_amdgcn_wave_barrier();
        __asm__ __volatile__("" ::: "memory");
      } else {
        __syncthreads();
      }
    }
  }
#undef PGS_RING_STAGE
#undef PGS_RING_PAIR
#undef PGS_RING_ENT
#undef PGS_RING_RHS
}

// G column groups of CG columns each share a workgroup of G * NT threads: the
// groups walk the same steps in lockstep (one barrier per step), so their
// identical loads of the block's entries and row pairs are served once by the
// CU's L1 instead of each group streaming them from L2; each group has its own
// ring in LDS.
template <typename V, int NT, int LPR, int EPL, int D, int CG, int G>
__global__ __launch_bounds__(NT * G) void k_trsv_ring(int nblocks, int wmask, const int* __restrict__ blk_row0,
                                                  const int* __restrict__ blk_step, const int* __restrict__ blk_stepU,
                                                  const int64_t* __restrict__ blk_ent0,
                                                  const int64_t* __restrict__ blk_ri0, const int4* __restrict__ desc,
                                                  const TPair<V>* __restrict__ ent, const TPair<V>* __restrict__ ri,
                                                  const V* __restrict__ rhs, V* __restrict__ x, int nrhs) {
  extern __shared__ unsigned char smem_raw[];
  const int grp = threadIdx.x / NT;
  V* xs = reinterpret_cast<V*>(smem_raw) + grp * (wmask + 2) * CG;
  const unsigned xsb = (unsigned)(grp * (wmask + 2) * CG * (int)sizeof(V));  // xs as an LDS byte address
  const int b = blockIdx.x % nblocks;
  const int c0 = (blockIdx.x / nblocks) * CG * G + grp * CG;
  const int r0 = blk_row0[b], nb = blk_row0[b + 1] - r0;
  const int sbeg = blk_step[b], send = blk_step[b + 1], sU = blk_stepU[b];
  const int64_t e0 = blk_ent0[b], e1 = blk_ent0[b + 1];
  const int64_t q0 = blk_ri0[b], q1 = blk_ri0[b + 1];
  const __amdgpu_buffer_rsrc_t rs_ent = make_rsrc(ent + e0, (e1 - e0) * (int64_t)sizeof(TPair<V>));
  const __amdgpu_buffer_rsrc_t rs_ri = make_rsrc(ri + q0, (q1 - q0) * (int64_t)sizeof(TPair<V>));
  const __amdgpu_buffer_rsrc_t rs_desc = make_rsrc(desc + sbeg, (int64_t)(send - sbeg) * 16);
  const int64_t vbytes = (int64_t)nb * nrhs * (int64_t)sizeof(V);
  V* xb = x + (int64_t)r0 * nrhs;
  V* xall = reinterpret_cast<V*>(smem_raw);
  for (int i = threadIdx.x; i < (wmask + 2) * CG * G; i += NT * G) xall[i] = 0;  // rings + dummy slots
  __syncthreads();
  ring_sweep<V, NT, LPR, EPL, D, CG, false, G>(sbeg, sU, sbeg, wmask, nrhs, c0, rs_desc, rs_ri, rs_ent,
                                               make_rsrc(rhs + (int64_t)r0 * nrhs, vbytes), xs, xsb,
                                               make_rsrc(xb, vbytes));
  ring_sweep<V, NT, LPR, EPL, D, CG, true, G>(sU, send, sbeg, wmask, nrhs, c0, rs_desc, rs_ri, rs_ent,
                                              make_rsrc(xb, vbytes), xs, xsb, make_rsrc(xb, vbytes));
}

}  // namespace dev
}  // namespace pgs
