// Device helpers shared by the blocked Bethe-Salpeter kernels: MFMA traits,
// lane exchanges (DPP, readlane), wave reductions and the accumulator planes.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace ldga_bse {

template <typename T>
struct mfma_traits;
template <>
struct mfma_traits<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  typedef double2 c2;
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // accumulator element j of lane group g holds tile row g + 4j (f64 16x16x4)
  static __device__ __forceinline__ int trow(int g, int j) { return g + 4 * j; }
};
template <>
struct mfma_traits<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  typedef float2 c2;
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // accumulator element j of lane group g holds tile row 4g + j (f32 16x16x4)
  static __device__ __forceinline__ int trow(int g, int j) { return 4 * g + j; }
};

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- lane exchange helpers ------------------------------------------------
template <int CTRL, int RM>
__device__ __forceinline__ int dpp32(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, RM, 0xf, false);
}
template <int CTRL, int RM>
__device__ __forceinline__ float dppv(float old, float v) {
  return __int_as_float(dpp32<CTRL, RM>(__float_as_int(old), __float_as_int(v)));
}
template <int CTRL, int RM>
__device__ __forceinline__ double dppv(double old, double v) {
  const int lo = dpp32<CTRL, RM>(__double2loint(old), __double2loint(v));
  const int hi = dpp32<CTRL, RM>(__double2hiint(old), __double2hiint(v));
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float rdlane(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double rdlane(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// f(std::integral_constant<int, j>) for a wave-uniform j in 0..3 (scalar branches)
template <typename F>
__device__ __forceinline__ void with_j(int j, F &&f) {
  switch (j) {
    case 0: f(std::integral_constant<int, 0>{}); break;
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    default: f(std::integral_constant<int, 3>{}); break;
  }
}

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double fast_rcp(double x) { return __builtin_amdgcn_rcp(x); }

// sum over the 16 lanes of each DPP row; every lane of the row gets the sum
template <typename T>
__device__ __forceinline__ T row_sum16(T v) {
  v += dppv<0xB1, 0xf>(v, v);   // quad xor 1
  v += dppv<0x4E, 0xf>(v, v);   // quad xor 2
  v += dppv<0x141, 0xf>(v, v);  // half-row mirror
  v += dppv<0x140, 0xf>(v, v);  // row mirror
  return v;
}
// FP32: each stage is one DPP add (s_nop 1 covers the two wait states between
// a VALU write and a DPP read of the register); same additions, same order
template <>
__device__ __forceinline__ float row_sum16<float>(float v) {
  __asm__("s_nop 1\n\tv_add_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(v));
  __asm__("s_nop 1\n\tv_add_f32_dpp %0, %0, %0 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf" : "+v"(v));
  __asm__("s_nop 1\n\tv_add_f32_dpp %0, %0, %0 row_half_mirror row_mask:0xf bank_mask:0xf" : "+v"(v));
  __asm__("s_nop 1\n\tv_add_f32_dpp %0, %0, %0 row_mirror row_mask:0xf bank_mask:0xf" : "+v"(v));
  return v;
}

// FP64: per stage two DPP moves of the partner's halves into fresh registers
// (every lane of a row reads a lane of the same row, so no old value is kept)
// and one add; same additions, same order as the generic form
template <int Q>
__device__ __forceinline__ double row_stage_f64(double v) {
  int lo, hi;
  if constexpr (Q == 0)
    __asm__("s_nop 1\n\tv_mov_b32_dpp %0, %2 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
            "v_mov_b32_dpp %1, %3 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf"
            : "=&v"(lo), "=&v"(hi) : "v"(__double2loint(v)), "v"(__double2hiint(v)));
  else if constexpr (Q == 1)
    __asm__("s_nop 1\n\tv_mov_b32_dpp %0, %2 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
            "v_mov_b32_dpp %1, %3 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf"
            : "=&v"(lo), "=&v"(hi) : "v"(__double2loint(v)), "v"(__double2hiint(v)));
  else if constexpr (Q == 2)
    __asm__("s_nop 1\n\tv_mov_b32_dpp %0, %2 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
            "v_mov_b32_dpp %1, %3 row_half_mirror row_mask:0xf bank_mask:0xf"
            : "=&v"(lo), "=&v"(hi) : "v"(__double2loint(v)), "v"(__double2hiint(v)));
  else
    __asm__("s_nop 1\n\tv_mov_b32_dpp %0, %2 row_mirror row_mask:0xf bank_mask:0xf\n\t"
            "v_mov_b32_dpp %1, %3 row_mirror row_mask:0xf bank_mask:0xf"
            : "=&v"(lo), "=&v"(hi) : "v"(__double2loint(v)), "v"(__double2hiint(v)));
  return v + __hiloint2double(hi, lo);
}
template <>
__device__ __forceinline__ double row_sum16<double>(double v) {
  v = row_stage_f64<0>(v);
  v = row_stage_f64<1>(v);
  v = row_stage_f64<2>(v);
  return row_stage_f64<3>(v);
}

// maximum over the 64 lanes of a wave, in lane 63 (unsigned keys)
__device__ __forceinline__ unsigned wave_max_u32(unsigned k) {
  __asm__("s_nop 1\n\tv_max_u32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(k));
  __asm__("s_nop 1\n\tv_max_u32_dpp %0, %0, %0 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf" : "+v"(k));
  __asm__("s_nop 1\n\tv_max_u32_dpp %0, %0, %0 row_half_mirror row_mask:0xf bank_mask:0xf" : "+v"(k));
  __asm__("s_nop 1\n\tv_max_u32_dpp %0, %0, %0 row_mirror row_mask:0xf bank_mask:0xf" : "+v"(k));
  __asm__("s_nop 1\n\tv_max_u32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf" : "+v"(k));
  __asm__("s_nop 1\n\tv_max_u32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf" : "+v"(k));
  return k;
}

// ---- complex rank-4 products for the FP32 path on bf16 MFMA -------------------
// An FP32 value is split into three bf16 parts, x = h + m + l (each part the
// round-to-nearest-even bf16 of the remainder; the remainders are exact in FP32
// and the three parts hold all 24 bits, so h + m + l == x exactly).  A product
// a b is taken as the sum of seven of the nine part products (all but l*m and
// l*l, which are below 2^-24 |a b|), each exact in FP32: its error is below
// that of one FP32 multiplication.  A complex update C += a b (K = 4) becomes
// two 16x16x32 bf16 MFMAs per plane: lane group g holds the eight K slots of
// panel column k = g, a dword pairing the real part (low half) with the
// imaginary part (high half) of one split level.
//   B operand (both planes, both MFMAs):  [h | m | l | h] of (br, bi)
//   real plane A:       MFMA 1 [h h h m], MFMA 2 [l m m 0] of (ar, -ai)
//   imaginary plane A:  the same levels of (ai, ar)
// i.e. slot pairs (A, B) = (h,h) (h,m) (h,l) (m,h) | (l,h) (m,m) (m,l) (0,h).
// Unlike the FP32 MFMA, bf16 MFMAs leave the VALU free to issue other waves'
// instructions while they run.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned pk_bf16(float lo, float hi) {
  const bf16x2 v = {(__bf16)lo, (__bf16)hi};
  return __builtin_bit_cast(unsigned, v);
}
__device__ __forceinline__ float bf_lo(unsigned u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf_hi(unsigned u) { return __uint_as_float(u & 0xffff0000u); }

struct Split3 {
  unsigned h, m, l;  // [re part | im part << 16] at each level
};
__device__ __forceinline__ Split3 split3(float x, float y) {
  Split3 s;
  s.h = pk_bf16(x, y);
  x -= bf_lo(s.h);
  y -= bf_hi(s.h);
  s.m = pk_bf16(x, y);
  x -= bf_lo(s.m);
  y -= bf_hi(s.m);
  s.l = pk_bf16(x, y);
  return s;
}

// the B operand of a value (br, bi): panel row k = g, one column
__device__ __forceinline__ u32x4 bf_pack_b(float br, float bi) {
  const Split3 s = split3(br, bi);
  return (u32x4){s.h, s.m, s.l, s.h};
}

// the A operand of one lane (matrix row n, panel column k = g), both planes
struct BfA {
  u32x4 r1, r2, i1, i2;
};
__device__ __forceinline__ BfA bf_prep_a(float ar, float ai) {
  const Split3 s = split3(ar, ai);
  auto neg = [](unsigned u) { return u ^ 0x80000000u; };                  // [re | -im]
  auto swp = [](unsigned u) { return __builtin_amdgcn_alignbit(u, u, 16); };  // [im | re]
  BfA a;
  a.r1 = (u32x4){neg(s.h), neg(s.h), neg(s.h), neg(s.m)};
  a.r2 = (u32x4){neg(s.l), neg(s.m), neg(s.m), 0u};
  a.i1 = (u32x4){swp(s.h), swp(s.h), swp(s.h), swp(s.m)};
  a.i2 = (u32x4){swp(s.l), swp(s.m), swp(s.m), 0u};
  return a;
}
// (cr, ci) += a b, b the packed B operand of this lane (bf_pack_b)
__device__ __forceinline__ void bf_cmma(const BfA &a, u32x4 bp, mfma_traits<float>::v4 &cr,
                                        mfma_traits<float>::v4 &ci) {
  const bf16x8 b = __builtin_bit_cast(bf16x8, bp);
  cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a.r1), b, cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a.i1), b, ci, 0, 0, 0);
  cr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a.r2), b, cr, 0, 0, 0);
  ci = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a.i2), b, ci, 0, 0, 0);
}

template <typename T>
__device__ __forceinline__ typename mfma_traits<T>::c2 mkc(T x, T y) {
  typename mfma_traits<T>::c2 r;
  r.x = x;
  r.y = y;
  return r;
}

// One plane (real or imaginary) of a wave's accumulators: NT tiles of four
// elements.  FP32 with NT <= 8 keeps them in one 32-register vector, so that a
// tile picked by a wave-uniform runtime index is read by VGPR indexing (no
// branches); otherwise they are NT 4-vectors and the tile is picked by scalar
// branches.
template <typename T, int NT, bool BANK = (sizeof(T) == 4 && NT <= 8)>
struct AccPlane;

template <typename T, int NT>
struct AccPlane<T, NT, true> {
  static constexpr bool bank = true;
  static constexpr int ntiles = NT;
  typedef typename mfma_traits<T>::v4 v4;
  typedef T bank_t __attribute__((ext_vector_type(NT <= 4 ? 16 : 32)));
  bank_t b;
  __device__ __forceinline__ void zero() { b = (bank_t)(T)0; }
  // t compile-time (unrolled); j compile-time or wave-uniform
  __device__ __forceinline__ T get(int t, int j) const { return b[4 * t + j]; }
  __device__ __forceinline__ void set(int t, int j, T v) { b[4 * t + j] = v; }
  __device__ __forceinline__ v4 tile(int t) const {
    v4 r;
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = b[4 * t + j];
    return r;
  }
  __device__ __forceinline__ void set_tile(int t, v4 v) {
#pragma unroll
    for (int j = 0; j < 4; ++j) b[4 * t + j] = v[j];
  }
  __device__ __forceinline__ bank_t &regs(int) { return b; }
  // tile tq, a wave-uniform runtime index
  __device__ __forceinline__ v4 select(int tq) const {
    tq = __builtin_amdgcn_readfirstlane(tq);
    v4 r;
#pragma unroll
    for (int j = 0; j < 4; ++j) r[j] = b[4 * tq + j];
    return r;
  }
};

template <typename T, int NT>
struct AccPlane<T, NT, false> {
  static constexpr bool bank = false;
  static constexpr int ntiles = NT;
  typedef typename mfma_traits<T>::v4 v4;
  v4 v[NT];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int t = 0; t < NT; ++t) v[t] = (v4)(T)0;
  }
  // t and j compile-time (unrolled)
  __device__ __forceinline__ T get(int t, int j) const { return v[t][j]; }
  __device__ __forceinline__ void set(int t, int j, T x) { v[t][j] = x; }
  __device__ __forceinline__ v4 tile(int t) const { return v[t]; }
  __device__ __forceinline__ void set_tile(int t, v4 x) { v[t] = x; }
  __device__ __forceinline__ v4 &regs(int t) { return v[t]; }
  // tile tq, a wave-uniform runtime index.  The empty asm in each branch keeps
  // the compiler from folding the branches into a dynamically indexed array,
  // which would move the accumulators to scratch memory.
  __device__ __forceinline__ v4 select(int tq) const {
    tq = __builtin_amdgcn_readfirstlane(tq);
    v4 r = v[0];
#pragma unroll
    for (int k = 1; k < NT; ++k)
      if (tq == k) {
        r = v[k];
        __asm__ volatile("" : "+v"(r));
      }
    return r;
  }
};

}  // namespace ldga_bse
