/*
 * tck_common.h -- device helpers shared by the HIP kernels (gfx950, wave64).
 */
#ifndef TCK_COMMON_H
#define TCK_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../tc_shared.h"

#define TCK_WAVE 64
#ifndef TCK_NT_STORE
#define TCK_NT_STORE 1     /* streamed results are stored non-temporally (not kept in cache):
                            * 3 % off the FP64 step, every stage but the predictor gaining */
#endif
#define TCK_NXCD 8        /* XCDs: blocks are dispatched to them round-robin */

/* launch-bounds hint used by every kernel below; keeps register allocation predictable */
#define TCK_BLOCK 256

__device__ __forceinline__ long long tck_idx(const tc_geom_t & g, int i, int j, int k){
  return (long long)(i + TC_XOFF) + (long long)g.px * ((long long)j + (long long)g.py * (long long)k);
}

/* Store an interior value (1 <= j <= ny, 1 <= k <= nz) together with its periodic halo
 * images (src/halo.c:18-142 fills j = 0 from ny, j = ny+1 from 1, then the z planes over the
 * full (x, y) extent, so corners are images of images).  With xcopy set the value is also
 * written to the x wall slot next to it (Neumann copy of src/fluid/boundary/{p,psi}.c).
 * MAIN false: the images and copies only, for a caller that has stored the value itself. */
typedef double tck_d2v __attribute__((ext_vector_type(2)));
typedef float tck_f2v __attribute__((ext_vector_type(2)));
#ifndef TCK_NT_LOAD
#define TCK_NT_LOAD 1      /* operands read once are loaded non-temporally */
#endif
/* load of a streamed operand (read once by this kernel) */
template <typename T>
__device__ __forceinline__ T tck_ld(const T * p){
#if TCK_NT_LOAD
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
__device__ __forceinline__ double2 tck_ld(const double2 * p){
#if TCK_NT_LOAD
  const tck_d2v w = __builtin_nontemporal_load((const tck_d2v *)p);
  return make_double2(w.x, w.y);
#else
  return *p;
#endif
}
__device__ __forceinline__ float2 tck_ld(const float2 * p){
#if TCK_NT_LOAD
  const tck_f2v w = __builtin_nontemporal_load((const tck_f2v *)p);
  return make_float2(w.x, w.y);
#else
  return *p;
#endif
}

/* store of a streamed result (read again only by a later kernel) */
template <typename T>
__device__ __forceinline__ void tck_st(T * p, T v){
#if TCK_NT_STORE
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}
__device__ __forceinline__ void tck_st(double2 * p, double2 v){
#if TCK_NT_STORE
  const tck_d2v w = {v.x, v.y};
  __builtin_nontemporal_store(w, (tck_d2v *)p);
#else
  *p = v;
#endif
}
__device__ __forceinline__ void tck_st(float2 * p, float2 v){
#if TCK_NT_STORE
  const tck_f2v w = {v.x, v.y};
  __builtin_nontemporal_store(w, (tck_f2v *)p);
#else
  *p = v;
#endif
}

template <typename T, bool MAIN = true>
__device__ __forceinline__ void tck_store_halo(T * a, const tc_geom_t & g, int i, int j, int k, T v, bool xcopy){
  const int nj = (j == 1) ? g.ny + 1 : (j == g.ny) ? 0 : -1;
  const int nk = (k == 1) ? g.nz + 1 : (k == g.nz) ? 0 : -1;
  const int ni = !xcopy ? -1 : (i == 1) ? 0 : (i == g.nx) ? g.nx + 1 : -1;
  #pragma unroll
  for(int a0 = 0; a0 < 2; a0++){
    const int ii = a0 ? ni : i;
    if(ii < 0) continue;
    if(MAIN || a0) tck_st(&a[tck_idx(g, ii, j, k)], v);
    if(nj >= 0) a[tck_idx(g, ii, nj, k)] = v;
    if(nk >= 0){
      a[tck_idx(g, ii, j, nk)] = v;
      if(nj >= 0) a[tck_idx(g, ii, nj, nk)] = v;
    }
  }
}

/* a value that is the same in every lane, as a scalar (so that a buffer resource built from it
 * needs no per-lane loop) */
__device__ __forceinline__ long long tck_uni64(long long x){
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(unsigned long long)x);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)x >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

/* Raw buffer access: a resource (base and size in bytes, in scalar registers, uniform across
 * the wave) plus a 32-bit per-lane byte offset and a uniform one, so that an access costs one
 * vector register of address.  Word 3 of the resource is the untyped-access value of gfx9
 * parts.  Accesses at or past the size read 0 and write nothing. */
#define TCK_BUF_WORD3 0x00020000
typedef unsigned int tck_u2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tck_rsrc(const void * base, unsigned bytes){
  return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)bytes, TCK_BUF_WORD3);
}
/* the same for a base and size that are the same in every lane but that the compiler may
 * not know to be (read from the first lane, so that the resource is built in scalar
 * registers and an access through it needs no per-lane loop) */
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tck_rsrc_u(const void * base, unsigned bytes){
  const unsigned long long a = (unsigned long long)base;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  return tck_rsrc((const void *)(((unsigned long long)hi << 32) | lo), __builtin_amdgcn_readfirstlane(bytes));
}
template <typename T> __device__ __forceinline__ T tck_bld(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so);
template <> __device__ __forceinline__ double tck_bld<double>(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so){
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}
template <> __device__ __forceinline__ float tck_bld<float>(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so){
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}
/* store of a streamed result (cache policy as tck_st) */
__device__ __forceinline__ void tck_bst(__amdgpu_buffer_rsrc_t r, unsigned vo, double v){
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(tck_u2v, v), r, vo, 0, TCK_NT_STORE ? 2 : 0);
}
/* load of a streamed operand (cache policy as tck_ld) and store of a streamed result (as
 * tck_st), FP64 or FP32, at byte offset vo + so of resource r */
template <typename T> __device__ __forceinline__ T tck_bldn(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so);
template <> __device__ __forceinline__ double tck_bldn<double>(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so){
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, TCK_NT_LOAD ? 2 : 0));
}
template <> __device__ __forceinline__ float tck_bldn<float>(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so){
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, TCK_NT_LOAD ? 2 : 0));
}
template <typename T> __device__ __forceinline__ void tck_bstn(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, T v);
template <> __device__ __forceinline__ void tck_bstn<double>(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, double v){
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(tck_u2v, v), r, vo, so, TCK_NT_STORE ? 2 : 0);
}
template <> __device__ __forceinline__ void tck_bstn<float>(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, float v){
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, vo, so, TCK_NT_STORE ? 2 : 0);
}
/* complex FP32 value: load of a streamed operand and store of a streamed result (cache
 * policies as tck_ld and tck_st) at byte offset vo + so of resource r */
__device__ __forceinline__ float2 tck_bldf2(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so){
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, TCK_NT_LOAD ? 2 : 0));
}
__device__ __forceinline__ void tck_bstf2(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so, float2 v){
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(tck_u2v, v), r, vo, so, TCK_NT_STORE ? 2 : 0);
}
/* complex FP64 value: load of a streamed operand (cache policy as tck_ld) at byte offset
 * vo + so of resource r */
typedef unsigned int tck_u4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ double2 tck_bld2(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so){
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, TCK_NT_LOAD ? 2 : 0));
}
/* the complex load of either precision */
template <typename V> __device__ __forceinline__ V tck_bldc(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so);
template <> __device__ __forceinline__ double2 tck_bldc<double2>(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so){ return tck_bld2(r, vo, so); }
template <> __device__ __forceinline__ float2 tck_bldc<float2>(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so){ return tck_bldf2(r, vo, so); }
/* 16-byte store of zeros at byte offset vo of resource r; with a zero-size resource it writes
 * nothing but is counted among the outstanding memory operations like any store (distinct
 * offsets keep the compiler from merging such stores) */
__device__ __forceinline__ void tck_bst0(__amdgpu_buffer_rsrc_t r, unsigned vo){
  const tck_u4v z = {0u, 0u, 0u, 0u};
  __builtin_amdgcn_raw_buffer_store_b128(z, r, vo, 0u, 0);
}
/* store of a halo image (plain, as tck_store_halo's) */
__device__ __forceinline__ void tck_bsti(__amdgpu_buffer_rsrc_t r, unsigned vo, double v){
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(tck_u2v, v), r, vo, 0, 0);
}

/* Store for the second neighbour case ny == 1 or nz == 1 is not supported (ny, nz >= 3 is
 * checked on the host). */

__device__ __forceinline__ double tck_shfl_up(double v, int off){
  return __shfl_up(v, off, TCK_WAVE);
}
__device__ __forceinline__ double tck_shfl_down(double v, int off){
  return __shfl_down(v, off, TCK_WAVE);
}
__device__ __forceinline__ double tck_shfl(double v, int lane){
  return __shfl(v, lane, TCK_WAVE);
}

/* order-preserving maps between non-negative doubles and their bit patterns (atomic min/max) */
__device__ __forceinline__ unsigned long long tck_bits(double v){
  return (unsigned long long)__double_as_longlong(v);
}

__device__ __forceinline__ double tck_wave_min(double v){
  for(int off = TCK_WAVE / 2; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, TCK_WAVE));
  return v;
}
__device__ __forceinline__ double tck_wave_max(double v){
  for(int off = TCK_WAVE / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, TCK_WAVE));
  return v;
}
__device__ __forceinline__ double tck_wave_sum(double v){
  for(int off = TCK_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, TCK_WAVE);
  return v;
}

#define TCK_CHECK_LAUNCH() do { hipError_t e_ = hipGetLastError(); if(e_ != hipSuccess) return (int)e_; } while(0)

#endif
