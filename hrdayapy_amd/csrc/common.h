// Shared helpers for the hrdaya HIP library: error reporting, math traits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

namespace hp {

// Last error message of the calling thread; read through hp_last_error().
void set_error(const char* fmt, ...);

// Checks the launch status of the most recent kernel on this thread.
int check_launch(const char* what);

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// Raw buffer access: a 128-bit resource over `bytes` bytes of one array, and
// loads/stores at a 32-bit element index.  The address is the scalar base plus
// the lane's 32-bit offset, so a run of per-variable loads costs no 64-bit
// vector address arithmetic; loads past `bytes` return 0 and such stores are
// dropped by the hardware.  Word 3 is the gfx9-family raw (untyped) format.
// ---------------------------------------------------------------------------
typedef __amdgpu_buffer_rsrc_t buf_t;
__device__ __forceinline__ buf_t buf_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
template <typename T> __device__ __forceinline__ T buf_ld(buf_t r, uint32_t idx);
template <> __device__ __forceinline__ float buf_ld<float>(buf_t r, uint32_t idx) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, idx * 4u, 0, 0));
}
template <> __device__ __forceinline__ double buf_ld<double>(buf_t r, uint32_t idx) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, idx * 8u, 0, 0));
}
__device__ __forceinline__ void buf_st(buf_t r, uint32_t idx, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, idx * 4u, 0, 0);
}
__device__ __forceinline__ void buf_st(buf_t r, uint32_t idx, double v) {
  typedef unsigned u2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, idx * 8u, 0, 0);
}
__device__ __forceinline__ uint32_t buf_ld_u8(buf_t r, uint32_t idx) {
  return __builtin_amdgcn_raw_buffer_load_b8(r, idx, 0, 0);
}

// ---------------------------------------------------------------------------
// Compact cell-state index of a brick-layout slot.  groups[g] = {myocardial
// slots before slot 64*g, 0, myocardial mask bits 0-31, bits 32-63}; the
// state entry of slot s is base + popcount(mask below s & 63), or the
// out-of-range index 0x1FFFFFFF for a slot without myocardium.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t state_index(const int32_t* __restrict__ groups, uint32_t s) {
  const int4 r = reinterpret_cast<const int4*>(groups)[s >> 6];
  const uint64_t m = (uint64_t)(uint32_t)r.z | ((uint64_t)(uint32_t)r.w << 32);
  const uint32_t b = s & 63u;
  const uint64_t below = m & ((1ull << b) - 1ull);
  return ((m >> b) & 1ull) ? (uint32_t)r.x + (uint32_t)__popcll(below) : 0x1FFFFFFFu;
}

// The same for a wave whose 64 lanes hold the 64 slots of group gidx in lane
// order (gidx wave-uniform: one scalar load, a bit test and mbcnt).
__device__ __forceinline__ uint32_t state_index_wave(const int32_t* __restrict__ groups,
                                                     uint32_t gidx) {
  const int4 r = reinterpret_cast<const int4*>(groups)[gidx];
  const uint32_t lo = (uint32_t)r.z, hi = (uint32_t)r.w;
  const uint32_t below = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0u));
  const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  const bool in = ((lane < 32 ? lo >> lane : hi >> (lane - 32)) & 1u) != 0;
  return in ? (uint32_t)r.x + below : 0x1FFFFFFFu;
}

// ---------------------------------------------------------------------------
// Arithmetic traits.  The FP32 path uses the hardware transcendental and
// reciprocal instructions (v_exp_f32 / v_log_f32 / v_rcp_f32 / v_sqrt_f32,
// ~1 ulp); the FP64 path uses the correctly rounded library routines so that
// it can be compared with the FP64 oracle at round-off level.  The host pass
// (used by the host entry points that check the cell algebra on a CPU) takes
// the C library routines in place of the FP32 hardware instructions.
// ---------------------------------------------------------------------------
#define HP_HD __host__ __device__ __forceinline__
template <typename T> struct M;

template <> struct M<float> {
#if defined(__HIP_DEVICE_COMPILE__)
  static HP_HD float ex(float x) { return __expf(x); }
  static HP_HD float ex2(float x) { return __builtin_amdgcn_exp2f(x); }
  static HP_HD float lg(float x) { return __logf(x); }
  static HP_HD float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  static HP_HD float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  static HP_HD float sq(float x) { return __builtin_amdgcn_sqrtf(x); }
  static HP_HD float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
#else
  static HP_HD float ex(float x) { return expf(x); }
  static HP_HD float ex2(float x) { return exp2f(x); }
  static HP_HD float lg(float x) { return logf(x); }
  static HP_HD float rcp(float x) { return 1.0f / x; }
  static HP_HD float dv(float a, float b) { return a / b; }
  static HP_HD float sq(float x) { return sqrtf(x); }
  static HP_HD float rsq(float x) { return 1.0f / sqrtf(x); }
#endif
  static HP_HD float mx(float a, float b) { return fmaxf(a, b); }
  static HP_HD float mn(float a, float b) { return fminf(a, b); }
  static HP_HD float ab(float a) { return fabsf(a); }
};

template <> struct M<double> {
  static HP_HD double ex(double x) { return exp(x); }
  static HP_HD double ex2(double x) { return exp2(x); }
  static HP_HD double lg(double x) { return log(x); }
  static HP_HD double rcp(double x) { return 1.0 / x; }
  static HP_HD double dv(double a, double b) { return a / b; }
  static HP_HD double sq(double x) { return sqrt(x); }
  static HP_HD double rsq(double x) { return 1.0 / sqrt(x); }
  static HP_HD double mx(double a, double b) { return fmax(a, b); }
  static HP_HD double mn(double a, double b) { return fmin(a, b); }
  static HP_HD double ab(double a) { return fabs(a); }
};

}  // namespace hp

#define HP_CHECK(expr)                                                              \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess) {                                                         \
      hp::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                    __LINE__);                                                      \
      return -(int)_e;                                                              \
    }                                                                               \
  } while (0)
