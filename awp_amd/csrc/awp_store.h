// Storage formats of the wavefield arrays.
//
// FP32: the reference's storage.  BF16 / FP16: 16-bit storage of the nine
// wavefield components and six memory variables with one power-of-two scale
// per field; every value is decoded to binary32 on load (q * s), all
// arithmetic is binary32, and results are encoded on store by round to
// nearest even of x / s.  With power-of-two scales, decode and the division
// are exact, so a stored value is x rounded to the 16-bit format of x / s.
// The medium stays binary32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

enum { AWP_FMT_F32 = 0, AWP_FMT_BF16 = 1, AWP_FMT_F16 = 2 };

// Field slots of the per-field scale table.
enum {
  AWP_SU = 0, AWP_SV, AWP_SW, AWP_SXX, AWP_SYY, AWP_SZZ, AWP_SXY, AWP_SXZ, AWP_SYZ,
  AWP_SR1, AWP_SR2, AWP_SR3, AWP_SR4, AWP_SR5, AWP_SR6, AWP_NSCALE
};

struct WaveScales {
  float s[AWP_NSCALE];    // decode: x = q * s
  float inv[AWP_NSCALE];  // encode: q = rne(x * inv)
};

__host__ __device__ __forceinline__ float awp_bf16_to_f32(uint16_t q) {
  return __builtin_bit_cast(float, (uint32_t)q << 16);
}

// binary32 <-> binary16 with the hardware conversions (v_cvt_f16_f32 rounds
// to nearest even and keeps binary16 subnormals, as the binary16 denormal
// mode is on).
__device__ __forceinline__ uint16_t awp_f32_to_f16(float x) {
  return __builtin_bit_cast(uint16_t, (_Float16)x);
}

__device__ __forceinline__ float awp_f16_to_f32(uint16_t q) {
  return (float)__builtin_bit_cast(_Float16, q);
}

// Element access by index (ld/st) or by a uniform row pointer plus a 32-bit
// byte offset (ldb/stb, which keeps the row base in scalar registers).
template <typename T>
__device__ __forceinline__ T awp_rawb(const T* row, unsigned kbyte) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(row) + kbyte);
}
template <typename T>
__device__ __forceinline__ void awp_rawb_st(T* row, unsigned kbyte, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(row) + kbyte) = v;
}

struct StF32 {
  typedef float T;
  static constexpr int fmt = AWP_FMT_F32;
  __device__ static __forceinline__ float dec(T q, float) { return q; }
  __device__ static __forceinline__ T enc(float x, float) { return x; }
  __device__ static __forceinline__ float ld(const T* a, long long i, float s) { return a[i]; }
  __device__ static __forceinline__ void st(T* a, long long i, float x, float) { a[i] = x; }
  __device__ static __forceinline__ float ldb(const T* r, unsigned kb, float) {
    return awp_rawb(r, kb);
  }
  __device__ static __forceinline__ void stb(T* r, unsigned kb, float x, float) {
    awp_rawb_st(r, kb, x);
  }
};

template <class Codec>
struct St16 {
  typedef uint16_t T;
  static constexpr int fmt = Codec::fmt;
  __device__ static __forceinline__ float dec(T q, float s) {
    return Codec::scale(Codec::to_f32(q), s);
  }
  __device__ static __forceinline__ T enc(float x, float inv) {
    return Codec::from_f32(Codec::scale(x, inv));
  }
  __device__ static __forceinline__ float ld(const T* a, long long i, float s) {
    return dec(a[i], s);
  }
  __device__ static __forceinline__ void st(T* a, long long i, float x, float inv) {
    a[i] = enc(x, inv);
  }
  __device__ static __forceinline__ float ldb(const T* r, unsigned kb, float s) {
    return dec(awp_rawb(r, kb), s);
  }
  __device__ static __forceinline__ void stb(T* r, unsigned kb, float x, float inv) {
    awp_rawb_st(r, kb, enc(x, inv));
  }
};

// binary32 -> bfloat16 is gfx950's v_cvt_pk_bf16_f32: round to nearest even,
// overflow to infinity, NaN stays NaN.  Encoder inputs are products x * inv
// computed with binary32 denormals flushed, so they are never subnormal.
struct CodecBF16 {
  static constexpr int fmt = AWP_FMT_BF16;
  __device__ static __forceinline__ float scale(float x, float s) { return x * s; }
  __device__ static __forceinline__ float to_f32(uint16_t q) { return awp_bf16_to_f32(q); }
  __device__ static __forceinline__ uint16_t from_f32(float x) {
    return __builtin_bit_cast(uint16_t, (__bf16)x);
  }
};
// binary16 scales by v_ldexp_f32 with the exponent of the power-of-two s
// (the same value as x * s, flushes included).  A multiply beside the
// conversion would be fused into v_fma_mix*, which follows the binary32
// denormal mode and so loses binary16 subnormals; ldexp has no mix form.
struct CodecF16 {
  static constexpr int fmt = AWP_FMT_F16;
  __device__ static __forceinline__ float scale(float x, float s) {
    return __builtin_ldexpf(x, __builtin_amdgcn_frexp_expf(s) - 1);
  }
  __device__ static __forceinline__ float to_f32(uint16_t q) { return awp_f16_to_f32(q); }
  __device__ static __forceinline__ uint16_t from_f32(float x) { return awp_f32_to_f16(x); }
};
typedef St16<CodecBF16> StBF16;
typedef St16<CodecF16> StF16;

// Host side: validate a 16-bit format code and a scale table (15 finite
// powers of two); fills ws.  Returns false and records an error otherwise.
bool awp_scales_from_host(const char* where, int fmt, const float* scale, WaveScales* ws);
