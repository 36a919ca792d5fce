// 16-bit wavefield storage: the scale table check, and the element-wise
// passes a driver needs around the 16-bit launchers (encode, decode, the
// largest magnitude of a field, exact power-of-two rescaling).
#include <math.h>

#include "awp_common.h"
#include "awp_store.h"
#include "../../include/awp_hip.h"

bool awp_scales_from_host(const char* where, int fmt, const float* scale, WaveScales* ws) {
  if (fmt != AWP_FMT_BF16 && fmt != AWP_FMT_F16) {
    awp_set_error(AWP_ERR_ARG, "%s: storage format %d is not 1 (bfloat16) or 2 (binary16)",
                  where, fmt);
    return false;
  }
  if (!scale) {
    awp_set_error(AWP_ERR_ARG, "%s: scale table is NULL", where);
    return false;
  }
  for (int f = 0; f < AWP_NSCALE; ++f) {
    int e;
    const float s = scale[f];
    if (!(s > 0.0f) || !isfinite(s) || frexpf(s, &e) != 0.5f || e < -120 || e > 120) {
      awp_set_error(AWP_ERR_ARG, "%s: scale[%d] = %g is not a power of two in [2^-121, 2^119]",
                    where, f, (double)s);
      return false;
    }
    ws->s[f] = s;
    ws->inv[f] = 1.0f / s;  // exact for a power of two
  }
  return true;
}

namespace {

template <class ST>
__global__ __launch_bounds__(256) void encode_kernel(const float* __restrict__ x,
                                                     uint16_t* __restrict__ q, long long n,
                                                     float inv) {
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    q[i] = ST::enc(x[i], inv);
}

template <class ST>
__global__ __launch_bounds__(256) void decode_kernel(const uint16_t* __restrict__ q,
                                                     float* __restrict__ x, long long n,
                                                     float s) {
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    x[i] = ST::dec(q[i], s);
}

// Eight 16-bit values per lane and 16-byte access where the array allows it
// (n8 > 0 only for a 16-byte aligned q); the remainder one value per lane.
// q <- q * 2^shift, rounded to the format: the scaling is v_ldexp_f32, exact
// like a multiply by 2^shift (flushes included) and never fused with the
// binary16 conversion into a v_fma_mix* (which would flush binary16
// subnormals).
template <class ST>
__device__ __forceinline__ uint16_t rescale1(uint16_t q, int shift) {
  return ST::enc(__builtin_ldexpf(ST::dec(q, 1.0f), shift), 1.0f);
}

template <class ST>
__global__ __launch_bounds__(256) void rescale_kernel(uint16_t* __restrict__ q, long long n,
                                                      long long n8, int shift) {
  const long long stride = (long long)gridDim.x * 256;
  uint4* q8 = reinterpret_cast<uint4*>(q);
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n8; i += stride) {
    uint4 v = q8[i];
    unsigned* w = &v.x;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const unsigned lo = rescale1<ST>((uint16_t)(w[c] & 0xffffu), shift);
      const unsigned hi = rescale1<ST>((uint16_t)(w[c] >> 16), shift);
      w[c] = lo | (hi << 16);
    }
    q8[i] = v;
  }
  for (long long i = n8 * 8 + blockIdx.x * 256ll + threadIdx.x; i < n; i += stride)
    q[i] = rescale1<ST>(q[i], shift);
}

// Largest |decoded value| (unit scale) into *out.  Both formats order
// non-negative values as unsigned integers (NaN patterns above infinity), so
// the pass takes the largest 15-bit magnitude pattern, decodes that one value,
// and issues one vector atomic max per block on its binary32 bit pattern.
template <class ST>
__global__ __launch_bounds__(256) void absmax_kernel(const uint16_t* __restrict__ q,
                                                     long long n, long long n8,
                                                     unsigned* __restrict__ out) {
  __shared__ unsigned red[4];
  const long long stride = (long long)gridDim.x * 256;
  const uint4* q8 = reinterpret_cast<const uint4*>(q);
  unsigned m = 0u;
  auto fold = [&m](const uint4 v) {
    const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const unsigned lo = w[c] & 0x7fffu, hi = (w[c] >> 16) & 0x7fffu;
      m = lo > m ? lo : m;
      m = hi > m ? hi : m;
    }
  };
  // four independent 16-byte loads in flight per lane
  long long i = blockIdx.x * 256ll + threadIdx.x;
  for (; i + 3 * stride < n8; i += 4 * stride) {
    const uint4 a = q8[i], b = q8[i + stride], c = q8[i + 2 * stride], d = q8[i + 3 * stride];
    fold(a);
    fold(b);
    fold(c);
    fold(d);
  }
  for (; i < n8; i += stride) fold(q8[i]);
  for (long long i = n8 * 8 + blockIdx.x * 256ll + threadIdx.x; i < n; i += stride) {
    const unsigned b = q[i] & 0x7fffu;
    m = b > m ? b : m;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned o = (unsigned)__shfl_xor((int)m, off, 64);
    m = o > m ? o : m;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned r = red[0];
    for (int w = 1; w < 4; ++w) r = red[w] > r ? red[w] : r;
    atomicMax(out, __float_as_uint(ST::dec((uint16_t)r, 1.0f)));
  }
}

// Blocks of the absmax pass: each block ends in one atomic on the same
// word, so the grid stays at four blocks per CU (1024) and every lane loops.
int absmax_grid(long long n) {
  const long long b = (n / 8 + 255) / 256;
  return (int)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}

long long vec8_count(const void* q, long long n) {
  return ((uintptr_t)q & 15u) == 0 ? n / 8 : 0;
}

int grid_for(long long n) {
  long long b = (n + 255) / 256;
  return (int)(b < 4096 ? (b < 1 ? 1 : b) : 4096);
}

bool fmt_ok(const char* W, int fmt) {
  if (fmt == AWP_FMT_BF16 || fmt == AWP_FMT_F16) return true;
  awp_set_error(AWP_ERR_ARG, "%s: storage format %d is not 1 (bfloat16) or 2 (binary16)", W, fmt);
  return false;
}

bool pow2_ok(const char* W, float s) {
  int e;
  if (s > 0.0f && isfinite(s) && frexpf(s, &e) == 0.5f && e >= -120 && e <= 120) return true;
  awp_set_error(AWP_ERR_ARG, "%s: scale %g is not a power of two in [2^-121, 2^119]", W,
                (double)s);
  return false;
}

}  // namespace

extern "C" {

void awp_s16_encode(const float* x, uint16_t* q, long long n, int fmt, float scale,
                    hipStream_t St) {
  const char* W = "awp_s16_encode";
  if (!fmt_ok(W, fmt) || !pow2_ok(W, scale) || n <= 0) return;
  if (fmt == AWP_FMT_BF16)
    hipLaunchKernelGGL(encode_kernel<StBF16>, dim3(grid_for(n)), dim3(256), 0, St, x, q, n,
                       1.0f / scale);
  else
    hipLaunchKernelGGL(encode_kernel<StF16>, dim3(grid_for(n)), dim3(256), 0, St, x, q, n,
                       1.0f / scale);
  awp_check_launch(W);
}

void awp_s16_decode(const uint16_t* q, float* x, long long n, int fmt, float scale,
                    hipStream_t St) {
  const char* W = "awp_s16_decode";
  if (!fmt_ok(W, fmt) || !pow2_ok(W, scale) || n <= 0) return;
  if (fmt == AWP_FMT_BF16)
    hipLaunchKernelGGL(decode_kernel<StBF16>, dim3(grid_for(n)), dim3(256), 0, St, q, x, n,
                       scale);
  else
    hipLaunchKernelGGL(decode_kernel<StF16>, dim3(grid_for(n)), dim3(256), 0, St, q, x, n,
                       scale);
  awp_check_launch(W);
}

void awp_s16_rescale(uint16_t* q, long long n, int fmt, int shift, hipStream_t St) {
  const char* W = "awp_s16_rescale";
  if (!fmt_ok(W, fmt) || n <= 0 || shift == 0) return;
  if (shift < -60 || shift > 60) {
    awp_set_error(AWP_ERR_ARG, "%s: shift %d outside [-60, 60]", W, shift);
    return;
  }
  if (fmt == AWP_FMT_BF16)
    hipLaunchKernelGGL(rescale_kernel<StBF16>, dim3(grid_for(n)), dim3(256), 0, St, q, n,
                       vec8_count(q, n), shift);
  else
    hipLaunchKernelGGL(rescale_kernel<StF16>, dim3(grid_for(n)), dim3(256), 0, St, q, n,
                       vec8_count(q, n), shift);
  awp_check_launch(W);
}

void awp_s16_absmax(const uint16_t* q, long long n, int fmt, float* out, hipStream_t St) {
  const char* W = "awp_s16_absmax";
  if (!fmt_ok(W, fmt)) return;
  if (!out) {
    awp_set_error(AWP_ERR_ARG, "%s: out is NULL", W);
    return;
  }
  if (hipMemsetAsync(out, 0, sizeof(float), St) != hipSuccess) {
    awp_set_error(AWP_ERR_HIP, "%s: hipMemsetAsync failed", W);
    return;
  }
  if (n <= 0) return;
  if (fmt == AWP_FMT_BF16)
    hipLaunchKernelGGL(absmax_kernel<StBF16>, dim3(absmax_grid(n)), dim3(256), 0, St, q, n,
                       vec8_count(q, n), (unsigned*)out);
  else
    hipLaunchKernelGGL(absmax_kernel<StF16>, dim3(absmax_grid(n)), dim3(256), 0, St, q, n,
                       vec8_count(q, n), (unsigned*)out);
  awp_check_launch(W);
}

}  // extern "C"
