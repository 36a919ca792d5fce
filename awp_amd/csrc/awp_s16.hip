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

template <class ST>
__global__ __launch_bounds__(256) void rescale_kernel(uint16_t* __restrict__ q, long long n,
                                                      float factor) {
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    q[i] = ST::enc(ST::dec(q[i], 1.0f) * factor, 1.0f);
}

// Largest |decoded value| (unit scale) into *out: per-block maximum over a
// grid-stride range, then one vector atomic max per block on the bit
// pattern (non-negative binary32 values order as unsigned integers).
template <class ST>
__global__ __launch_bounds__(256) void absmax_kernel(const uint16_t* __restrict__ q,
                                                     long long n, unsigned* __restrict__ out) {
  __shared__ unsigned red[4];
  unsigned m = 0u;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const unsigned b = __float_as_uint(fabsf(ST::dec(q[i], 1.0f)));
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
    atomicMax(out, r);
  }
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
  const float factor = ldexpf(1.0f, shift);
  if (fmt == AWP_FMT_BF16)
    hipLaunchKernelGGL(rescale_kernel<StBF16>, dim3(grid_for(n)), dim3(256), 0, St, q, n, factor);
  else
    hipLaunchKernelGGL(rescale_kernel<StF16>, dim3(grid_for(n)), dim3(256), 0, St, q, n, factor);
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
    hipLaunchKernelGGL(absmax_kernel<StBF16>, dim3(grid_for(n)), dim3(256), 0, St, q, n,
                       (unsigned*)out);
  else
    hipLaunchKernelGGL(absmax_kernel<StF16>, dim3(grid_for(n)), dim3(256), 0, St, q, n,
                       (unsigned*)out);
  awp_check_launch(W);
}

}  // extern "C"
