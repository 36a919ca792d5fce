// HBM streaming probe (no reference counterpart): the practical bandwidth
// of an access mix, read nin arrays and write nout arrays of n floats each,
// every element once, as the stencil kernels stream their per-cell arrays.
// bench.py --workload stream times it for the copy mix (1 read, 1 write)
// and the mixes of the velocity (10/3), stress (23/12, 19/12 with the
// texture rule) and fused (20/15) updates, so the stencil kernels' achieved
// bandwidth can be read against what the same number of streams reaches with
// no stencil, no halo and no arithmetic.
//
// Per-lane width: 4 bytes (256 bytes per wave per array, the stencil kernels'
// access width), 8 bytes (dwordx2) or 16 bytes (dwordx4).
//
// Two load schedules:
// - serial: a runtime loop over the arrays, each load waited for before the
//   next is issued (one load in flight per wave);
// - batched: the loads of all nin arrays issued before the first is used
//   (nin loads in flight per wave, as a stencil kernel that loads a plane's
//   operands ahead of its arithmetic), nin a compile-time constant.
#include <array>
#include <utility>

#include "awp_common.h"
#include "../../include/awp_hip.h"

namespace {

constexpr int PROBE_MAX = 32;

struct ProbePtrs {
  const float* in[PROBE_MAX];
  float* out[PROBE_MAX];
};

template <int W> struct VecOf;
template <> struct VecOf<1> { using T = float; };
template <> struct VecOf<2> { using T = float2; };
template <> struct VecOf<4> { using T = float4; };

__device__ __forceinline__ float vadd(float a, float b) { return a + b; }
__device__ __forceinline__ float2 vadd(float2 a, float2 b) {
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float4 vadd(float4 a, float4 b) {
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
template <class V> __device__ __forceinline__ V vsplat(float s);
template <> __device__ __forceinline__ float vsplat<float>(float s) { return s; }
template <> __device__ __forceinline__ float2 vsplat<float2>(float s) { return make_float2(s, s); }
template <> __device__ __forceinline__ float4 vsplat<float4>(float s) {
  return make_float4(s, s, s, s);
}

// serial schedule: nv = n / W vectors per array
template <int W>
__global__ __launch_bounds__(256) void probe_kernel(const ProbePtrs P, int nin, int nout,
                                                    long long nv) {
  using V = typename VecOf<W>::T;
  for (long long e = blockIdx.x * 256ll + threadIdx.x; e < nv; e += (long long)gridDim.x * 256) {
    V acc = vsplat<V>(0.f);
    for (int a = 0; a < nin; ++a) acc = vadd(acc, reinterpret_cast<const V*>(P.in[a])[e]);
    for (int a = 0; a < nout; ++a)
      reinterpret_cast<V*>(P.out[a])[e] = vadd(acc, vsplat<V>((float)a));
  }
}

// batched schedule: every array's vector for element e is loaded before any
// is summed, so the wave has NIN loads in flight
template <int W, int NIN>
__global__ __launch_bounds__(256) void probe_batched_kernel(const ProbePtrs P, int nout,
                                                            long long nv) {
  using V = typename VecOf<W>::T;
  for (long long e = blockIdx.x * 256ll + threadIdx.x; e < nv; e += (long long)gridDim.x * 256) {
    V x[NIN];
#pragma unroll
    for (int a = 0; a < NIN; ++a) x[a] = reinterpret_cast<const V*>(P.in[a])[e];
    V acc = vsplat<V>(0.f);
#pragma unroll
    for (int a = 0; a < NIN; ++a) acc = vadd(acc, x[a]);
    for (int a = 0; a < nout; ++a)
      reinterpret_cast<V*>(P.out[a])[e] = vadd(acc, vsplat<V>((float)a));
  }
}

using BatchedFn = void (*)(const ProbePtrs, int, long long);

template <int W, int... I>
constexpr std::array<BatchedFn, sizeof...(I)> batched_table(std::integer_sequence<int, I...>) {
  return {{&probe_batched_kernel<W, I + 1>...}};
}

// batched kernels for nin = 1 .. PROBE_MAX
template <int W>
BatchedFn batched_for(int nin) {
  static const auto tab = batched_table<W>(std::make_integer_sequence<int, PROBE_MAX>{});
  return tab[nin - 1];
}

}  // namespace

extern "C" int awp_stream_probe(const float* const* in, int nin, float* const* out, int nout,
                                long long n, int vec4, int blocks, int batched, hipStream_t St) {
  const char* W = "awp_stream_probe";
  // vec4: 0 = 4 bytes per lane and array, 1 = 16 bytes, 2 = 8 bytes
  const int width = vec4 == 1 ? 4 : vec4 == 2 ? 2 : 1;
  if (nin < 0 || nout < 0 || nin > PROBE_MAX || nout > PROBE_MAX || n <= 0 || blocks < 1 ||
      vec4 < 0 || vec4 > 2 || n % width != 0 || (batched && nin < 1)) {
    awp_set_error(AWP_ERR_ARG, "%s: nin=%d nout=%d n=%lld blocks=%d vec4=%d batched=%d out of range",
                  W, nin, nout, n, blocks, vec4, batched);
    return AWP_ERR_ARG;
  }
  ProbePtrs P = {};
  for (int a = 0; a < nin; ++a) {
    P.in[a] = in[a];
    if (reinterpret_cast<uintptr_t>(in[a]) % (4 * width) != 0) {
      awp_set_error(AWP_ERR_ARG, "%s: input %d is not %d-byte aligned", W, a, 4 * width);
      return AWP_ERR_ARG;
    }
  }
  for (int a = 0; a < nout; ++a) {
    P.out[a] = out[a];
    if (reinterpret_cast<uintptr_t>(out[a]) % (4 * width) != 0) {
      awp_set_error(AWP_ERR_ARG, "%s: output %d is not %d-byte aligned", W, a, 4 * width);
      return AWP_ERR_ARG;
    }
  }
  const long long nv = n / width;
  if (batched) {
    const BatchedFn k = width == 4 ? batched_for<4>(nin)
                        : width == 2 ? batched_for<2>(nin)
                                     : batched_for<1>(nin);
    hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, St, P, nout, nv);
  } else if (width == 4) {
    hipLaunchKernelGGL(probe_kernel<4>, dim3(blocks), dim3(256), 0, St, P, nin, nout, nv);
  } else if (width == 2) {
    hipLaunchKernelGGL(probe_kernel<2>, dim3(blocks), dim3(256), 0, St, P, nin, nout, nv);
  } else {
    hipLaunchKernelGGL(probe_kernel<1>, dim3(blocks), dim3(256), 0, St, P, nin, nout, nv);
  }
  return awp_check_launch(W) ? 0 : AWP_ERR_HIP;
}
