// HBM streaming probe (no reference counterpart): the practical bandwidth
// of an access mix, read nin arrays and write nout arrays of n floats each,
// every element once, as the stencil kernels stream their per-cell arrays.
// bench.py --workload stream times it for the copy mix (1 read, 1 write)
// and the mixes of the velocity (10/3), stress (23/12, 20/12 with the
// texture rule) and fused (20/15) updates, so the stencil kernels' achieved
// bandwidth can be read against what the same number of streams reaches with
// no stencil, no halo and no arithmetic.
//
// Per-lane width: 4 bytes (256 bytes per wave per array, the stencil kernels'
// access width), 8 bytes (dwordx2) or 16 bytes (dwordx4).
#include "awp_common.h"
#include "../../include/awp_hip.h"

namespace {

constexpr int PROBE_MAX = 32;

struct ProbePtrs {
  const float* in[PROBE_MAX];
  float* out[PROBE_MAX];
};

template <int W>
__global__ __launch_bounds__(256) void probe_kernel(const ProbePtrs P, int nin, int nout,
                                                    long long n) {
  if (W == 2) {
    const long long nv = n / 2;
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < nv; e += (long long)gridDim.x * 256) {
      float2 acc = make_float2(0.f, 0.f);
      for (int a = 0; a < nin; ++a) {
        const float2 x = reinterpret_cast<const float2*>(P.in[a])[e];
        acc.x += x.x; acc.y += x.y;
      }
      for (int a = 0; a < nout; ++a) {
        const float s = (float)a;
        reinterpret_cast<float2*>(P.out[a])[e] = make_float2(acc.x + s, acc.y + s);
      }
    }
  } else if (W == 4) {
    const long long nv = n / 4;
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < nv; e += (long long)gridDim.x * 256) {
      float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
      for (int a = 0; a < nin; ++a) {
        const float4 x = reinterpret_cast<const float4*>(P.in[a])[e];
        acc.x += x.x; acc.y += x.y; acc.z += x.z; acc.w += x.w;
      }
      for (int a = 0; a < nout; ++a) {
        const float s = (float)a;
        reinterpret_cast<float4*>(P.out[a])[e] =
            make_float4(acc.x + s, acc.y + s, acc.z + s, acc.w + s);
      }
    }
  } else {
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
      float acc = 0.f;
      for (int a = 0; a < nin; ++a) acc += P.in[a][e];
      for (int a = 0; a < nout; ++a) P.out[a][e] = acc + (float)a;
    }
  }
}

}  // namespace

extern "C" int awp_stream_probe(const float* const* in, int nin, float* const* out, int nout,
                                long long n, int vec4, int blocks, hipStream_t St) {
  const char* W = "awp_stream_probe";
  // vec4: 0 = 4 bytes per lane and array, 1 = 16 bytes, 2 = 8 bytes
  const int width = vec4 == 1 ? 4 : vec4 == 2 ? 2 : 1;
  if (nin < 0 || nout < 0 || nin > PROBE_MAX || nout > PROBE_MAX || n <= 0 || blocks < 1 ||
      vec4 < 0 || vec4 > 2 || n % width != 0) {
    awp_set_error(AWP_ERR_ARG, "%s: nin=%d nout=%d n=%lld blocks=%d vec4=%d out of range", W,
                  nin, nout, n, blocks, vec4);
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
  if (width == 4)
    hipLaunchKernelGGL(probe_kernel<4>, dim3(blocks), dim3(256), 0, St, P, nin, nout, n);
  else if (width == 2)
    hipLaunchKernelGGL(probe_kernel<2>, dim3(blocks), dim3(256), 0, St, P, nin, nout, n);
  else
    hipLaunchKernelGGL(probe_kernel<1>, dim3(blocks), dim3(256), 0, St, P, nin, nout, n);
  return awp_check_launch(W) ? 0 : AWP_ERR_HIP;
}
