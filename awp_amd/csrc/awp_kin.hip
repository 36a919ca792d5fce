// Kinematic finite-fault sources and prescribed-velocity sources (reference
// src/awp/kernel.cu: SetDeviceFilterParameters :119-123, addkinsrc_H/_cu
// :3677-3858 with compmt, brune, liu and lfilter, frcvel_H/_cu :1808-1929).
//
// One lane per subfault, as the reference: subfaults at distinct cells are
// independent (two at one cell race there as in the reference).  The moment
// tensor, the source time functions and the IIR filter keep the reference's
// mix of single and double precision; sinf/cosf/expf/powf are the device
// library's, so results agree with a CPU evaluation to a few ulp rather than
// bit for bit.
#include <math.h>

#include "awp_common.h"
#include "../../include/awp_hip.h"

#define AWP_MAXFILT 20  // include/awp/pmcl3d_cons.h:95
#define AWP_PI 3.14159265358979323846

namespace {

struct KinFilter {
  int order;
  double b[AWP_MAXFILT], a[AWP_MAXFILT];
};
KinFilter g_filter = {0, {0.0}, {0.0}};

struct KinLaunch {
  int step, dim, npsrc, yline;
  long long slice;
  float DT, vtst;
};

// compmt: the double-couple moment tensor of strike, dip and rake (radians).
// Order of m: xx, yy, zz, xz, yz, xy.
__device__ void moment_tensor(float str, float dip, float rake, float* m) {
  const float sd = sinf(dip), cd = cosf(dip), sr = sinf(rake), cr = cosf(rake);
  const float ss = sinf(str), cs = cosf(str);
  const float s2s = sinf((float)(2.0 * str)), c2s = cosf((float)(2.0 * str));
  const float s2d = sinf((float)(2.0 * dip)), c2d = cosf((float)(2.0 * dip));
  m[1] = -(sd * cr * s2s + s2d * sr * ss * ss);
  m[5] = (float)((double)(sd * cr * c2s) + 0.5 * (double)(s2d * sr * s2s));
  m[4] = cd * cr * cs + c2d * sr * ss;
  m[0] = sd * cr * s2s - s2d * sr * cs * cs;
  m[3] = cd * cr * ss - c2d * sr * cs;
  m[2] = s2d * sr;
}

// brune: omega^2 t exp(-omega t), omega = 2 pi freq.
__device__ float brune(float freq, float t) {
  const float omega = (float)((double)freq * AWP_PI * 2.0);
  if (t > 0.0f) return powf(omega, 2.0f) * t * expf(-omega * t);
  return 0.0f;
}

// liu: Liu et al. (2006) with rise time tau.
__device__ float liu(float tau, float t) {
  const float tau1 = (float)(0.13 * (double)tau);
  const float tau2 = tau - tau1;
  const float CN = (float)(AWP_PI / (1.4 * AWP_PI * (double)tau1 + 1.2 * (double)tau1 +
                                     0.3 * AWP_PI * (double)tau2));
  if (t < tau1)
    return (float)((double)CN * (0.7 - 0.7 * (double)cosf((float)(AWP_PI * t / tau1)) +
                                 0.6 * (double)sinf((float)(0.5 * AWP_PI * t / tau1))));
  if (t < 2 * tau1)
    return (float)((double)CN * (1.0 - 0.7 * (double)cosf((float)(AWP_PI * t / tau1)) +
                                 0.3 * (double)cosf((float)(AWP_PI * (t - tau1) / tau2))));
  if (t < tau)
    return (float)((double)CN * (0.3 + 0.3 * (double)cosf((float)(AWP_PI * (t - tau1) / tau2))));
  return 0.0f;
}

// lfilter: one sample of a direct-form-II-transposed IIR filter; d holds
// order + 1 state values.
__device__ double iir_step(const KinFilter& F, double x, double* d) {
  const double y = (F.b[0] * x + d[0]) / F.a[0];
  for (int n = 0; n < F.order - 1; ++n) d[n] = F.b[n + 1] * x - F.a[n + 1] * y + d[n + 1];
  d[F.order - 1] = F.b[F.order] * x - F.a[F.order] * y;
  return y;
}

__global__ __launch_bounds__(256) void addkinsrc_kernel(
    const int* __restrict__ psrc, const float* __restrict__ mu, const float* __restrict__ axx,
    const float* __restrict__ ayy, const float* __restrict__ azz, const float* __restrict__ axz,
    const float* __restrict__ ayz, float* xx, float* yy, float* zz, float* xy, float* yz,
    float* xz, float* mom, double* filt_d, const KinLaunch A, const KinFilter F) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= A.npsrc) return;
  const int r = 2 * j;  // two parameter rows per subfault
  const int stf_type = (int)axx[r];
  const float slip = ayy[r], ruptime = azz[r], risetime = axz[r];
  const float atime = A.step * A.DT;
  if (!(atime > ruptime)) return;
  const float area = ayz[r];
  const float strike = (float)((double)axx[r + 1] / 180.0 * AWP_PI);
  const float dip = (float)((double)ayy[r + 1] / 180.0 * AWP_PI);
  const float rake = (float)((double)azz[r + 1] / 180.0 * AWP_PI);
  float m[6];
  moment_tensor(strike, dip, rake, m);
  float stf;
  if (stf_type == 1)
    stf = brune((float)(1.0 / (double)risetime), atime - ruptime);
  else if (stf_type == 2)
    stf = liu(risetime, atime - ruptime);
  else
    stf = 0.0f;
  if (F.order > 0) stf = (float)iir_step(F, (double)stf, filt_d + (long long)j * (F.order + 1));
  const int idx = psrc[j * A.dim] + 1 + AWP_NGSL;
  const int idy = psrc[j * A.dim + 1] + 1 + AWP_NGSL;
  const int idz = psrc[j * A.dim + 2] + AWP_ALIGN - 1;
  const long long pos = (long long)idx * A.slice + (long long)idy * A.yline + idz;
  stf = stf * (slip * area / mu[pos]);
  mom[j] = mom[j] + stf * A.DT;
  stf = stf * A.vtst;
  xx[pos] = xx[pos] - stf * m[0];
  yy[pos] = yy[pos] - stf * m[1];
  zz[pos] = zz[pos] - stf * m[2];
  xz[pos] = xz[pos] - stf * m[3];
  yz[pos] = yz[pos] - stf * m[4];
  xy[pos] = xy[pos] - stf * m[5];
}

struct FrcLaunch {
  int step, read_step, dim, npsrc, tskp, xmin, xmax, yline;
  long long slice;
  float DT;
};

__global__ __launch_bounds__(256) void frcvel_kernel(const int* __restrict__ psrc, float* axx,
                                                     float* ayy, float* azz, float* u1, float* v1,
                                                     float* w1, const FrcLaunch A) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= A.npsrc) return;
  const int i = A.step - 1;
  const int tskp = A.tskp, np = A.npsrc;
  int i0 = (int)((float)tskp * floorf((float)(i + 1.0) / (float)tskp));
  int i1 = (int)((float)tskp * ceilf((float)(i + 1.0) / (float)tskp));
  const float pfact = (float)(i + 1 - i0) / (float)tskp;  // linear interpolation
  i0 /= tskp;
  i1 /= tskp;
  const int idx = psrc[j * A.dim] + 1 + AWP_NGSL;
  const int idy = psrc[j * A.dim + 1] + 1 + AWP_NGSL;
  const int idz = psrc[j * A.dim + 2] + AWP_ALIGN - 1;
  const long long pos = (long long)idx * A.slice + (long long)idy * A.yline + idz;
  const bool abvmin = A.xmin == -1 || idx >= A.xmin;
  const bool blwmax = A.xmax == -1 || idx <= A.xmax;
  if (!(abvmin && blwmax)) return;
  if (i < A.read_step * tskp) {
    const float up = axx[i0 * np + j], vp = ayy[i0 * np + j], wp = azz[i0 * np + j];
    const float un = axx[i1 * np + j], vn = ayy[i1 * np + j], wn = azz[i1 * np + j];
    const float ui = up + (un - up) * pfact;
    const float vi = vp + (vn - vp) * pfact;
    const float wi = wp + (wn - wp) * pfact;
    if (i == A.read_step * tskp - 1) {
      // the last prescribed sample moves to the front; the force that holds
      // the last velocity replaces it
      axx[j] = axx[i1 * np + j];
      ayy[j] = ayy[i1 * np + j];
      azz[j] = azz[i1 * np + j];
      axx[i1 * np + j] = (un - u1[pos]) / A.DT;
      ayy[i1 * np + j] = (vn - v1[pos]) / A.DT;
      azz[i1 * np + j] = (wn - w1[pos]) / A.DT;
    }
    u1[pos] = ui;
    v1[pos] = vi;
    w1[pos] = wi;
  } else {
    u1[pos] = u1[pos] + axx[A.read_step * np + j] * A.DT;
    v1[pos] = v1[pos] + ayy[A.read_step * np + j] * A.DT;
    w1[pos] = w1[pos] + azz[A.read_step * np + j] * A.DT;
  }
}

}  // namespace

extern "C" {

void SetDeviceFilterParameters(int filtorder, double* srcfilt_b, double* srcfilt_a) {
  if (filtorder < 0 || filtorder > AWP_MAXFILT - 1 ||
      (filtorder > 0 && (!srcfilt_b || !srcfilt_a))) {
    awp_set_error(AWP_ERR_ARG, "SetDeviceFilterParameters: order %d outside [0, %d]", filtorder,
                  AWP_MAXFILT - 1);
    return;
  }
  g_filter.order = filtorder;
  for (int k = 0; k < AWP_MAXFILT; ++k) {
    g_filter.b[k] = (filtorder > 0 && k <= filtorder) ? srcfilt_b[k] : 0.0;
    g_filter.a[k] = (filtorder > 0 && k <= filtorder) ? srcfilt_a[k] : 0.0;
  }
}

void addkinsrc_H(int i, int dim, int* psrc, int npsrc, hipStream_t St, float* mu, float* axx,
                 float* ayy, float* azz, float* axz, float* ayz, float* axy, float* xx, float* yy,
                 float* zz, float* xy, float* yz, float* xz, float* mom, double* srcfilt_d,
                 int d_i) {
  (void)axy;  // the reference passes it; the parameter rows use five arrays
  const char* W = "addkinsrc_H";
  if (!awp_grid_ok(d_i, W)) return;
  if (npsrc <= 0) return;
  if (dim < 3) {
    awp_set_error(AWP_ERR_ARG, "%s: dim=%d < 3", W, dim);
    return;
  }
  if (g_filter.order > 0 && !srcfilt_d) {
    awp_set_error(AWP_ERR_ARG, "%s: a filter of order %d needs its state array", W,
                  g_filter.order);
    return;
  }
  const AwpGrid& g = awp_state().grid[d_i];
  KinLaunch A;
  A.step = i;
  A.dim = dim;
  A.npsrc = npsrc;
  A.yline = g.yline;
  A.slice = g.slice;
  A.DT = g.DT;
  A.vtst = g.DT / (g.DH * g.DH * g.DH);  // kernel.cu:3828
  const int block = npsrc < 256 ? npsrc : 256;
  hipLaunchKernelGGL(addkinsrc_kernel, dim3(awp_cdiv(npsrc, block)), dim3(block), 0, St, psrc, mu,
                     axx, ayy, azz, axz, ayz, xx, yy, zz, xy, yz, xz, mom, srcfilt_d, A,
                     g_filter);
  awp_check_launch(W);
}

void frcvel_H(int i, int READ_STEP, int dim, int* psrc, int npsrc, int tskp, hipStream_t St,
              float* axx, float* ayy, float* azz, float* axz, float* ayz, float* axy, float* u1,
              float* v1, float* w1, int ymin, int ymax, int d_i) {
  (void)axz;
  (void)ayz;
  (void)axy;
  const char* W = "frcvel_H";
  if (!awp_grid_ok(d_i, W)) return;
  if (npsrc <= 0) return;
  if (dim < 3 || tskp < 1 || READ_STEP < 1 || i < 1) {
    awp_set_error(AWP_ERR_ARG, "%s: dim=%d, tskp=%d, READ_STEP=%d, step=%d", W, dim, tskp,
                  READ_STEP, i);
    return;
  }
  const AwpGrid& g = awp_state().grid[d_i];
  FrcLaunch A;
  A.step = i;
  A.read_step = READ_STEP;
  A.dim = dim;
  A.npsrc = npsrc;
  A.tskp = tskp;
  A.xmin = ymin;  // the launcher's ymin/ymax bound the kernel's idx (kernel.cu:1836-1875)
  A.xmax = ymax;
  A.yline = g.yline;
  A.slice = g.slice;
  A.DT = g.DT;
  const int block = npsrc < 256 ? npsrc : 256;
  hipLaunchKernelGGL(frcvel_kernel, dim3(awp_cdiv(npsrc, block)), dim3(block), 0, St, psrc, axx,
                     ayy, azz, u1, v1, w1, A);
  awp_check_launch(W);
}

}  // extern "C"
