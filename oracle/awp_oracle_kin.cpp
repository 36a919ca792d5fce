// CPU oracle (test infrastructure only) for kinematic finite-fault sources
// and prescribed-velocity sources.  Each function restates the reference:
//   awp_oracle_compmt          src/awp/kernel.cu:3706-3727 (compmt)
//   awp_oracle_brune           src/awp/kernel.cu:3729-3738 (brune)
//   awp_oracle_liu             src/awp/kernel.cu:3740-3761 (liu)
//   awp_oracle_lfilter         src/awp/kernel.cu:3763-3775 (lfilter)
//   awp_oracle_addkinsrc       src/awp/kernel.cu:3777-3858 (addkinsrc_cu)
//   awp_oracle_frcvel          src/awp/kernel.cu:1836-1929 (frcvel_cu)
// with the reference's single/double precision mix and glibc's sinf, cosf,
// expf and powf.  Single-precision denormals are flushed, as on the GPU.
#include <math.h>
#include <xmmintrin.h>

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kNgsl = 8, kAlign = 32;

inline void ftz_daz() { _mm_setcsr(_mm_getcsr() | 0x8040); }

}  // namespace

extern "C" {

// m = xx, yy, zz, xz, yz, xy
void awp_oracle_compmt(float str, float dip, float rake, float* m) {
  const float a = sinf(dip) * cosf(rake);
  m[1] = -(a * sinf((float)(2. * str)) +
           sinf((float)(2. * dip)) * sinf(rake) * sinf(str) * sinf(str));
  const float b = sinf((float)(2. * dip)) * sinf(rake) * sinf((float)(2. * str));
  m[5] = (float)((double)(a * cosf((float)(2. * str))) + 0.5 * (double)b);
  m[4] = cosf(dip) * cosf(rake) * cosf(str) + cosf((float)(2. * dip)) * sinf(rake) * sinf(str);
  m[0] = a * sinf((float)(2. * str)) -
         sinf((float)(2. * dip)) * sinf(rake) * cosf(str) * cosf(str);
  m[3] = cosf(dip) * cosf(rake) * sinf(str) - cosf((float)(2. * dip)) * sinf(rake) * cosf(str);
  m[2] = sinf((float)(2. * dip)) * sinf(rake);
}

float awp_oracle_brune(float freq, float time) {
  const float omega = (float)(freq * kPi * 2.);
  if (time > 0.) return powf(omega, 2.f) * time * expf(-omega * time);
  return 0.f;
}

float awp_oracle_liu(float tau, float time) {
  const float tau1 = (float)(0.13 * tau);
  const float tau2 = tau - tau1;
  const float CN = (float)(kPi / (1.4 * kPi * tau1 + 1.2 * tau1 + 0.3 * kPi * tau2));
  if (time < tau1)
    return (float)(CN * (0.7 - 0.7 * cosf((float)(kPi * time / tau1)) +
                         0.6 * sinf((float)(0.5 * kPi * time / tau1))));
  if (time < 2 * tau1)
    return (float)(CN * (1.0 - 0.7 * cosf((float)(kPi * time / tau1)) +
                         0.3 * cosf((float)(kPi * (time - tau1) / tau2))));
  if (time < tau) return (float)(CN * (0.3 + 0.3 * cosf((float)(kPi * (time - tau1) / tau2))));
  return 0.f;
}

double awp_oracle_lfilter(int order, const double* b, const double* a, double x, double* d) {
  const double y = (b[0] * x + d[0]) / a[0];
  for (int n = 0; n < order - 1; ++n) d[n] = b[n + 1] * x - a[n + 1] * y + d[n + 1];
  d[order - 1] = b[order] * x - a[order] * y;
  return y;
}

// One launch of addkinsrc_cu over npsrc subfaults, in subfault order.
void awp_oracle_addkinsrc(int step, int dim, const int* psrc, int npsrc, const float* mu,
                          const float* axx, const float* ayy, const float* azz, const float* axz,
                          const float* ayz, float* xx, float* yy, float* zz, float* xy, float* yz,
                          float* xz, float* mom, double* filt_d, int order, const double* b,
                          const double* a, float DT, float DH, int yline, long long slice) {
  ftz_daz();
  for (int j = 0; j < npsrc; ++j) {
    const int stf_type = (int)axx[j * 2];
    const float slip = ayy[j * 2], ruptime = azz[j * 2], risetime = axz[j * 2];
    const float atime = step * DT;
    if (!(atime > ruptime)) continue;
    const float area = ayz[j * 2];
    const float strike = (float)(axx[j * 2 + 1] / 180. * kPi);
    const float dip = (float)(ayy[j * 2 + 1] / 180. * kPi);
    const float rake = (float)(azz[j * 2 + 1] / 180. * kPi);
    float m[6];
    awp_oracle_compmt(strike, dip, rake, m);
    float stf;
    if (stf_type == 1.0f) {
      const float freq = (float)(1. / risetime);
      stf = awp_oracle_brune(freq, atime - ruptime);
    } else if (stf_type == 2.0f) {
      stf = awp_oracle_liu(risetime, atime - ruptime);
    } else {
      stf = 0.f;
    }
    if (order > 0) stf = (float)awp_oracle_lfilter(order, b, a, (double)stf, filt_d + j * (order + 1));
    const float vtst = DT / (DH * DH * DH);
    const int idx = psrc[j * dim] + 1 + kNgsl;
    const int idy = psrc[j * dim + 1] + 1 + kNgsl;
    const int idz = psrc[j * dim + 2] + kAlign - 1;
    const long long pos = idx * slice + (long long)idy * yline + idz;
    const float sam = slip * area / mu[pos];
    stf = stf * sam;
    mom[j] = mom[j] + stf * DT;
    stf = stf * vtst;
    xx[pos] = xx[pos] - stf * m[0];
    yy[pos] = yy[pos] - stf * m[1];
    zz[pos] = zz[pos] - stf * m[2];
    xz[pos] = xz[pos] - stf * m[3];
    yz[pos] = yz[pos] - stf * m[4];
    xy[pos] = xy[pos] - stf * m[5];
  }
}

// One launch of frcvel_cu over npsrc cells, in source order.
void awp_oracle_frcvel(int step, int read_step, int dim, const int* psrc, int npsrc, int tskp,
                       float* axx, float* ayy, float* azz, float* u1, float* v1, float* w1,
                       int xmin, int xmax, float DT, int yline, long long slice) {
  ftz_daz();
  for (int j = 0; j < npsrc; ++j) {
    const int i = step - 1;
    int i0 = int(tskp * floorf(float(i + 1.) / float(tskp)));
    int i1 = int(tskp * ceilf(float(i + 1.) / float(tskp)));
    const float pfact = float(i + 1 - i0) / float(tskp);
    i0 /= tskp;
    i1 /= tskp;
    const int idx = psrc[j * dim] + 1 + kNgsl;
    const int idy = psrc[j * dim + 1] + 1 + kNgsl;
    const int idz = psrc[j * dim + 2] + kAlign - 1;
    const long long pos = idx * slice + (long long)idy * yline + idz;
    const bool abvmin = (xmin == -1) || (idx >= xmin);
    const bool blwmax = (xmax == -1) || (idx <= xmax);
    if (!(abvmin && blwmax)) continue;
    if (i < read_step * tskp) {
      const float u1_p = axx[i0 * npsrc + j], v1_p = ayy[i0 * npsrc + j],
                  w1_p = azz[i0 * npsrc + j];
      const float u1_n = axx[i1 * npsrc + j], v1_n = ayy[i1 * npsrc + j],
                  w1_n = azz[i1 * npsrc + j];
      const float u1_i = u1_p + (u1_n - u1_p) * pfact;
      const float v1_i = v1_p + (v1_n - v1_p) * pfact;
      const float w1_i = w1_p + (w1_n - w1_p) * pfact;
      if (i == read_step * tskp - 1) {
        axx[j] = axx[i1 * npsrc + j];
        ayy[j] = ayy[i1 * npsrc + j];
        azz[j] = azz[i1 * npsrc + j];
        axx[i1 * npsrc + j] = (u1_n - u1[pos]) / DT;
        ayy[i1 * npsrc + j] = (v1_n - v1[pos]) / DT;
        azz[i1 * npsrc + j] = (w1_n - w1[pos]) / DT;
      }
      u1[pos] = u1_i;
      v1[pos] = v1_i;
      w1[pos] = w1_i;
    } else {
      u1[pos] = u1[pos] + axx[read_step * npsrc + j] * DT;
      v1[pos] = v1[pos] + ayy[read_step * npsrc + j] * DT;
      w1[pos] = w1[pos] + azz[read_step * npsrc + j] * DT;
    }
  }
}

}  // extern "C"
