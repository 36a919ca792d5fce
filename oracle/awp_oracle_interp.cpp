// CPU oracle (test infrastructure only: the tests and the CPU baseline call
// it; the product never does) for Lagrange-interpolated receivers and
// sources.  Each function restates one reference function:
//   awp_oracle_lagrange_node_poly  src/interpolation/lagrange.c:70-77
//   awp_oracle_lagrange_basis      src/interpolation/lagrange.c:20-52
//   awp_oracle_argnearest          src/interpolation/interpolation.c:14-38
//   awp_oracle_argnearest_range    src/interpolation/interpolation.c:49-63
//   awp_oracle_lagrange1_coef      src/interpolation/interpolation.c:90-128
//   awp_oracle_lagrange1           src/interpolation/interpolation.c:65-88
//   awp_oracle_lagrange_h          src/interpolation/interpolation.cu:43-67
//   awp_oracle_interp              src/interpolation/interpolation.cu:177-199
//   awp_oracle_source_add          src/topography/sources/source.cu:24-158
// Host functions keep the reference's single/double precision mix.  The
// device restatements (interp, source_add) flush single-precision denormals
// as the GPU build does, and add the sources in ascending query order, the
// order a race-free run of the reference's scatter would use.
#include <math.h>
#include <xmmintrin.h>

#include <algorithm>

namespace {

constexpr double kFltol = 1e-5;  // include/awp/definitions.h:43
constexpr int kAlign = 32, kNgsl = 8;

inline void ftz_daz() { _mm_setcsr(_mm_getcsr() | 0x8040); }

inline long long grid_index(int off_x, int off_y, int off_z, int line, int slice, int i, int j,
                            int k) {
  return (long long)off_z + k + (long long)(off_y + j) * line + (long long)(off_x + i) * slice;
}

}  // namespace

extern "C" {

double awp_oracle_lagrange_node_poly(float x0, const float* x, int n) {
  double l = 1.0;
  for (int i = 0; i < n; ++i) {
    const float d = x0 - x[i];
    l = l * (double)d;
  }
  return l;
}

int awp_oracle_lagrange_basis(float* l, int n, const float* x, float x0) {
  if (n == 1) {
    l[0] = 1.0f;
    return 0;
  }
  const float d0 = x0 - x[0], dn = x0 - x[n - 1];
  if ((double)d0 < -kFltol) return 600;
  if ((double)dn > kFltol) return 601;
  const double Ld = awp_oracle_lagrange_node_poly(x0, x, n);
  const float L = (float)Ld;  // `_prec L` receives the double product
  for (int j = 0; j < n; ++j) {
    l[j] = 1.0f;
    const float dj = x0 - x[j];
    if (fabs((double)dj) < kFltol) continue;
    double denom = 1.0;
    for (int k = 0; k < n; ++k) {
      if (k == j) continue;
      const float dk = x[j] - x[k];
      denom = denom * (double)dk;
    }
    const double lambda = 1.0 / denom;
    const double v = ((double)L * lambda) / (double)dj;
    l[j] = (float)v;
  }
  return 0;
}

int awp_oracle_argnearest(int* nearest, const float* points, int n, float query) {
  if (query < points[0] || isnan(query)) {
    *nearest = 0;
    return 700;
  }
  if (query > points[n - 1]) {
    *nearest = n - 1;
    return 701;
  }
  int argmin = n + 1;
  float dist = INFINITY;
  for (int i = 0; i < n; ++i) {
    const float cur = (float)fabs((double)(query - points[i]));
    if (cur < dist) {
      dist = cur;
      argmin = i;
    }
  }
  *nearest = argmin;
  return 0;
}

int awp_oracle_argnearest_range(int* first, int* last, int lower, int upper, int nearest, int n,
                                float /*query*/) {
  const int diffleft = nearest - lower;
  const int diffright = n - nearest - upper - 1;
  *first = std::max(diffleft, 0) + std::min(diffright, 0);
  *last = std::min(nearest + upper + 1, n) - std::min(diffleft, 0);
  return 0;
}

int awp_oracle_lagrange1_coef(float* xloc, float* l, int* first, const float* x, int n,
                              float query, int deg) {
  int nearest = 0;
  int lower = (int)ceil((double)deg * 0.5);
  int upper = (int)ceil((double)deg * 0.5);
  float q = query;
  int err = awp_oracle_argnearest(&nearest, x, n, query);
  if (err == 700) q = x[0];
  if (err == 701) q = x[n - 1];
  if (deg % 2 == 1) {
    const float side = x[nearest] - query;
    if (side > 0) {
      lower = (int)ceil((double)deg * 0.5);
      upper = (int)floor((double)deg * 0.5);
    } else {
      lower = (int)floor((double)deg * 0.5);
      upper = (int)ceil((double)deg * 0.5);
    }
  }
  int lo = 0, hi = 0;
  err |= awp_oracle_argnearest_range(&lo, &hi, lower, upper, nearest, n, q);
  for (int j = 0; j < deg + 1; ++j) xloc[j] = x[lo + j];
  *first = lo;
  err |= awp_oracle_lagrange_basis(l, deg + 1, xloc, q);
  return err;
}

// One-dimensional interpolation of `in` (on nodes x) at m queries.
int awp_oracle_lagrange1(float* out, const float* x, const float* in, int n, const float* query,
                         int m, int deg) {
  float l[32], xloc[32];
  if (deg + 1 > 32) return 902;
  for (int i = 0; i < m; ++i) {
    int lower = 0;
    awp_oracle_lagrange1_coef(xloc, l, &lower, x, n, query[i], deg);
    float acc = 0.0f;
    for (int j = 0; j < deg + 1; ++j) acc = acc + l[j] * in[lower + j];
    out[i] = acc;
  }
  return 0;
}

int awp_oracle_lagrange_h(float* lx, float* ly, float* lz, int* ix, int* iy, int* iz,
                          const float* x, const float* y, const float* z, int nx, int ny, int nz,
                          const float* qx, const float* qy, const float* qz, int num_query,
                          int deg) {
  float loc[32];
  if (deg + 1 > 32) return 902;
  const int nb = deg + 1;
  int err = 0;
  for (int q = 0; q < num_query; ++q) {
    err = awp_oracle_lagrange1_coef(loc, lx + q * nb, ix + q, x, nx, qx[q], deg);
    err = awp_oracle_lagrange1_coef(loc, ly + q * nb, iy + q, y, ny, qy[q], deg);
    err = awp_oracle_lagrange1_coef(loc, lz + q * nb, iz + q, z, nz, qz[q], deg);
  }
  return err;
}

void awp_oracle_interp(float* out, const float* in, const float* lx, const float* ly,
                       const float* lz, int nb, const int* ix, const int* iy, const int* iz,
                       const int* ridx, int num_query, int off_x, int off_y, int off_z, int line,
                       int slice) {
  ftz_daz();
  for (int q = 0; q < num_query; ++q) {
    out[ridx[q]] = 0.0f;
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j)
        for (int k = 0; k < nb; ++k) {
          const long long pos =
              grid_index(off_x, off_y, off_z, line, slice, ix[q] + i, iy[q] + j, iz[q] + k);
          const float w = lx[q * nb + i] * ly[q * nb + j];
          const float t = (w * lz[q * nb + k]) * in[pos];
          out[ridx[q]] = out[ridx[q]] + t;
        }
  }
}

// kind 0: cusource_add_cartesian, 1: cusource_add_curvilinear,
// 2: cusource_add_force (d1 = density and quad_weight are read only there).
void awp_oracle_source_add(int kind, float* out, const float* in, const float* lx,
                           const float* ly, const float* lz, int nb, const int* ix,
                           const int* iy, const int* iz, const int* lidx, int num_query,
                           int off_x, int off_y, int off_z, int line, int slice, float h,
                           float dt, float quad_weight, const float* f, const float* dg,
                           const float* d1, int ny, int nz) {
  ftz_daz();
  const float h3 = (h * h) * h;
  const float dth = dt / h3;
  const int fline = 2 * kAlign + 2 * kNgsl + ny + 4;
  for (int q = 0; q < num_query; ++q)
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j)
        for (int k = 0; k < nb; ++k) {
          const int gi = ix[q] + i, gj = iy[q] + j, gk = iz[q] + k;
          const long long pos = grid_index(off_x, off_y, off_z, line, slice, gi, gj, gk);
          float t = -dth;
          t = t * lx[q * nb + i];
          t = t * ly[q * nb + j];
          t = t * lz[q * nb + k];
          t = t * in[lidx[q]];
          if (kind != 0) {
            const float fv = f[gj + kAlign + (long long)(gi + 2) * fline + 2];
            const float gv = dg[gk + kAlign];
            float Ji;
            if (kind == 1) {
              const float fg = fv * gv;
              Ji = (float)(1.0 / (double)fg);
            } else {
              const long long r = gk + kAlign +
                                  (long long)(2 * kAlign + nz) * (gi + kNgsl + 2) *
                                      (2 * kNgsl + ny + 4) +
                                  (long long)(2 * kAlign + nz) * (gj + kNgsl + 2);
              const float den = (fv * gv) * d1[r];
              Ji = (-quad_weight) / den;
            }
            t = t * Ji;
          }
          out[pos] = out[pos] + t;
        }
}

}  // extern "C"
