// Curvilinear-topography oracle: CPU restatement of the reference's
// topography kernels, metric builders and geometry.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP library, like
// awp_oracle.cpp): only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.
//
// Each function follows the reference in operation order and literal type:
// for T=float the arithmetic is the reference's single precision (the
// sub-expressions it writes with double literals are evaluated in double
// here too), for T=double it is the FP64 oracle.  Built with
// -ffp-contract=off, FTZ/DAZ as in the reference's -use_fast_math build.
//
// Coordinates.  The kernels index the wavefield with topographic
// coordinates (i, j, k) in [0,nx) x [0,ny) x [0,nz), the array element being
// (k+align) + (2 align+nz)*(j+ngsl+2) + (2 align+nz)(2 ngsl+ny+4)*(i+ngsl+2)
// (velocity_unroll.cu:103-140); the stress kernels use array indices
// directly (stress_attenuation.cu:9-22).  The 2-D metrics f* are
// (nx+4+2 ngsl) x (ny+4+2 ngsl+2 align) arrays, element (i, j) at
// (j+ngsl+2) + align + (i+ngsl+2)*(2 align+2 ngsl+ny+4); the 1-D metrics g*
// have nz+2 align elements, level k at k+align (metrics.c:14-33, 206-219).
//
// Pinning (tests/test_oracle_topo.py): the coefficient tables against
// polynomial exactness on the node / cell-centre grids in z; the metric
// builders against constant, linear and quadratic f and g (the reference's
// test_metrics.c cases); both kernels against the Cartesian oracle kernels on
// flat topography (f = 1, g linear; the reference's test_attenuation.cu
// invariant); the curvilinear derivative terms against the physical z
// coordinate of a tilted, curved grid (d z / d x = 0, d z / d z = 1 in the
// interior).
#include <cmath>
#include <cstdint>
#include <omp.h>
#include <xmmintrin.h>

#include "topo_tables.h"

namespace {

inline void ftz_daz() { _mm_setcsr(_mm_getcsr() | 0x8040); }

constexpr int NGSL = 8;    // include/awp/pmcl3d_cons.h:22
constexpr int ALIGN = 32;  // include/awp/pmcl3d_cons.h:16

// Interior stencils (velocity_unroll.cu:24-58, stress_attenuation.cu:112-150):
// 4-point interpolation p4, 4-point staggered derivative d4.
constexpr double P4[4] = {-0.0625000000000000, 0.5625000000000000, 0.5625000000000000,
                          -0.0625000000000000};
constexpr double D4[4] = {0.0416666666666667, -1.1250000000000000, 1.1250000000000000,
                          -0.0416666666666667};

// Array geometry of one grid.
struct TGrid {
  int nx, ny, nz;
  long sl, yl;  // wavefield strides
  long frow;    // metric row length
  // topographic (i, j, k) -> wavefield element
  long p(int i, int j, int k) const {
    return (long)(k + ALIGN) + yl * (j + NGSL + 2) + sl * (i + NGSL + 2);
  }
  // topographic (i, j) -> metric element
  long f(int i, int j) const { return (long)(j + ALIGN + NGSL + 2) + (long)(i + NGSL + 2) * frow; }
};

TGrid make_tgrid(int nx, int ny, int nz) {
  TGrid g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.yl = nz + 2 * ALIGN;
  g.sl = (long)(ny + 4 + 2 * NGSL) * g.yl;
  g.frow = 2 * ALIGN + 2 * NGSL + ny + 4;
  return g;
}

// Metric arrays in the order of the C ABI.
template <typename T>
struct Metrics {
  const T *f, *f_1, *f_2, *f_c, *f1_1, *f1_2, *f1_c, *f2_1, *f2_2, *f2_c;
  const T *g, *g3, *g3_c, *g_c;
};

template <typename T>
Metrics<T> metrics_from(const T* const* mf, const T* const* mg) {
  Metrics<T> m;
  m.f = mf[0];
  m.f_1 = mf[1];
  m.f_2 = mf[2];
  m.f_c = mf[3];
  m.f1_1 = mf[4];
  m.f1_2 = mf[5];
  m.f1_c = mf[6];
  m.f2_1 = mf[7];
  m.f2_2 = mf[8];
  m.f2_c = mf[9];
  m.g = mg[0];
  m.g3 = mg[1];
  m.g3_c = mg[2];
  m.g_c = mg[3];
  return m;
}

// nu * 1.0 / x with nu a float: the quotient is a double expression
// (velocity_unroll.cu:156-161).
template <typename T>
inline T dbl_quot(T num, T den) {
  return (T)((double)num * 1.0 / (double)den);
}

// ---------------------------------------------------------------------------
// Velocity.  dtopo_vel_111_unroll (kernels/velocity_unroll.cu:1-461) for
// levels dm_offset=3 <= k < nz-6, dtopo_vel_112 (optimized_velocity.cu:
// 1040-1627) for the top six levels k = nz-1-kk, kk = 0..5.  Bounds i in
// [bi, ei), j in [bj, ej) (topographic).
template <typename T>
void topo_vel(T* u1, T* u2, T* u3, const T* dcrjx, const T* dcrjy, const T* dcrjz,
              const Metrics<T>& M, const T* rho, const T* s11, const T* s12, const T* s13,
              const T* s22, const T* s23, const T* s33, T a, T nu, const TGrid& G, int bi,
              int bj, int ei, int ej) {
  const int nz = G.nz;
  const int dm_offset = 3;
  // const float coefficient tables (the reference's initialisers)
  T p4[4], d4[4], dhpz4[7], dphz4[7], dhpz4r[6][9], dhz4r[6][8], dphz4r[6][9], dz4r[6][7];
  for (int q = 0; q < 4; ++q) {
    p4[q] = (T)P4[q];
    d4[q] = (T)D4[q];
  }
  for (int q = 0; q < 7; ++q) {
    dhpz4[q] = (T)TV_DHPZ4[q];
    dphz4[q] = (T)TV_DPHZ4[q];
  }
  for (int r = 0; r < 6; ++r) {
    for (int q = 0; q < 9; ++q) {
      dhpz4r[r][q] = (T)TV_DHPZ4R[r][q];
      dphz4r[r][q] = (T)TV_DPHZ4R[r][q];
    }
    for (int q = 0; q < 8; ++q) dhz4r[r][q] = (T)TV_DHZ4R[r][q];
    for (int q = 0; q < 7; ++q) dz4r[r][q] = (T)TV_DZ4R[r][q];
  }
  // phx4 = phy4 = px4 = py4 = p4; dhx4 = dhy4 = dhz4 = dx4 = dy4 = dz4 = d4
  const T *phx4 = p4, *phy4 = p4, *px4 = p4, *py4 = p4;
  const T *dhx4 = d4, *dhy4 = d4, *dhz4 = d4, *dx4 = d4, *dy4 = d4, *dz4 = d4;
  const int ie = ei < G.nx ? ei : G.nx, je = ej < G.ny ? ej : G.ny;

#define RHO(i, j, k) rho[G.p(i, j, k)]
#define S11(i, j, k) s11[G.p(i, j, k)]
#define S12(i, j, k) s12[G.p(i, j, k)]
#define S13(i, j, k) s13[G.p(i, j, k)]
#define S22(i, j, k) s22[G.p(i, j, k)]
#define S23(i, j, k) s23[G.p(i, j, k)]
#define S33(i, j, k) s33[G.p(i, j, k)]
#define MF(name, i, j) M.name[G.f(i, j)]
#define MG(name, k) M.name[(k) + ALIGN]
#define DCRJ(i, j, k) (dcrjx[(i) + NGSL + 2] * dcrjy[(j) + NGSL + 2] * dcrjz[(k) + ALIGN])

#pragma omp parallel for schedule(static)
  for (int i = bi; i < ie; ++i) {
    ftz_daz();
    for (int j = bj; j < je; ++j) {
      // interior levels
      for (int k = dm_offset; k < nz - 6; ++k) {
        const T c = (T)0.25f;
        const T rho1 = c * (RHO(i, j, k - 1) + RHO(i, j - 1, k - 1)) +
                       c * (RHO(i, j, k) + RHO(i, j - 1, k));
        const T rho2 = c * (RHO(i, j, k - 1) + RHO(i - 1, j, k - 1)) +
                       c * (RHO(i, j, k) + RHO(i - 1, j, k));
        const T rho3 = c * (RHO(i, j, k) + RHO(i - 1, j, k)) +
                       c * (RHO(i, j - 1, k) + RHO(i - 1, j - 1, k));
        T Ai1 = MF(f_1, i, j) * MG(g3_c, k) * rho1;
        Ai1 = dbl_quot(nu, Ai1);
        T Ai2 = MF(f_2, i, j) * MG(g3_c, k) * rho2;
        Ai2 = dbl_quot(nu, Ai2);
        T Ai3 = MF(f_c, i, j) * MG(g3, k) * rho3;
        Ai3 = dbl_quot(nu, Ai3);
        const T f_dcrj = DCRJ(i, j, k);
        // 7-point z derivative of 4-point horizontal interpolants, terms in
        // the order 3, 0, 1, 2, 4, 5, 6 of the stencil.
        static const int ZO[7] = {3, 0, 1, 2, 4, 5, 6};
        T z11 = 0, z12 = 0, z12x = 0, z22 = 0, z13 = 0, z23 = 0;
        for (int q = 0; q < 7; ++q) {
          const int m = ZO[q], kk = k + m - 3;
          const T t11 = dhpz4[m] * MG(g_c, kk) *
                        (phx4[2] * S11(i, j, kk) + phx4[0] * S11(i - 2, j, kk) +
                         phx4[1] * S11(i - 1, j, kk) + phx4[3] * S11(i + 1, j, kk));
          const T t12 = dhpz4[m] * MG(g_c, kk) *
                        (phy4[2] * S12(i, j, kk) + phy4[0] * S12(i, j - 2, kk) +
                         phy4[1] * S12(i, j - 1, kk) + phy4[3] * S12(i, j + 1, kk));
          const T t12x = dhpz4[m] * MG(g_c, kk) *
                         (px4[1] * S12(i, j, kk) + px4[0] * S12(i - 1, j, kk) +
                          px4[2] * S12(i + 1, j, kk) + px4[3] * S12(i + 2, j, kk));
          const T t22 = dhpz4[m] * MG(g_c, kk) *
                        (py4[1] * S22(i, j, kk) + py4[0] * S22(i, j - 1, kk) +
                         py4[2] * S22(i, j + 1, kk) + py4[3] * S22(i, j + 2, kk));
          const T t13 = dphz4[m] * MG(g, kk) *
                        (px4[1] * S13(i, j, kk) + px4[0] * S13(i - 1, j, kk) +
                         px4[2] * S13(i + 1, j, kk) + px4[3] * S13(i + 2, j, kk));
          const T t23 = dphz4[m] * MG(g, kk) *
                        (phy4[2] * S23(i, j, kk) + phy4[0] * S23(i, j - 2, kk) +
                         phy4[1] * S23(i, j - 1, kk) + phy4[3] * S23(i, j + 1, kk));
          if (q == 0) {
            z11 = t11; z12 = t12; z12x = t12x; z22 = t22; z13 = t13; z23 = t23;
          } else {
            z11 = z11 + t11; z12 = z12 + t12; z12x = z12x + t12x; z22 = z22 + t22;
            z13 = z13 + t13; z23 = z23 + t23;
          }
        }
        T s;
        // u1 (velocity_unroll.cu:164-254)
        s = dhx4[2] * MF(f_c, i, j) * MG(g3_c, k) * S11(i, j, k);
        s = s + dhx4[0] * MF(f_c, i - 2, j) * MG(g3_c, k) * S11(i - 2, j, k);
        s = s + dhx4[1] * MF(f_c, i - 1, j) * MG(g3_c, k) * S11(i - 1, j, k);
        s = s + dhx4[3] * MF(f_c, i + 1, j) * MG(g3_c, k) * S11(i + 1, j, k);
        s = s + dhy4[2] * MF(f, i, j) * MG(g3_c, k) * S12(i, j, k);
        s = s + dhy4[0] * MF(f, i, j - 2) * MG(g3_c, k) * S12(i, j - 2, k);
        s = s + dhy4[1] * MF(f, i, j - 1) * MG(g3_c, k) * S12(i, j - 1, k);
        s = s + dhy4[3] * MF(f, i, j + 1) * MG(g3_c, k) * S12(i, j + 1, k);
        s = s + dhz4[2] * S13(i, j, k);
        s = s + dhz4[0] * S13(i, j, k - 2);
        s = s + dhz4[1] * S13(i, j, k - 1);
        s = s + dhz4[3] * S13(i, j, k + 1);
        s = s - MF(f1_1, i, j) * z11;
        s = s - MF(f2_1, i, j) * z12;
        const T v1 = (a * u1[G.p(i, j, k)] + Ai1 * s) * f_dcrj;
        // u2 (velocity_unroll.cu:255-343)
        s = dhz4[2] * S23(i, j, k);
        s = s + dhz4[0] * S23(i, j, k - 2);
        s = s + dhz4[1] * S23(i, j, k - 1);
        s = s + dhz4[3] * S23(i, j, k + 1);
        s = s + dx4[1] * MF(f, i, j) * MG(g3_c, k) * S12(i, j, k);
        s = s + dx4[0] * MF(f, i - 1, j) * MG(g3_c, k) * S12(i - 1, j, k);
        s = s + dx4[2] * MF(f, i + 1, j) * MG(g3_c, k) * S12(i + 1, j, k);
        s = s + dx4[3] * MF(f, i + 2, j) * MG(g3_c, k) * S12(i + 2, j, k);
        s = s + dy4[1] * MF(f_c, i, j) * MG(g3_c, k) * S22(i, j, k);
        s = s + dy4[0] * MF(f_c, i, j - 1) * MG(g3_c, k) * S22(i, j - 1, k);
        s = s + dy4[2] * MF(f_c, i, j + 1) * MG(g3_c, k) * S22(i, j + 1, k);
        s = s + dy4[3] * MF(f_c, i, j + 2) * MG(g3_c, k) * S22(i, j + 2, k);
        s = s - MF(f1_2, i, j) * z12x;
        s = s - MF(f2_2, i, j) * z22;
        const T v2 = (a * u2[G.p(i, j, k)] + Ai2 * s) * f_dcrj;
        // u3 (velocity_unroll.cu:344-432)
        s = dhy4[2] * MF(f_2, i, j) * MG(g3, k) * S23(i, j, k);
        s = s + dhy4[0] * MF(f_2, i, j - 2) * MG(g3, k) * S23(i, j - 2, k);
        s = s + dhy4[1] * MF(f_2, i, j - 1) * MG(g3, k) * S23(i, j - 1, k);
        s = s + dhy4[3] * MF(f_2, i, j + 1) * MG(g3, k) * S23(i, j + 1, k);
        s = s + dx4[1] * MF(f_1, i, j) * MG(g3, k) * S13(i, j, k);
        s = s + dx4[0] * MF(f_1, i - 1, j) * MG(g3, k) * S13(i - 1, j, k);
        s = s + dx4[2] * MF(f_1, i + 1, j) * MG(g3, k) * S13(i + 1, j, k);
        s = s + dx4[3] * MF(f_1, i + 2, j) * MG(g3, k) * S13(i + 2, j, k);
        s = s + dz4[1] * S33(i, j, k);
        s = s + dz4[0] * S33(i, j, k - 1);
        s = s + dz4[2] * S33(i, j, k + 1);
        s = s + dz4[3] * S33(i, j, k + 2);
        s = s - MF(f1_c, i, j) * z13;
        s = s - MF(f2_c, i, j) * z23;
        const T v3 = (a * u3[G.p(i, j, k)] + Ai3 * s) * f_dcrj;
        // results are written after all reads of the column's neighbours:
        // the reference keeps them in registers until the end of the thread
        // (velocity_unroll.cu:436-446) and no output feeds another output.
        u1[G.p(i, j, k)] = v1;
        u2[G.p(i, j, k)] = v2;
        u3[G.p(i, j, k)] = v3;
      }
      // top six levels: one-sided boundary stencils
      for (int kk = 0; kk < 6; ++kk) {
        const int k = nz - 1 - kk, kb = nz - kk - 2;
        const T rho1 = (T)(0.25 * (double)(RHO(i, j, kb + 0) + RHO(i, j - 1, kb + 0)) +
                           0.25 * (double)(RHO(i, j, kb + 1) + RHO(i, j - 1, kb + 1)));
        const T rho2 = (T)(0.25 * (double)(RHO(i, j, kb + 0) + RHO(i - 1, j, kb + 0)) +
                           0.25 * (double)(RHO(i, j, kb + 1) + RHO(i - 1, j, kb + 1)));
        const T rho3 = (T)(0.25 * (double)(RHO(i, j, kb + 1) + RHO(i - 1, j, kb + 1)) +
                           0.25 * (double)(RHO(i, j - 1, kb + 1) + RHO(i - 1, j - 1, kb + 1)));
        T Ai1 = MF(f_1, i, j) * MG(g3_c, k) * rho1;
        Ai1 = dbl_quot(nu, Ai1);
        T Ai2 = MF(f_2, i, j) * MG(g3_c, k) * rho2;
        Ai2 = dbl_quot(nu, Ai2);
        T Ai3 = MF(f_c, i, j) * MG(g3, k) * rho3;
        Ai3 = dbl_quot(nu, Ai3);
        const T f_dcrj = DCRJ(i, j, k);
        // z sums over levels nz-9 .. nz-1, table column m at level nz-1-m,
        // terms in the order m = 8, 7, ..., 0.
        T z11 = 0, z12 = 0, z12x = 0, z22 = 0, z13 = 0, z23 = 0;
        for (int m = 8; m >= 0; --m) {
          const int l = nz - 1 - m;
          const T t11 = dhpz4r[kk][m] * MG(g_c, l) *
                        (phx4[2] * S11(i, j, l) + phx4[0] * S11(i - 2, j, l) +
                         phx4[1] * S11(i - 1, j, l) + phx4[3] * S11(i + 1, j, l));
          const T t12 = dhpz4r[kk][m] * MG(g_c, l) *
                        (phy4[2] * S12(i, j, l) + phy4[0] * S12(i, j - 2, l) +
                         phy4[1] * S12(i, j - 1, l) + phy4[3] * S12(i, j + 1, l));
          const T t12x = dhpz4r[kk][m] * MG(g_c, l) *
                         (px4[1] * S12(i, j, l) + px4[0] * S12(i - 1, j, l) +
                          px4[2] * S12(i + 1, j, l) + px4[3] * S12(i + 2, j, l));
          const T t22 = dhpz4r[kk][m] * MG(g_c, l) *
                        (py4[1] * S22(i, j, l) + py4[0] * S22(i, j - 1, l) +
                         py4[2] * S22(i, j + 1, l) + py4[3] * S22(i, j + 2, l));
          const T t13 = dphz4r[kk][m] * MG(g, l) *
                        (px4[1] * S13(i, j, l) + px4[0] * S13(i - 1, j, l) +
                         px4[2] * S13(i + 1, j, l) + px4[3] * S13(i + 2, j, l));
          const T t23 = dphz4r[kk][m] * MG(g, l) *
                        (phy4[2] * S23(i, j, l) + phy4[0] * S23(i, j - 2, l) +
                         phy4[1] * S23(i, j - 1, l) + phy4[3] * S23(i, j + 1, l));
          if (m == 8) {
            z11 = t11; z12 = t12; z12x = t12x; z22 = t22; z13 = t13; z23 = t23;
          } else {
            z11 = z11 + t11; z12 = z12 + t12; z12x = z12x + t12x; z22 = z22 + t22;
            z13 = z13 + t13; z23 = z23 + t23;
          }
        }
        T s;
        // u1 (optimized_velocity.cu:1223-1340)
        s = dhx4[2] * MF(f_c, i, j) * MG(g3_c, k) * S11(i, j, k);
        s = s + dhx4[0] * MF(f_c, i - 2, j) * MG(g3_c, k) * S11(i - 2, j, k);
        s = s + dhx4[1] * MF(f_c, i - 1, j) * MG(g3_c, k) * S11(i - 1, j, k);
        s = s + dhx4[3] * MF(f_c, i + 1, j) * MG(g3_c, k) * S11(i + 1, j, k);
        s = s + dhy4[2] * MF(f, i, j) * MG(g3_c, k) * S12(i, j, k);
        s = s + dhy4[0] * MF(f, i, j - 2) * MG(g3_c, k) * S12(i, j - 2, k);
        s = s + dhy4[1] * MF(f, i, j - 1) * MG(g3_c, k) * S12(i, j - 1, k);
        s = s + dhy4[3] * MF(f, i, j + 1) * MG(g3_c, k) * S12(i, j + 1, k);
        for (int m = 7; m >= 0; --m) s = s + dhz4r[kk][m] * S13(i, j, nz - 1 - m);
        s = s - MF(f1_1, i, j) * z11;
        s = s - MF(f2_1, i, j) * z12;
        const T v1 = (a * u1[G.p(i, j, k)] + Ai1 * s) * f_dcrj;
        // u2 (optimized_velocity.cu:1342-1450): z terms first
        s = dhz4r[kk][7] * S23(i, j, nz - 8);
        for (int m = 6; m >= 0; --m) s = s + dhz4r[kk][m] * S23(i, j, nz - 1 - m);
        s = s + dx4[1] * MF(f, i, j) * MG(g3_c, k) * S12(i, j, k);
        s = s + dx4[0] * MF(f, i - 1, j) * MG(g3_c, k) * S12(i - 1, j, k);
        s = s + dx4[2] * MF(f, i + 1, j) * MG(g3_c, k) * S12(i + 1, j, k);
        s = s + dx4[3] * MF(f, i + 2, j) * MG(g3_c, k) * S12(i + 2, j, k);
        s = s + dy4[1] * MF(f_c, i, j) * MG(g3_c, k) * S22(i, j, k);
        s = s + dy4[0] * MF(f_c, i, j - 1) * MG(g3_c, k) * S22(i, j - 1, k);
        s = s + dy4[2] * MF(f_c, i, j + 1) * MG(g3_c, k) * S22(i, j + 1, k);
        s = s + dy4[3] * MF(f_c, i, j + 2) * MG(g3_c, k) * S22(i, j + 2, k);
        s = s - MF(f1_2, i, j) * z12x;
        s = s - MF(f2_2, i, j) * z22;
        const T v2 = (a * u2[G.p(i, j, k)] + Ai2 * s) * f_dcrj;
        // u3 (optimized_velocity.cu:1451-1600)
        s = dhy4[2] * MF(f_2, i, j) * MG(g3, k) * S23(i, j, k);
        s = s + dhy4[0] * MF(f_2, i, j - 2) * MG(g3, k) * S23(i, j - 2, k);
        s = s + dhy4[1] * MF(f_2, i, j - 1) * MG(g3, k) * S23(i, j - 1, k);
        s = s + dhy4[3] * MF(f_2, i, j + 1) * MG(g3, k) * S23(i, j + 1, k);
        s = s + dx4[1] * MF(f_1, i, j) * MG(g3, k) * S13(i, j, k);
        s = s + dx4[0] * MF(f_1, i - 1, j) * MG(g3, k) * S13(i - 1, j, k);
        s = s + dx4[2] * MF(f_1, i + 1, j) * MG(g3, k) * S13(i + 1, j, k);
        s = s + dx4[3] * MF(f_1, i + 2, j) * MG(g3, k) * S13(i + 2, j, k);
        for (int m = 6; m >= 0; --m) s = s + dz4r[kk][m] * S33(i, j, nz - 1 - m);
        s = s - MF(f1_c, i, j) * z13;
        s = s - MF(f2_c, i, j) * z23;
        const T v3 = (a * u3[G.p(i, j, k)] + Ai3 * s) * f_dcrj;
        u1[G.p(i, j, k)] = v1;
        u2[G.p(i, j, k)] = v2;
        u3[G.p(i, j, k)] = v3;
      }
    }
  }
#undef RHO
#undef S11
#undef S12
#undef S13
#undef S22
#undef S23
#undef S33
#undef DCRJ
}

// ---------------------------------------------------------------------------
// Stress with coarse-grained Q.  dtopo_str_111 (kernels/stress_attenuation.cu:
// 74-668) for array levels align+3 <= k < align+nz-6 and dtopo_str_112
// (:670-1410) for the top six; array index bounds i in [s_i, e_i),
// j in [s_j, e_j).  The material and memory-variable arithmetic is the
// dstrqc_new one (xl, xm, xmu1 with float literals, xmu2/xmu3 and the qpa
// quadratic in double); the constants are those topo_set_constants derives
// (stress_attenuation.cu:19-72): DT' = (dt/h)*h in _prec, dth = DT'/h,
// dh1 = 1.0/h.
template <typename T>
struct TopoConst {
  T DT, dth, dh1;
};

template <typename T>
TopoConst<T> topo_constants(double DH, double DT) {
  const T h = (T)DH, dt = (T)DT;
  const T dth0 = dt / h;        // topo_t.dth (topography.c:71)
  TopoConst<T> c;
  c.DT = dth0 * h;              // set_constants' dt argument
  c.dth = c.DT / h;
  c.dh1 = (T)(1.0 / (double)h);
  return c;
}

template <typename T>
inline T qweight(T q, const T* coeff, int ww, T wwo, bool dbl_poly) {
  T w;
  if (1.0f / (q * 2.0f) <= 200.0f) {
    if (dbl_poly)
      w = (T)((double)coeff[ww * 2 - 2] * (2. * (double)q) * (2. * (double)q) +
              (double)coeff[ww * 2 - 1] * (2. * (double)q));
    else
      w = coeff[ww * 2 - 2] * (2.0f * q) * (2.0f * q) + coeff[ww * 2 - 1] * (2.0f * q);
  } else {
    w = 2.0f * wwo * q;
  }
  return w / wwo;
}

template <typename T>
void topo_str(T* xx, T* yy, T* zz, T* xy, T* xz, T* yz, T* r1, T* r2, T* r3, T* r4, T* r5,
              T* r6, const T* u1, const T* v1, const T* w1, const Metrics<T>& M,
              const T* lam, const T* mu, const T* qp, const T* coeff, const T* qs,
              const T* dcrjx, const T* dcrjy, const T* dcrjz, const T* d_vx1, const T* d_vx2,
              const int* d_ww, const T* d_wwo, const TGrid& G, const TopoConst<T>& C,
              int s_i, int e_i, int s_j, int e_j) {
  const int nz = G.nz;
  const long s1 = G.sl, y1 = G.yl;
  T p4[4], d4[4], phdz4[7], pdhz4[7], phdz4r[6][9], dz4r[6][7], dhz4r[6][8], pdhz4r[6][9];
  for (int q = 0; q < 4; ++q) {
    p4[q] = (T)P4[q];
    d4[q] = (T)D4[q];
  }
  for (int q = 0; q < 7; ++q) {
    phdz4[q] = (T)TS_PHDZ4[q];
    pdhz4[q] = (T)TS_PDHZ4[q];
  }
  for (int r = 0; r < 6; ++r) {
    for (int q = 0; q < 9; ++q) {
      phdz4r[r][q] = (T)TS_PHDZ4R[r][q];
      pdhz4r[r][q] = (T)TS_PDHZ4R[r][q];
    }
    for (int q = 0; q < 7; ++q) dz4r[r][q] = (T)TS_DZ4R[r][q];
    for (int q = 0; q < 8; ++q) dhz4r[r][q] = (T)TS_DHZ4R[r][q];
  }
  const T *phx4 = p4, *phy4 = p4, *px4 = p4, *py4 = p4;
  const T *dhx4 = d4, *dhy4 = d4, *dhz4 = d4, *dx4 = d4, *dy4 = d4, *dz4 = d4;
  const T DT = C.DT, dth = C.dth, dh1 = C.dh1;

  // array-index accessors (stress_attenuation.cu:9-22, 148-158)
#define U1(i, j, k) u1[(long)(k) + y1 * (j) + s1 * (i)]
#define V1(i, j, k) v1[(long)(k) + y1 * (j) + s1 * (i)]
#define W1(i, j, k) w1[(long)(k) + y1 * (j) + s1 * (i)]
#define MFA(name, i, j) M.name[(long)(j) + ALIGN + (long)(i) * G.frow]

#pragma omp parallel for schedule(static)
  for (int i = s_i; i < e_i; ++i) {
    ftz_daz();
    for (int j = s_j; j < e_j; ++j) {
      for (int k = ALIGN + 3; k < ALIGN + nz; ++k) {
        const bool top = k >= nz - 6 + ALIGN;
        const int kb = top ? ALIGN + nz - 1 - k : 0;  // boundary table row
        const long pos = (long)i * s1 + (long)j * y1 + k;
        const long pos_ip1 = pos + s1, pos_jm1 = pos - y1, pos_km1 = pos - 1;
        const long pos_jk1 = pos - y1 - 1, pos_ik1 = pos + s1 - 1, pos_ijk = pos + s1 - y1,
                   pos_ijk1 = pos + s1 - y1 - 1;
        const T f_vx1 = d_vx1[pos], f_vx2 = d_vx2[pos], f_wwo = d_wwo[pos];
        const int f_ww = d_ww[pos];
        const T f_dcrj = dcrjx[i] * dcrjy[j] * dcrjz[k];
        T xl = 8.0f / (lam[pos] + lam[pos_ip1] + lam[pos_jm1] + lam[pos_ijk] + lam[pos_km1] +
                       lam[pos_ik1] + lam[pos_jk1] + lam[pos_ijk1]);
        T xm = 16.0f / (mu[pos] + mu[pos_ip1] + mu[pos_jm1] + mu[pos_ijk] + mu[pos_km1] +
                        mu[pos_ik1] + mu[pos_jk1] + mu[pos_ijk1]);
        T xmu1 = 2.0f / (mu[pos] + mu[pos_km1]);
        T xmu2 = (T)(2.0 / (double)(mu[pos] + mu[pos_jm1]));
        T xmu3 = (T)(2.0 / (double)(mu[pos] + mu[pos_ip1]));
        xl = xl + xm;
        T qpa = 0.0625f * (qp[pos] + qp[pos_ip1] + qp[pos_jm1] + qp[pos_ijk] + qp[pos_km1] +
                           qp[pos_ik1] + qp[pos_jk1] + qp[pos_ijk1]);
        const T qpaw = qweight(qpa, coeff, f_ww, f_wwo, true);
        T h = 0.0625f * (qs[pos] + qs[pos_ip1] + qs[pos_jm1] + qs[pos_ijk] + qs[pos_km1] +
                         qs[pos_ik1] + qs[pos_jk1] + qs[pos_ijk1]);
        const T hw = qweight(h, coeff, f_ww, f_wwo, false);
        T h1 = 0.250f * (qs[pos] + qs[pos_km1]);
        const T h1w = qweight(h1, coeff, f_ww, f_wwo, false);
        T h2 = 0.250f * (qs[pos] + qs[pos_jm1]);
        const T h2w = qweight(h2, coeff, f_ww, f_wwo, false);
        T h3 = 0.250f * (qs[pos] + qs[pos_ip1]);
        const T h3w = qweight(h3, coeff, f_ww, f_wwo, false);
        h = -xm * hw * dh1;
        h1 = -xmu1 * h1w * dh1;
        h2 = -xmu2 * h2w * dh1;
        h3 = -xmu3 * h3w * dh1;
        qpa = -qpaw * xl * dh1;
        xm = xm * dth;
        xmu1 = xmu1 * dth;
        xmu2 = xmu2 * dth;
        xmu3 = xmu3 * dth;
        xl = xl * dth;
        h = h * f_vx1;
        h1 = h1 * f_vx1;
        h2 = h2 * f_vx1;
        h3 = h3 * f_vx1;
        qpa = qpa * f_vx1;
        xm = xm + DT * h;
        xmu1 = xmu1 + DT * h1;
        xmu2 = xmu2 + DT * h2;
        xmu3 = xmu3 + DT * h3;
        const T vx1 = DT * (1 + f_vx2 * f_vx1);

        // z sums of a column at horizontal position (a, b):
        //  interior: 7-point stencil c[0..6] at k-3..k+3, in the given order;
        //  top:      table row kb, column m at level nz-1-m (array
        //            align+nz-1-m), in the order m = 8..0.
        auto zsum7 = [&](const T* c, const T* arr, int a, int b, const int* ord) {
          T acc = 0;
          for (int q = 0; q < 7; ++q) {
            const int m = ord[q];
            const T t = c[m] * arr[(long)(k + m - 3) + y1 * b + s1 * a];
            acc = q == 0 ? t : acc + t;
          }
          return acc;
        };
        auto zsumr = [&](const T(*c)[9], const T* arr, int a, int b) {
          T acc = 0;
          for (int m = 8; m >= 0; --m) {
            const T t = c[kb][m] * arr[(long)(ALIGN + nz - 1 - m) + y1 * b + s1 * a];
            acc = m == 8 ? t : acc + t;
          }
          return acc;
        };
        static const int ASC[7] = {0, 1, 2, 3, 4, 5, 6};
        static const int MID[7] = {3, 0, 1, 2, 4, 5, 6};
        const int kz = top ? (ALIGN + nz - 1 - kb) : k;  // == k
        (void)kz;

        T vs1, vs2, vs3, tmp, a1, f_r, f_rtmp, sv;
        // ---- xx, yy, zz ----
        {
          T Jii = MFA(f_c, i, j) * M.g3_c[k];
          Jii = (T)(1.0 * 1.0 / (double)Jii);
          T acc;
          if (!top) {
            // stress_attenuation.cu:336-372: groups in the order i-1, i, i+1, i+2
            acc = px4[0] * MFA(f1_1, i - 1, j) * zsum7(phdz4, u1, i - 1, j, ASC);
            acc = acc + px4[1] * MFA(f1_1, i, j) * zsum7(phdz4, u1, i, j, ASC);
            acc = acc + px4[2] * MFA(f1_1, i + 1, j) * zsum7(phdz4, u1, i + 1, j, ASC);
            acc = acc + px4[3] * MFA(f1_1, i + 2, j) * zsum7(phdz4, u1, i + 2, j, ASC);
          } else {
            // stress_attenuation.cu:1021-1066: groups in the order i, i-1, i+1, i+2
            acc = px4[1] * MFA(f1_1, i, j) * zsumr(phdz4r, u1, i, j);
            acc = acc + px4[0] * MFA(f1_1, i - 1, j) * zsumr(phdz4r, u1, i - 1, j);
            acc = acc + px4[2] * MFA(f1_1, i + 1, j) * zsumr(phdz4r, u1, i + 1, j);
            acc = acc + px4[3] * MFA(f1_1, i + 2, j) * zsumr(phdz4r, u1, i + 2, j);
          }
          vs1 = dx4[1] * U1(i, j, k) + dx4[0] * U1(i - 1, j, k) + dx4[2] * U1(i + 1, j, k) +
                dx4[3] * U1(i + 2, j, k) - Jii * M.g_c[k] * acc;
          if (!top) {
            acc = phy4[2] * MFA(f2_2, i, j) * zsum7(phdz4, v1, i, j, ASC);
            acc = acc + phy4[0] * MFA(f2_2, i, j - 2) * zsum7(phdz4, v1, i, j - 2, ASC);
            acc = acc + phy4[1] * MFA(f2_2, i, j - 1) * zsum7(phdz4, v1, i, j - 1, ASC);
            acc = acc + phy4[3] * MFA(f2_2, i, j + 1) * zsum7(phdz4, v1, i, j + 1, ASC);
          } else {
            acc = phy4[2] * MFA(f2_2, i, j) * zsumr(phdz4r, v1, i, j);
            acc = acc + phy4[0] * MFA(f2_2, i, j - 2) * zsumr(phdz4r, v1, i, j - 2);
            acc = acc + phy4[1] * MFA(f2_2, i, j - 1) * zsumr(phdz4r, v1, i, j - 1);
            acc = acc + phy4[3] * MFA(f2_2, i, j + 1) * zsumr(phdz4r, v1, i, j + 1);
          }
          vs2 = dhy4[2] * V1(i, j, k) + dhy4[0] * V1(i, j - 2, k) + dhy4[1] * V1(i, j - 1, k) +
                dhy4[3] * V1(i, j + 1, k) - Jii * M.g_c[k] * acc;
          if (!top) {
            vs3 = Jii * (dhz4[2] * W1(i, j, k) + dhz4[0] * W1(i, j, k - 2) +
                         dhz4[1] * W1(i, j, k - 1) + dhz4[3] * W1(i, j, k + 1));
          } else {
            T zs = dhz4r[kb][7] * W1(i, j, ALIGN + nz - 8);
            for (int m = 6; m >= 0; --m) zs = zs + dhz4r[kb][m] * W1(i, j, ALIGN + nz - 1 - m);
            vs3 = Jii * zs;
          }
        }
        tmp = xl * (vs1 + vs2 + vs3);
        a1 = qpa * (vs1 + vs2 + vs3);
        tmp = tmp + DT * a1;

        f_r = r1[pos];
        f_rtmp = -h * (vs2 + vs3) + a1;
        sv = xx[pos] + tmp - xm * (vs2 + vs3) + vx1 * f_r;
        r1[pos] = f_vx2 * f_r + f_wwo * f_rtmp;
        f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
        xx[pos] = (sv + DT * f_rtmp) * f_dcrj;

        f_r = r2[pos];
        f_rtmp = -h * (vs1 + vs3) + a1;
        sv = (yy[pos] + tmp - xm * (vs1 + vs3) + vx1 * f_r) * f_dcrj;
        r2[pos] = f_vx2 * f_r + f_wwo * f_rtmp;
        f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
        yy[pos] = (sv + DT * f_rtmp) * f_dcrj;

        f_r = r3[pos];
        f_rtmp = -h * (vs1 + vs2) + a1;
        sv = (zz[pos] + tmp - xm * (vs1 + vs2) + vx1 * f_r) * f_dcrj;
        r3[pos] = f_vx2 * f_r + f_wwo * f_rtmp;
        f_rtmp = f_rtmp * (f_wwo - 1.0f) + f_vx2 * f_r * (1.0f - f_vx1);
        zz[pos] = (sv + DT * f_rtmp) * f_dcrj;

        // ---- xy ----
        {
          // the interior kernel reads g3_c six levels up (stress_attenuation.cu:473)
          T J12i = MFA(f, i, j) * (top ? M.g3_c[k] : M.g3_c[k + 6]);
          J12i = (T)(1.0 / (double)J12i);
          T acc;
          if (!top) {
            acc = py4[1] * MFA(f2_1, i, j) * zsum7(phdz4, u1, i, j, MID);
            acc = acc + py4[0] * MFA(f2_1, i, j - 1) * zsum7(phdz4, u1, i, j - 1, MID);
            acc = acc + py4[2] * MFA(f2_1, i, j + 1) * zsum7(phdz4, u1, i, j + 1, MID);
            acc = acc + py4[3] * MFA(f2_1, i, j + 2) * zsum7(phdz4, u1, i, j + 2, MID);
          } else {
            acc = py4[1] * MFA(f2_1, i, j) * zsumr(phdz4r, u1, i, j);
            acc = acc + py4[0] * MFA(f2_1, i, j - 1) * zsumr(phdz4r, u1, i, j - 1);
            acc = acc + py4[2] * MFA(f2_1, i, j + 1) * zsumr(phdz4r, u1, i, j + 1);
            acc = acc + py4[3] * MFA(f2_1, i, j + 2) * zsumr(phdz4r, u1, i, j + 2);
          }
          vs1 = dy4[1] * U1(i, j, k) + dy4[0] * U1(i, j - 1, k) + dy4[2] * U1(i, j + 1, k) +
                dy4[3] * U1(i, j + 2, k) - J12i * M.g_c[k] * acc;
          if (!top) {
            acc = phx4[2] * MFA(f1_2, i, j) * zsum7(phdz4, v1, i, j, MID);
            acc = acc + phx4[0] * MFA(f1_2, i - 2, j) * zsum7(phdz4, v1, i - 2, j, MID);
            acc = acc + phx4[1] * MFA(f1_2, i - 1, j) * zsum7(phdz4, v1, i - 1, j, MID);
            acc = acc + phx4[3] * MFA(f1_2, i + 1, j) * zsum7(phdz4, v1, i + 1, j, MID);
          } else {
            acc = phx4[2] * MFA(f1_2, i, j) * zsumr(phdz4r, v1, i, j);
            acc = acc + phx4[0] * MFA(f1_2, i - 2, j) * zsumr(phdz4r, v1, i - 2, j);
            acc = acc + phx4[1] * MFA(f1_2, i - 1, j) * zsumr(phdz4r, v1, i - 1, j);
            acc = acc + phx4[3] * MFA(f1_2, i + 1, j) * zsumr(phdz4r, v1, i + 1, j);
          }
          vs2 = dhx4[2] * V1(i, j, k) + dhx4[0] * V1(i - 2, j, k) + dhx4[1] * V1(i - 1, j, k) +
                dhx4[3] * V1(i + 1, j, k) - J12i * M.g_c[k] * acc;
        }
        f_r = r4[pos];
        f_rtmp = h1 * (vs1 + vs2);
        sv = xy[pos] + xmu1 * (vs1 + vs2) + vx1 * f_r;
        r4[pos] = f_vx2 * f_r + f_wwo * f_rtmp;
        f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
        xy[pos] = (sv + DT * f_rtmp) * f_dcrj;

        // ---- xz ----
        {
          T J13i = MFA(f_1, i, j) * M.g3[k];
          J13i = (T)(1.0 * 1.0 / (double)J13i);
          if (!top) {
            vs1 = J13i * (dz4[1] * U1(i, j, k) + dz4[0] * U1(i, j, k - 1) +
                          dz4[2] * U1(i, j, k + 1) + dz4[3] * U1(i, j, k + 2));
          } else {
            T zs = dz4r[kb][6] * U1(i, j, ALIGN + nz - 7);
            for (int m = 5; m >= 0; --m) zs = zs + dz4r[kb][m] * U1(i, j, ALIGN + nz - 1 - m);
            vs1 = J13i * zs;
          }
          T acc;
          if (!top) {
            acc = phx4[2] * MFA(f1_c, i, j) * zsum7(pdhz4, w1, i, j, MID);
            acc = acc + phx4[0] * MFA(f1_c, i - 2, j) * zsum7(pdhz4, w1, i - 2, j, MID);
            acc = acc + phx4[1] * MFA(f1_c, i - 1, j) * zsum7(pdhz4, w1, i - 1, j, MID);
            acc = acc + phx4[3] * MFA(f1_c, i + 1, j) * zsum7(pdhz4, w1, i + 1, j, MID);
          } else {
            acc = phx4[2] * MFA(f1_c, i, j) * zsumr(pdhz4r, w1, i, j);
            acc = acc + phx4[0] * MFA(f1_c, i - 2, j) * zsumr(pdhz4r, w1, i - 2, j);
            acc = acc + phx4[1] * MFA(f1_c, i - 1, j) * zsumr(pdhz4r, w1, i - 1, j);
            acc = acc + phx4[3] * MFA(f1_c, i + 1, j) * zsumr(pdhz4r, w1, i + 1, j);
          }
          vs2 = dhx4[2] * W1(i, j, k) + dhx4[0] * W1(i - 2, j, k) + dhx4[1] * W1(i - 1, j, k) +
                dhx4[3] * W1(i + 1, j, k) - J13i * M.g[k] * acc;
        }
        f_r = r5[pos];
        f_rtmp = h2 * (vs1 + vs2);
        sv = xz[pos] + xmu2 * (vs1 + vs2) + vx1 * f_r;
        r5[pos] = f_vx2 * f_r + f_wwo * f_rtmp;
        f_rtmp = f_rtmp * (f_wwo - 1.0f) + f_vx2 * f_r * (1.0f - f_vx1);
        xz[pos] = (sv + DT * f_rtmp) * f_dcrj;

        // ---- yz ----
        {
          T J23i = MFA(f_2, i, j) * M.g3[k];
          J23i = (T)(1.0 * 1.0 / (double)J23i);
          if (!top) {
            vs1 = J23i * (dz4[1] * V1(i, j, k) + dz4[0] * V1(i, j, k - 1) +
                          dz4[2] * V1(i, j, k + 1) + dz4[3] * V1(i, j, k + 2));
          } else {
            T zs = dz4r[kb][6] * V1(i, j, ALIGN + nz - 7);
            for (int m = 5; m >= 0; --m) zs = zs + dz4r[kb][m] * V1(i, j, ALIGN + nz - 1 - m);
            vs1 = J23i * zs;
          }
          T acc;
          if (!top) {
            acc = py4[1] * MFA(f2_c, i, j) * zsum7(pdhz4, w1, i, j, MID);
            acc = acc + py4[0] * MFA(f2_c, i, j - 1) * zsum7(pdhz4, w1, i, j - 1, MID);
            acc = acc + py4[2] * MFA(f2_c, i, j + 1) * zsum7(pdhz4, w1, i, j + 1, MID);
            acc = acc + py4[3] * MFA(f2_c, i, j + 2) * zsum7(pdhz4, w1, i, j + 2, MID);
          } else {
            acc = py4[1] * MFA(f2_c, i, j) * zsumr(pdhz4r, w1, i, j);
            acc = acc + py4[0] * MFA(f2_c, i, j - 1) * zsumr(pdhz4r, w1, i, j - 1);
            acc = acc + py4[2] * MFA(f2_c, i, j + 1) * zsumr(pdhz4r, w1, i, j + 1);
            acc = acc + py4[3] * MFA(f2_c, i, j + 2) * zsumr(pdhz4r, w1, i, j + 2);
          }
          vs2 = dy4[1] * W1(i, j, k) + dy4[0] * W1(i, j - 1, k) + dy4[2] * W1(i, j + 1, k) +
                dy4[3] * W1(i, j + 2, k) - J23i * M.g[k] * acc;
        }
        f_r = r6[pos];
        f_rtmp = h3 * (vs1 + vs2);
        sv = yz[pos] + xmu3 * (vs1 + vs2) + vx1 * f_r;
        r6[pos] = f_vx2 * f_r + f_wwo * f_rtmp;
        f_rtmp = f_rtmp * (f_wwo - 1.0f) + f_vx2 * f_r * (1.0f - f_vx1);
        yz[pos] = (sv + DT * f_rtmp) * f_dcrj;
      }
    }
  }
#undef U1
#undef V1
#undef W1
#undef MFA
}

// ---------------------------------------------------------------------------
// Metrics.  metrics_build_f (metrics.c:35-41, 141-170) with the kernels of
// metrics/kernel.c:3-124: over topographic (i, j) in [-6, n+6) in each
// direction, f_1 = P_y f, f_2 = P_x f, f_c = P_y P_x f; f1_1 = D_x f_c,
// f1_2 = D_x f, f1_c = D_x f_1, f2_1 = D_y f, f2_2 = D_y f_c, f2_c = D_y f_2
// (derivatives scaled by hi = 1/h).  f holds the geometry on [-8, n+8).
template <typename T>
void metrics_f(T* const* mf, int nx, int ny, double DH) {
  const TGrid G = make_tgrid(nx, ny, 1);
  T* f = mf[0];
  T *f_1 = mf[1], *f_2 = mf[2], *f_c = mf[3], *f1_1 = mf[4], *f1_2 = mf[5], *f1_c = mf[6],
    *f2_1 = mf[7], *f2_2 = mf[8], *f2_c = mf[9];
  const T hi = (T)(1.0 / (double)(T)DH);  // f_grid_t.hi (metrics.c:24)
  T p[4], d[4];
  for (int q = 0; q < 4; ++q) {
    p[q] = (T)P4[q];
    d[q] = (T)D4[q];
  }
  const int i0 = -NGSL + 2, i1 = NGSL + nx - 2, j0 = -NGSL + 2, j1 = NGSL + ny - 2;
#define FA(arr, a, b) arr[G.f(a, b)]
  // interpolants first (metrics_interpolate_f), then derivatives
  for (int a = i0; a < i1; ++a)
    for (int b = j0; b < j1; ++b) {
      FA(f_1, a, b) = p[0] * FA(f, a, b - 2) + p[1] * FA(f, a, b - 1) + p[2] * FA(f, a, b) +
                      p[3] * FA(f, a, b + 1);
    }
  for (int a = i0; a < i1; ++a)
    for (int b = j0; b < j1; ++b) {
      FA(f_2, a, b) = p[0] * FA(f, a - 1, b) + p[1] * FA(f, a, b) + p[2] * FA(f, a + 1, b) +
                      p[3] * FA(f, a + 2, b);
    }
  for (int a = i0; a < i1; ++a)
    for (int b = j0; b < j1; ++b) {
      T acc = 0;
      for (int q = 0; q < 4; ++q) {
        const int bb = b - 2 + q;
        const T r = p[0] * FA(f, a - 1, bb) + p[1] * FA(f, a, bb) + p[2] * FA(f, a + 1, bb) +
                    p[3] * FA(f, a + 2, bb);
        acc = q == 0 ? p[0] * r : acc + p[q] * r;
      }
      FA(f_c, a, b) = acc;
    }
  // D_x at edge-1 positions: taps a-2 .. a+1 (diff_1_1); edge-2: a-1 .. a+2
  // (diff_1_2); D_y: taps b-2 .. b+1 (diff_2_1), b-1 .. b+2 (diff_2_2).
  auto dx11 = [&](const T* src, int a, int b) {
    return hi * (d[0] * FA(src, a - 2, b) + d[1] * FA(src, a - 1, b) + d[2] * FA(src, a, b) +
                 d[3] * FA(src, a + 1, b));
  };
  auto dx12 = [&](const T* src, int a, int b) {
    return hi * (d[0] * FA(src, a - 1, b) + d[1] * FA(src, a, b) + d[2] * FA(src, a + 1, b) +
                 d[3] * FA(src, a + 2, b));
  };
  auto dy21 = [&](const T* src, int a, int b) {
    return hi * (d[0] * FA(src, a, b - 2) + d[1] * FA(src, a, b - 1) + d[2] * FA(src, a, b) +
                 d[3] * FA(src, a, b + 1));
  };
  auto dy22 = [&](const T* src, int a, int b) {
    return hi * (d[0] * FA(src, a, b - 1) + d[1] * FA(src, a, b) + d[2] * FA(src, a, b + 1) +
                 d[3] * FA(src, a, b + 2));
  };
  for (int a = i0; a < i1; ++a)
    for (int b = j0; b < j1; ++b) {
      FA(f1_1, a, b) = dx11(f_c, a, b);
      FA(f1_2, a, b) = dx12(f, a, b);
      FA(f1_c, a, b) = dx12(f_1, a, b);
      FA(f2_1, a, b) = dy21(f, a, b);
      FA(f2_2, a, b) = dy22(f_c, a, b);
      FA(f2_c, a, b) = dy21(f_2, a, b);
    }
#undef FA
}

// metrics_build_g (metrics.c:223-229, 293-318) with metrics/kernel.c:126-292:
// g_c = P g, g3 = D g_c (g3(nz-1) = 1), g3_c = D g with the boundary rows of
// the tables.  With fill, g is first set as geom_no_grid_stretching does
// (geometry.c:33-42, grid_3d.c:258-278): g(k) = h' k for k < nz-1 and 0 at
// the top ghost, h' = (_prec)(1.0/hi), hi = (_prec)(1.0/h).
template <typename T>
void metrics_g(T* g, T* g_c, T* g3, T* g3_c, int nz, double DH, bool fill) {
  const T hi = (T)(1.0 / (double)(T)DH);
  const T hp = (T)(1.0 / (double)hi);
  T* G = g + ALIGN;
  T* Gc = g_c + ALIGN;
  T* G3 = g3 + ALIGN;
  T* G3c = g3_c + ALIGN;
  if (fill) {
    for (int k = 0; k < nz; ++k) G[k] = (T)((double)hp * ((double)k - 0.5 * 0 - 0));
    G[nz - 1] = (T)0.0;
  }
  const T p[4] = {(T)P4[0], (T)P4[1], (T)P4[2], (T)P4[3]};
  const T d[4] = {(T)D4[0], (T)D4[1], (T)D4[2], (T)D4[3]};
  // g_c
  for (int k = 0; k < 6; ++k) {
    T acc = (T)TM_PHZL[k][0] * G[0];
    for (int m = 1; m < 7; ++m) acc = acc + (T)TM_PHZL[k][m] * G[m];
    Gc[k] = acc;
  }
  for (int k = 0; k < nz - 12; ++k)
    Gc[k + 6] = p[0] * G[k + 4] + p[1] * G[k + 5] + p[2] * G[k + 6] + p[3] * G[k + 7];
  for (int k = 0; k < 6; ++k) {
    T acc = (T)TM_PHZR[k][7] * G[nz - 8];
    for (int m = 6; m >= 0; --m) acc = acc + (T)TM_PHZR[k][m] * G[nz - 1 - m];
    Gc[nz - 1 - k] = acc;
  }
  // g3 = D g_c at the nodes
  for (int k = 0; k < 6; ++k) {
    T acc = (T)TM_DZL[k][0] * Gc[0];
    for (int m = 1; m < 8; ++m) acc = acc + (T)TM_DZL[k][m] * Gc[m];
    G3[k] = hi * acc;
  }
  for (int k = 0; k < nz - 12; ++k)
    G3[k + 6] = hi * (d[0] * Gc[k + 5] + d[1] * Gc[k + 6] + d[2] * Gc[k + 7] + d[3] * Gc[k + 8]);
  for (int k = 0; k < 6; ++k) {
    T acc = (T)TM_DZR[k][6] * Gc[nz - 7];
    for (int m = 5; m >= 0; --m) acc = acc + (T)TM_DZR[k][m] * Gc[nz - 1 - m];
    G3[nz - 1 - k] = hi * acc;
  }
  G3[nz - 1] = (T)1.0;
  // g3_c = D g at the cell centres
  for (int k = 0; k < 6; ++k) {
    T acc = (T)TM_DHZL[k][0] * G[0];
    for (int m = 1; m < 7; ++m) acc = acc + (T)TM_DHZL[k][m] * G[m];
    G3c[k] = hi * acc;
  }
  for (int k = 0; k < nz - 12; ++k)
    G3c[k + 6] = hi * (d[0] * G[k + 4] + d[1] * G[k + 5] + d[2] * G[k + 6] + d[3] * G[k + 7]);
  for (int k = 0; k < 6; ++k) {
    T acc = (T)TM_DHZR[k][7] * G[nz - 8];
    for (int m = 6; m >= 0; --m) acc = acc + (T)TM_DHZR[k][m] * G[nz - 1 - m];
    G3c[nz - 1 - k] = hi * acc;
  }
}

}  // namespace

extern "C" {

// geom_custom (geometry.c:188-210): f = 1.0 + elevation * normalize on the
// f grid [-8, n+8)^2, normalize = 1.0 / h / (nz - 2) (in double, stored as
// _prec); f holds the elevation on entry.
#define AWP_ORACLE_GEOM(SUF, T)                                                     \
  void awp_oracle_topo_geom_##SUF(T* f, int nx, int ny, int nz, double DH) {        \
    ftz_daz();                                                                      \
    const TGrid G = make_tgrid(nx, ny, 1);                                          \
    const T normalize = (T)(1.0 / (double)(T)DH / (double)(nz - 2));                \
    for (int i = -NGSL; i < nx + NGSL; ++i)                                         \
      for (int j = -NGSL; j < ny + NGSL; ++j) {                                     \
        const long p = G.f(i, j);                                                   \
        f[p] = (T)(1.0 + (double)(f[p] * normalize));                               \
      }                                                                             \
  }
AWP_ORACLE_GEOM(f32, float)
AWP_ORACLE_GEOM(f64, double)

#define AWP_ORACLE_METRICS(SUF, T)                                                  \
  void awp_oracle_topo_metrics_f_##SUF(T* const* mf, int nx, int ny, double DH) {   \
    ftz_daz();                                                                      \
    metrics_f<T>(mf, nx, ny, DH);                                                   \
  }                                                                                 \
  void awp_oracle_topo_metrics_g_##SUF(T* g, T* g_c, T* g3, T* g3_c, int nz,        \
                                       double DH, int fill) {                       \
    ftz_daz();                                                                      \
    metrics_g<T>(g, g_c, g3, g3_c, nz, DH, fill != 0);                              \
  }
AWP_ORACLE_METRICS(f32, float)
AWP_ORACLE_METRICS(f64, double)

// mf: f, f_1, f_2, f_c, f1_1, f1_2, f1_c, f2_1, f2_2, f2_c; mg: g, g3, g3_c, g_c.
#define AWP_ORACLE_TOPO_VEL(SUF, T)                                                      \
  void awp_oracle_topo_vel_##SUF(T* u1, T* u2, T* u3, const T* dcrjx, const T* dcrjy,    \
                                 const T* dcrjz, const T* const* mf, const T* const* mg, \
                                 const T* rho, const T* s11, const T* s12, const T* s13, \
                                 const T* s22, const T* s23, const T* s33, T a, T nu,    \
                                 int nx, int ny, int nz, int bi, int bj, int ei,         \
                                 int ej) {                                               \
    ftz_daz();                                                                           \
    const TGrid G = make_tgrid(nx, ny, nz);                                              \
    const Metrics<T> M = metrics_from<T>(mf, mg);                                        \
    topo_vel<T>(u1, u2, u3, dcrjx, dcrjy, dcrjz, M, rho, s11, s12, s13, s22, s23, s33,   \
                a, nu, G, bi, bj, ei, ej);                                               \
  }
AWP_ORACLE_TOPO_VEL(f32, float)
AWP_ORACLE_TOPO_VEL(f64, double)

#define AWP_ORACLE_TOPO_STR(SUF, T)                                                         \
  void awp_oracle_topo_str_##SUF(T* xx, T* yy, T* zz, T* xy, T* xz, T* yz, T* r1, T* r2,    \
                                 T* r3, T* r4, T* r5, T* r6, const T* u1, const T* v1,      \
                                 const T* w1, const T* const* mf, const T* const* mg,       \
                                 const T* lam, const T* mu, const T* qp, const T* coeff,    \
                                 const T* qs, const T* dcrjx, const T* dcrjy,               \
                                 const T* dcrjz, const T* vx1, const T* vx2, const int* ww, \
                                 const T* wwo, int nx, int ny, int nz, double DH,           \
                                 double DT, int s_i, int e_i, int s_j, int e_j) {           \
    ftz_daz();                                                                              \
    const TGrid G = make_tgrid(nx, ny, nz);                                                 \
    const Metrics<T> M = metrics_from<T>(mf, mg);                                           \
    const TopoConst<T> C = topo_constants<T>(DH, DT);                                       \
    topo_str<T>(xx, yy, zz, xy, xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1, M, lam, mu, qp, \
                coeff, qs, dcrjx, dcrjy, dcrjz, vx1, vx2, ww, wwo, G, C, s_i, e_i, s_j,    \
                e_j);                                                                       \
  }
AWP_ORACLE_TOPO_STR(f32, float)
AWP_ORACLE_TOPO_STR(f64, double)

}  // extern "C"
