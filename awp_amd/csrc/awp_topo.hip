// Curvilinear topography: the velocity and stress (with coarse-grained Q)
// updates on a grid z = g(r) f(x, y) that follows the free surface, and the
// builders of the metric arrays f*, g*.
//
//   topo_velocity_interior_H  dtopo_vel_111 (levels 3 <= k < nz-6) and
//                             dtopo_vel_112 (top six levels), reference
//                             src/topography/velocity.cu:28-268,
//                             kernels/velocity_unroll.cu, optimized_velocity.cu:1040-1627
//   topo_stress_interior_H    dtopo_str_111 / dtopo_str_112,
//                             src/topography/stress_attenuation.cu:25-130,
//                             kernels/stress_attenuation.cu:74-1410
//   awp_topo_metrics_f/_g     metrics_build_f/_g, metrics/metrics.c:35-41,
//                             223-229 with metrics/kernel.c:3-292
//   awp_topo_geom_custom      geom_custom, geometry/geometry.c:188-210
//
// Lane map: one lane per level k (64 consecutive levels per wave, so every
// field access of a wave is one contiguous 256-byte run), waves along j,
// workgroups along i; the top six levels run in kernels of their own.
// Every expression keeps the reference's operand order and literal types
// (double where it writes double literals), so results match the CPU
// oracle bit for bit.
#include "awp_common.h"
#include "awp_store.h"
#include "awp_topo_coef.h"
#include "../../include/awp_hip.h"

namespace {

// float tables, as the reference's `const float` initialisers compute them
template <int R, int C>
struct FTab {
  float v[R][C];
};
template <int R, int C>
constexpr FTab<R, C> ftab(const double (&a)[R][C]) {
  FTab<R, C> t{};
  for (int r = 0; r < R; ++r)
    for (int c = 0; c < C; ++c) t.v[r][c] = (float)a[r][c];
  return t;
}

__constant__ FTab<6, 9> c_vel_dhpz4r = ftab(TV_DHPZ4R);
__constant__ FTab<6, 8> c_vel_dhz4r = ftab(TV_DHZ4R);
__constant__ FTab<6, 9> c_vel_dphz4r = ftab(TV_DPHZ4R);
__constant__ FTab<6, 7> c_vel_dz4r = ftab(TV_DZ4R);
__constant__ FTab<6, 9> c_str_phdz4r = ftab(TS_PHDZ4R);
__constant__ FTab<6, 7> c_str_dz4r = ftab(TS_DZ4R);
__constant__ FTab<6, 8> c_str_dhz4r = ftab(TS_DHZ4R);
__constant__ FTab<6, 9> c_str_pdhz4r = ftab(TS_PDHZ4R);

#define P4F(q) ((float)TOPO_P4[q])
#define D4F(q) ((float)TOPO_D4[q])

// Geometry of one grid: wavefield strides (32-bit: checked on the host) and
// the metric row length.
struct TGeom {
  int nx, ny, nz;
  int yl, sl, frow;
  int bi, bj, ei, ej;  // bounds of the launch (topographic for velocity,
                       // array indices for stress)
};

template <class ST>
struct TField {
  const typename ST::T* p;
  float s;
  __device__ __forceinline__ float operator[](int idx) const { return ST::ld(p, idx, s); }
};

struct TMetrics {
  const float *f, *f_1, *f_2, *f_c, *f1_1, *f1_2, *f1_c, *f2_1, *f2_2, *f2_c;
  const float *g, *g3, *g3_c, *g_c;
};

template <class ST>
struct TopoVelArgs {
  typename ST::T *u1, *u2, *u3;
  TField<ST> s11, s12, s13, s22, s23, s33;
  const float *rho, *dcrjx, *dcrjy, *dcrjz;
  TMetrics M;
  float iu1, iu2, iu3, su1, su2, su3;  // velocity scales (decode / encode)
  float a, nu;
};

__device__ __forceinline__ float dbl_quot(float num, float den) {
  return (float)((double)num * 1.0 / (double)den);
}

// ---------------------------------------------------------------------------
// Velocity, interior levels: dtopo_vel_111_unroll (velocity_unroll.cu:1-461).
template <class ST>
__global__ __launch_bounds__(256) void topo_vel_int_kernel(const TopoVelArgs<ST> A,
                                                           const TGeom G) {
  const int k = (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = G.bj + (int)(blockIdx.y * 4 + threadIdx.y);
  const int i = G.bi + (int)blockIdx.z;
  if (k < 3 || k >= G.nz - 6 || j >= G.ej || i >= G.ei) return;
  const int sl = G.sl, yl = G.yl, fr = G.frow;
  const int p = (k + AWP_ALIGN) + yl * (j + AWP_NGSL + 2) + sl * (i + AWP_NGSL + 2);
  const int q = (j + AWP_ALIGN + AWP_NGSL + 2) + (i + AWP_NGSL + 2) * fr;
  const int kg = k + AWP_ALIGN;
  const TMetrics& M = A.M;
  const float* rho = A.rho;

  const float c = 0.25f;
  const float rho1 = c * (rho[p - 1] + rho[p - yl - 1]) + c * (rho[p] + rho[p - yl]);
  const float rho2 = c * (rho[p - 1] + rho[p - sl - 1]) + c * (rho[p] + rho[p - sl]);
  const float rho3 = c * (rho[p] + rho[p - sl]) + c * (rho[p - yl] + rho[p - sl - yl]);
  float Ai1 = M.f_1[q] * M.g3_c[kg] * rho1;
  Ai1 = dbl_quot(A.nu, Ai1);
  float Ai2 = M.f_2[q] * M.g3_c[kg] * rho2;
  Ai2 = dbl_quot(A.nu, Ai2);
  float Ai3 = M.f_c[q] * M.g3[kg] * rho3;
  Ai3 = dbl_quot(A.nu, Ai3);
  const float f_dcrj =
      A.dcrjx[i + AWP_NGSL + 2] * A.dcrjy[j + AWP_NGSL + 2] * A.dcrjz[kg];

  // z derivatives of horizontal 4-point interpolants over k-3..k+3, terms in
  // the stencil order 3, 0, 1, 2, 4, 5, 6 (dhpz4 = dphz4: [3] is +0).
  constexpr float dhpz4[7] = {(float)TV_DHPZ4[0], (float)TV_DHPZ4[1], (float)TV_DHPZ4[2],
                              (float)TV_DHPZ4[3], (float)TV_DHPZ4[4], (float)TV_DHPZ4[5],
                              (float)TV_DHPZ4[6]};
  constexpr float dphz4[7] = {(float)TV_DPHZ4[0], (float)TV_DPHZ4[1], (float)TV_DPHZ4[2],
                              (float)TV_DPHZ4[3], (float)TV_DPHZ4[4], (float)TV_DPHZ4[5],
                              (float)TV_DPHZ4[6]};
  float z11 = 0.f, z12 = 0.f, z12x = 0.f, z22 = 0.f, z13 = 0.f, z23 = 0.f;
#pragma unroll
  for (int t = 0; t < 7; ++t) {
    const int m = t == 0 ? 3 : (t <= 3 ? t - 1 : t);
    const int o = p + m - 3, kk = kg + m - 3;
    const float cz = dhpz4[m] * M.g_c[kk];
    const float czp = dphz4[m] * M.g[kk];
    const float t11 = cz * (P4F(2) * A.s11[o] + P4F(0) * A.s11[o - 2 * sl] +
                            P4F(1) * A.s11[o - sl] + P4F(3) * A.s11[o + sl]);
    const float t12 = cz * (P4F(2) * A.s12[o] + P4F(0) * A.s12[o - 2 * yl] +
                            P4F(1) * A.s12[o - yl] + P4F(3) * A.s12[o + yl]);
    const float t12x = cz * (P4F(1) * A.s12[o] + P4F(0) * A.s12[o - sl] +
                             P4F(2) * A.s12[o + sl] + P4F(3) * A.s12[o + 2 * sl]);
    const float t22 = cz * (P4F(1) * A.s22[o] + P4F(0) * A.s22[o - yl] +
                            P4F(2) * A.s22[o + yl] + P4F(3) * A.s22[o + 2 * yl]);
    const float t13 = czp * (P4F(1) * A.s13[o] + P4F(0) * A.s13[o - sl] +
                             P4F(2) * A.s13[o + sl] + P4F(3) * A.s13[o + 2 * sl]);
    const float t23 = czp * (P4F(2) * A.s23[o] + P4F(0) * A.s23[o - 2 * yl] +
                             P4F(1) * A.s23[o - yl] + P4F(3) * A.s23[o + yl]);
    if (t == 0) {
      z11 = t11; z12 = t12; z12x = t12x; z22 = t22; z13 = t13; z23 = t23;
    } else {
      z11 = z11 + t11; z12 = z12 + t12; z12x = z12x + t12x; z22 = z22 + t22;
      z13 = z13 + t13; z23 = z23 + t23;
    }
  }
  const float g3c = M.g3_c[kg], g3 = M.g3[kg];
  float s;
  // u1
  s = D4F(2) * M.f_c[q] * g3c * A.s11[p];
  s = s + D4F(0) * M.f_c[q - 2 * fr] * g3c * A.s11[p - 2 * sl];
  s = s + D4F(1) * M.f_c[q - fr] * g3c * A.s11[p - sl];
  s = s + D4F(3) * M.f_c[q + fr] * g3c * A.s11[p + sl];
  s = s + D4F(2) * M.f[q] * g3c * A.s12[p];
  s = s + D4F(0) * M.f[q - 2] * g3c * A.s12[p - 2 * yl];
  s = s + D4F(1) * M.f[q - 1] * g3c * A.s12[p - yl];
  s = s + D4F(3) * M.f[q + 1] * g3c * A.s12[p + yl];
  s = s + D4F(2) * A.s13[p];
  s = s + D4F(0) * A.s13[p - 2];
  s = s + D4F(1) * A.s13[p - 1];
  s = s + D4F(3) * A.s13[p + 1];
  s = s - M.f1_1[q] * z11;
  s = s - M.f2_1[q] * z12;
  const float v1 = (A.a * ST::ld(A.u1, p, A.su1) + Ai1 * s) * f_dcrj;
  // u2
  s = D4F(2) * A.s23[p];
  s = s + D4F(0) * A.s23[p - 2];
  s = s + D4F(1) * A.s23[p - 1];
  s = s + D4F(3) * A.s23[p + 1];
  s = s + D4F(1) * M.f[q] * g3c * A.s12[p];
  s = s + D4F(0) * M.f[q - fr] * g3c * A.s12[p - sl];
  s = s + D4F(2) * M.f[q + fr] * g3c * A.s12[p + sl];
  s = s + D4F(3) * M.f[q + 2 * fr] * g3c * A.s12[p + 2 * sl];
  s = s + D4F(1) * M.f_c[q] * g3c * A.s22[p];
  s = s + D4F(0) * M.f_c[q - 1] * g3c * A.s22[p - yl];
  s = s + D4F(2) * M.f_c[q + 1] * g3c * A.s22[p + yl];
  s = s + D4F(3) * M.f_c[q + 2] * g3c * A.s22[p + 2 * yl];
  s = s - M.f1_2[q] * z12x;
  s = s - M.f2_2[q] * z22;
  const float v2 = (A.a * ST::ld(A.u2, p, A.su2) + Ai2 * s) * f_dcrj;
  // u3
  s = D4F(2) * M.f_2[q] * g3 * A.s23[p];
  s = s + D4F(0) * M.f_2[q - 2] * g3 * A.s23[p - 2 * yl];
  s = s + D4F(1) * M.f_2[q - 1] * g3 * A.s23[p - yl];
  s = s + D4F(3) * M.f_2[q + 1] * g3 * A.s23[p + yl];
  s = s + D4F(1) * M.f_1[q] * g3 * A.s13[p];
  s = s + D4F(0) * M.f_1[q - fr] * g3 * A.s13[p - sl];
  s = s + D4F(2) * M.f_1[q + fr] * g3 * A.s13[p + sl];
  s = s + D4F(3) * M.f_1[q + 2 * fr] * g3 * A.s13[p + 2 * sl];
  s = s + D4F(1) * A.s33[p];
  s = s + D4F(0) * A.s33[p - 1];
  s = s + D4F(2) * A.s33[p + 1];
  s = s + D4F(3) * A.s33[p + 2];
  s = s - M.f1_c[q] * z13;
  s = s - M.f2_c[q] * z23;
  const float v3 = (A.a * ST::ld(A.u3, p, A.su3) + Ai3 * s) * f_dcrj;
  ST::st(A.u1, p, v1, A.iu1);
  ST::st(A.u2, p, v2, A.iu2);
  ST::st(A.u3, p, v3, A.iu3);
}

// Velocity, top six levels k = nz-1-kk: dtopo_vel_112
// (optimized_velocity.cu:1040-1627); one lane per (kk, j).
template <class ST>
__global__ __launch_bounds__(192) void topo_vel_top_kernel(const TopoVelArgs<ST> A,
                                                           const TGeom G) {
  const int kk = (int)threadIdx.x;  // 0..5
  const int j = G.bj + (int)(blockIdx.y * 32 + threadIdx.y);
  const int i = G.bi + (int)blockIdx.z;
  if (kk >= 6 || j >= G.ej || i >= G.ei) return;
  const int nz = G.nz;
  const int sl = G.sl, yl = G.yl, fr = G.frow;
  const int k = nz - 1 - kk;
  const int col = yl * (j + AWP_NGSL + 2) + sl * (i + AWP_NGSL + 2) + AWP_ALIGN;  // level 0
  const int p = col + k;
  const int q = (j + AWP_ALIGN + AWP_NGSL + 2) + (i + AWP_NGSL + 2) * fr;
  const int kg = k + AWP_ALIGN;
  const TMetrics& M = A.M;
  const float* rho = A.rho;
  const int kb = col + nz - kk - 2;

  const float rho1 = (float)(0.25 * (double)(rho[kb] + rho[kb - yl]) +
                             0.25 * (double)(rho[kb + 1] + rho[kb + 1 - yl]));
  const float rho2 = (float)(0.25 * (double)(rho[kb] + rho[kb - sl]) +
                             0.25 * (double)(rho[kb + 1] + rho[kb + 1 - sl]));
  const float rho3 = (float)(0.25 * (double)(rho[kb + 1] + rho[kb + 1 - sl]) +
                             0.25 * (double)(rho[kb + 1 - yl] + rho[kb + 1 - sl - yl]));
  float Ai1 = M.f_1[q] * M.g3_c[kg] * rho1;
  Ai1 = dbl_quot(A.nu, Ai1);
  float Ai2 = M.f_2[q] * M.g3_c[kg] * rho2;
  Ai2 = dbl_quot(A.nu, Ai2);
  float Ai3 = M.f_c[q] * M.g3[kg] * rho3;
  Ai3 = dbl_quot(A.nu, Ai3);
  const float f_dcrj =
      A.dcrjx[i + AWP_NGSL + 2] * A.dcrjy[j + AWP_NGSL + 2] * A.dcrjz[kg];

  float z11 = 0.f, z12 = 0.f, z12x = 0.f, z22 = 0.f, z13 = 0.f, z23 = 0.f;
#pragma unroll
  for (int m = 8; m >= 0; --m) {
    const int o = col + nz - 1 - m, lg = AWP_ALIGN + nz - 1 - m;
    const float cz = c_vel_dhpz4r.v[kk][m] * M.g_c[lg];
    const float czp = c_vel_dphz4r.v[kk][m] * M.g[lg];
    const float t11 = cz * (P4F(2) * A.s11[o] + P4F(0) * A.s11[o - 2 * sl] +
                            P4F(1) * A.s11[o - sl] + P4F(3) * A.s11[o + sl]);
    const float t12 = cz * (P4F(2) * A.s12[o] + P4F(0) * A.s12[o - 2 * yl] +
                            P4F(1) * A.s12[o - yl] + P4F(3) * A.s12[o + yl]);
    const float t12x = cz * (P4F(1) * A.s12[o] + P4F(0) * A.s12[o - sl] +
                             P4F(2) * A.s12[o + sl] + P4F(3) * A.s12[o + 2 * sl]);
    const float t22 = cz * (P4F(1) * A.s22[o] + P4F(0) * A.s22[o - yl] +
                            P4F(2) * A.s22[o + yl] + P4F(3) * A.s22[o + 2 * yl]);
    const float t13 = czp * (P4F(1) * A.s13[o] + P4F(0) * A.s13[o - sl] +
                             P4F(2) * A.s13[o + sl] + P4F(3) * A.s13[o + 2 * sl]);
    const float t23 = czp * (P4F(2) * A.s23[o] + P4F(0) * A.s23[o - 2 * yl] +
                             P4F(1) * A.s23[o - yl] + P4F(3) * A.s23[o + yl]);
    if (m == 8) {
      z11 = t11; z12 = t12; z12x = t12x; z22 = t22; z13 = t13; z23 = t23;
    } else {
      z11 = z11 + t11; z12 = z12 + t12; z12x = z12x + t12x; z22 = z22 + t22;
      z13 = z13 + t13; z23 = z23 + t23;
    }
  }
  const float g3c = M.g3_c[kg], g3 = M.g3[kg];
  float s;
  // u1
  s = D4F(2) * M.f_c[q] * g3c * A.s11[p];
  s = s + D4F(0) * M.f_c[q - 2 * fr] * g3c * A.s11[p - 2 * sl];
  s = s + D4F(1) * M.f_c[q - fr] * g3c * A.s11[p - sl];
  s = s + D4F(3) * M.f_c[q + fr] * g3c * A.s11[p + sl];
  s = s + D4F(2) * M.f[q] * g3c * A.s12[p];
  s = s + D4F(0) * M.f[q - 2] * g3c * A.s12[p - 2 * yl];
  s = s + D4F(1) * M.f[q - 1] * g3c * A.s12[p - yl];
  s = s + D4F(3) * M.f[q + 1] * g3c * A.s12[p + yl];
#pragma unroll
  for (int m = 7; m >= 0; --m) s = s + c_vel_dhz4r.v[kk][m] * A.s13[col + nz - 1 - m];
  s = s - M.f1_1[q] * z11;
  s = s - M.f2_1[q] * z12;
  const float v1 = (A.a * ST::ld(A.u1, p, A.su1) + Ai1 * s) * f_dcrj;
  // u2: z terms first
  s = c_vel_dhz4r.v[kk][7] * A.s23[col + nz - 8];
#pragma unroll
  for (int m = 6; m >= 0; --m) s = s + c_vel_dhz4r.v[kk][m] * A.s23[col + nz - 1 - m];
  s = s + D4F(1) * M.f[q] * g3c * A.s12[p];
  s = s + D4F(0) * M.f[q - fr] * g3c * A.s12[p - sl];
  s = s + D4F(2) * M.f[q + fr] * g3c * A.s12[p + sl];
  s = s + D4F(3) * M.f[q + 2 * fr] * g3c * A.s12[p + 2 * sl];
  s = s + D4F(1) * M.f_c[q] * g3c * A.s22[p];
  s = s + D4F(0) * M.f_c[q - 1] * g3c * A.s22[p - yl];
  s = s + D4F(2) * M.f_c[q + 1] * g3c * A.s22[p + yl];
  s = s + D4F(3) * M.f_c[q + 2] * g3c * A.s22[p + 2 * yl];
  s = s - M.f1_2[q] * z12x;
  s = s - M.f2_2[q] * z22;
  const float v2 = (A.a * ST::ld(A.u2, p, A.su2) + Ai2 * s) * f_dcrj;
  // u3
  s = D4F(2) * M.f_2[q] * g3 * A.s23[p];
  s = s + D4F(0) * M.f_2[q - 2] * g3 * A.s23[p - 2 * yl];
  s = s + D4F(1) * M.f_2[q - 1] * g3 * A.s23[p - yl];
  s = s + D4F(3) * M.f_2[q + 1] * g3 * A.s23[p + yl];
  s = s + D4F(1) * M.f_1[q] * g3 * A.s13[p];
  s = s + D4F(0) * M.f_1[q - fr] * g3 * A.s13[p - sl];
  s = s + D4F(2) * M.f_1[q + fr] * g3 * A.s13[p + sl];
  s = s + D4F(3) * M.f_1[q + 2 * fr] * g3 * A.s13[p + 2 * sl];
#pragma unroll
  for (int m = 6; m >= 0; --m) s = s + c_vel_dz4r.v[kk][m] * A.s33[col + nz - 1 - m];
  s = s - M.f1_c[q] * z13;
  s = s - M.f2_c[q] * z23;
  const float v3 = (A.a * ST::ld(A.u3, p, A.su3) + Ai3 * s) * f_dcrj;
  ST::st(A.u1, p, v1, A.iu1);
  ST::st(A.u2, p, v2, A.iu2);
  ST::st(A.u3, p, v3, A.iu3);
}

// ---------------------------------------------------------------------------
// Stress with coarse-grained Q: dtopo_str_111 / dtopo_str_112.
template <class ST>
struct TopoStrArgs {
  typename ST::T *xx, *yy, *zz, *xy, *xz, *yz, *r1, *r2, *r3, *r4, *r5, *r6;
  TField<ST> u1, v1, w1;
  const float *lam, *mu, *qp, *coeff, *qs, *dcrjx, *dcrjy, *dcrjz, *vx1, *vx2, *wwo;
  const int* ww;
  TMetrics M;
  WaveScales W;
  float DT, dth, dh1;
};

__device__ __forceinline__ float topo_qweight(float q, const float* coeff, int ww, float wwo,
                                              bool dbl_poly) {
  float w;
  if (1.0f / (q * 2.0f) <= 200.0f) {
    if (dbl_poly)
      w = (float)((double)coeff[ww * 2 - 2] * (2. * (double)q) * (2. * (double)q) +
                  (double)coeff[ww * 2 - 1] * (2. * (double)q));
    else
      w = coeff[ww * 2 - 2] * (2.0f * q) * (2.0f * q) + coeff[ww * 2 - 1] * (2.0f * q);
  } else {
    w = 2.0f * wwo * q;
  }
  return w / wwo;
}

// The material averages, Q weights and the memory-variable recursion, shared
// by the interior and the top kernel (stress_attenuation.cu:200-327, 462-467,
// 534-540, 595-600, 651-656).
template <class ST>
__device__ __forceinline__ void topo_str_update(const TopoStrArgs<ST>& A, int pos, int sl, int yl,
                                                float f_dcrj, float nvs1, float nvs2,
                                                float nvs3, float xyv1, float xyv2,
                                                float xzv1, float xzv2, float yzv1,
                                                float yzv2) {
  const int pos_ip1 = pos + sl, pos_jm1 = pos - yl, pos_km1 = pos - 1;
  const int pos_jk1 = pos - yl - 1, pos_ik1 = pos + sl - 1, pos_ijk = pos + sl - yl,
            pos_ijk1 = pos + sl - yl - 1;
  const float* lam = A.lam;
  const float* mu = A.mu;
  const float* qp = A.qp;
  const float* qs = A.qs;
  const float f_vx1 = A.vx1[pos], f_vx2 = A.vx2[pos], f_wwo = A.wwo[pos];
  const int f_ww = A.ww[pos];
  float xl = 8.0f / (lam[pos] + lam[pos_ip1] + lam[pos_jm1] + lam[pos_ijk] + lam[pos_km1] +
                     lam[pos_ik1] + lam[pos_jk1] + lam[pos_ijk1]);
  float xm = 16.0f / (mu[pos] + mu[pos_ip1] + mu[pos_jm1] + mu[pos_ijk] + mu[pos_km1] +
                      mu[pos_ik1] + mu[pos_jk1] + mu[pos_ijk1]);
  float xmu1 = 2.0f / (mu[pos] + mu[pos_km1]);
  float xmu2 = (float)(2.0 / (double)(mu[pos] + mu[pos_jm1]));
  float xmu3 = (float)(2.0 / (double)(mu[pos] + mu[pos_ip1]));
  xl = xl + xm;
  float qpa = 0.0625f * (qp[pos] + qp[pos_ip1] + qp[pos_jm1] + qp[pos_ijk] + qp[pos_km1] +
                         qp[pos_ik1] + qp[pos_jk1] + qp[pos_ijk1]);
  const float qpaw = topo_qweight(qpa, A.coeff, f_ww, f_wwo, true);
  float h = 0.0625f * (qs[pos] + qs[pos_ip1] + qs[pos_jm1] + qs[pos_ijk] + qs[pos_km1] +
                       qs[pos_ik1] + qs[pos_jk1] + qs[pos_ijk1]);
  const float hw = topo_qweight(h, A.coeff, f_ww, f_wwo, false);
  float h1 = 0.250f * (qs[pos] + qs[pos_km1]);
  const float h1w = topo_qweight(h1, A.coeff, f_ww, f_wwo, false);
  float h2 = 0.250f * (qs[pos] + qs[pos_jm1]);
  const float h2w = topo_qweight(h2, A.coeff, f_ww, f_wwo, false);
  float h3 = 0.250f * (qs[pos] + qs[pos_ip1]);
  const float h3w = topo_qweight(h3, A.coeff, f_ww, f_wwo, false);
  const float DT = A.DT, dth = A.dth, dh1 = A.dh1;
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
  const float vx1 = DT * (1 + f_vx2 * f_vx1);
  const WaveScales& W = A.W;

  const float vs1 = nvs1, vs2 = nvs2, vs3 = nvs3;
  float tmp = xl * (vs1 + vs2 + vs3);
  const float a1 = qpa * (vs1 + vs2 + vs3);
  tmp = tmp + DT * a1;
  float f_r, f_rtmp, sv;

  f_r = ST::ld(A.r1, pos, W.s[AWP_SR1]);
  f_rtmp = -h * (vs2 + vs3) + a1;
  sv = ST::ld(A.xx, pos, W.s[AWP_SXX]) + tmp - xm * (vs2 + vs3) + vx1 * f_r;
  ST::st(A.r1, pos, f_vx2 * f_r + f_wwo * f_rtmp, W.inv[AWP_SR1]);
  f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
  ST::st(A.xx, pos, (sv + DT * f_rtmp) * f_dcrj, W.inv[AWP_SXX]);

  f_r = ST::ld(A.r2, pos, W.s[AWP_SR2]);
  f_rtmp = -h * (vs1 + vs3) + a1;
  sv = (ST::ld(A.yy, pos, W.s[AWP_SYY]) + tmp - xm * (vs1 + vs3) + vx1 * f_r) * f_dcrj;
  ST::st(A.r2, pos, f_vx2 * f_r + f_wwo * f_rtmp, W.inv[AWP_SR2]);
  f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
  ST::st(A.yy, pos, (sv + DT * f_rtmp) * f_dcrj, W.inv[AWP_SYY]);

  f_r = ST::ld(A.r3, pos, W.s[AWP_SR3]);
  f_rtmp = -h * (vs1 + vs2) + a1;
  sv = (ST::ld(A.zz, pos, W.s[AWP_SZZ]) + tmp - xm * (vs1 + vs2) + vx1 * f_r) * f_dcrj;
  ST::st(A.r3, pos, f_vx2 * f_r + f_wwo * f_rtmp, W.inv[AWP_SR3]);
  f_rtmp = f_rtmp * (f_wwo - 1.0f) + f_vx2 * f_r * (1.0f - f_vx1);
  ST::st(A.zz, pos, (sv + DT * f_rtmp) * f_dcrj, W.inv[AWP_SZZ]);

  f_r = ST::ld(A.r4, pos, W.s[AWP_SR4]);
  f_rtmp = h1 * (xyv1 + xyv2);
  sv = ST::ld(A.xy, pos, W.s[AWP_SXY]) + xmu1 * (xyv1 + xyv2) + vx1 * f_r;
  ST::st(A.r4, pos, f_vx2 * f_r + f_wwo * f_rtmp, W.inv[AWP_SR4]);
  f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
  ST::st(A.xy, pos, (sv + DT * f_rtmp) * f_dcrj, W.inv[AWP_SXY]);

  f_r = ST::ld(A.r5, pos, W.s[AWP_SR5]);
  f_rtmp = h2 * (xzv1 + xzv2);
  sv = ST::ld(A.xz, pos, W.s[AWP_SXZ]) + xmu2 * (xzv1 + xzv2) + vx1 * f_r;
  ST::st(A.r5, pos, f_vx2 * f_r + f_wwo * f_rtmp, W.inv[AWP_SR5]);
  f_rtmp = f_rtmp * (f_wwo - 1.0f) + f_vx2 * f_r * (1.0f - f_vx1);
  ST::st(A.xz, pos, (sv + DT * f_rtmp) * f_dcrj, W.inv[AWP_SXZ]);

  f_r = ST::ld(A.r6, pos, W.s[AWP_SR6]);
  f_rtmp = h3 * (yzv1 + yzv2);
  sv = ST::ld(A.yz, pos, W.s[AWP_SYZ]) + xmu3 * (yzv1 + yzv2) + vx1 * f_r;
  ST::st(A.r6, pos, f_vx2 * f_r + f_wwo * f_rtmp, W.inv[AWP_SR6]);
  f_rtmp = f_rtmp * (f_wwo - 1.0f) + f_vx2 * f_r * (1.0f - f_vx1);
  ST::st(A.yz, pos, (sv + DT * f_rtmp) * f_dcrj, W.inv[AWP_SYZ]);
}

// Interior stress levels align+3 <= k < align+nz-6 (array indices), one lane
// per k; i, j are array indices in [s_i, e_i) x [s_j, e_j).
template <class ST>
__global__ __launch_bounds__(256) void topo_str_int_kernel(const TopoStrArgs<ST> A,
                                                           const TGeom G) {
  const int k = (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = G.bj + (int)(blockIdx.y * 4 + threadIdx.y);
  const int i = G.bi + (int)blockIdx.z;
  if (k < AWP_ALIGN + 3 || k >= AWP_ALIGN + G.nz - 6 || j >= G.ej || i >= G.ei) return;
  const int sl = G.sl, yl = G.yl, fr = G.frow;
  const int pos = k + yl * j + sl * i;
  const int q = j + AWP_ALIGN + i * fr;
  const TMetrics& M = A.M;

  // 7-point z sum of a column at offset o, stencil c (the stress tables'
  // [3] is -0), terms in ascending order (ASC) or 3, 0, 1, 2, 4, 5, 6.
  constexpr float phdz4[7] = {(float)TS_PHDZ4[0], (float)TS_PHDZ4[1], (float)TS_PHDZ4[2],
                              (float)TS_PHDZ4[3], (float)TS_PHDZ4[4], (float)TS_PHDZ4[5],
                              (float)TS_PHDZ4[6]};
  constexpr float pdhz4[7] = {(float)TS_PDHZ4[0], (float)TS_PDHZ4[1], (float)TS_PDHZ4[2],
                              (float)TS_PDHZ4[3], (float)TS_PDHZ4[4], (float)TS_PDHZ4[5],
                              (float)TS_PDHZ4[6]};
  auto zs_asc = [&](const TField<ST>& fld, int o) {
    float acc = phdz4[0] * fld[o - 3];
#pragma unroll
    for (int m = 1; m < 7; ++m) acc = acc + phdz4[m] * fld[o + m - 3];
    return acc;
  };
  auto zs_mid = [&](const float (&c)[7], const TField<ST>& fld, int o) {
    float acc = c[3] * fld[o];
    acc = acc + c[0] * fld[o - 3];
    acc = acc + c[1] * fld[o - 2];
    acc = acc + c[2] * fld[o - 1];
    acc = acc + c[4] * fld[o + 1];
    acc = acc + c[5] * fld[o + 2];
    acc = acc + c[6] * fld[o + 3];
    return acc;
  };
  const float f_dcrj = A.dcrjx[i] * A.dcrjy[j] * A.dcrjz[k];
  float acc;
  // xx, yy, zz
  float Jii = M.f_c[q] * M.g3_c[k];
  Jii = (float)(1.0 * 1.0 / (double)Jii);
  acc = P4F(0) * M.f1_1[q - fr] * zs_asc(A.u1, pos - sl);
  acc = acc + P4F(1) * M.f1_1[q] * zs_asc(A.u1, pos);
  acc = acc + P4F(2) * M.f1_1[q + fr] * zs_asc(A.u1, pos + sl);
  acc = acc + P4F(3) * M.f1_1[q + 2 * fr] * zs_asc(A.u1, pos + 2 * sl);
  const float vs1 = D4F(1) * A.u1[pos] + D4F(0) * A.u1[pos - sl] + D4F(2) * A.u1[pos + sl] +
                    D4F(3) * A.u1[pos + 2 * sl] - Jii * M.g_c[k] * acc;
  acc = P4F(2) * M.f2_2[q] * zs_asc(A.v1, pos);
  acc = acc + P4F(0) * M.f2_2[q - 2] * zs_asc(A.v1, pos - 2 * yl);
  acc = acc + P4F(1) * M.f2_2[q - 1] * zs_asc(A.v1, pos - yl);
  acc = acc + P4F(3) * M.f2_2[q + 1] * zs_asc(A.v1, pos + yl);
  const float vs2 = D4F(2) * A.v1[pos] + D4F(0) * A.v1[pos - 2 * yl] + D4F(1) * A.v1[pos - yl] +
                    D4F(3) * A.v1[pos + yl] - Jii * M.g_c[k] * acc;
  const float vs3 = Jii * (D4F(2) * A.w1[pos] + D4F(0) * A.w1[pos - 2] +
                           D4F(1) * A.w1[pos - 1] + D4F(3) * A.w1[pos + 1]);
  // xy (the interior kernel reads g3_c six levels up, stress_attenuation.cu:473)
  float J12i = M.f[q] * M.g3_c[k + 6];
  J12i = (float)(1.0 / (double)J12i);
  acc = P4F(1) * M.f2_1[q] * zs_mid(phdz4, A.u1, pos);
  acc = acc + P4F(0) * M.f2_1[q - 1] * zs_mid(phdz4, A.u1, pos - yl);
  acc = acc + P4F(2) * M.f2_1[q + 1] * zs_mid(phdz4, A.u1, pos + yl);
  acc = acc + P4F(3) * M.f2_1[q + 2] * zs_mid(phdz4, A.u1, pos + 2 * yl);
  const float xyv1 = D4F(1) * A.u1[pos] + D4F(0) * A.u1[pos - yl] + D4F(2) * A.u1[pos + yl] +
                     D4F(3) * A.u1[pos + 2 * yl] - J12i * M.g_c[k] * acc;
  acc = P4F(2) * M.f1_2[q] * zs_mid(phdz4, A.v1, pos);
  acc = acc + P4F(0) * M.f1_2[q - 2 * fr] * zs_mid(phdz4, A.v1, pos - 2 * sl);
  acc = acc + P4F(1) * M.f1_2[q - fr] * zs_mid(phdz4, A.v1, pos - sl);
  acc = acc + P4F(3) * M.f1_2[q + fr] * zs_mid(phdz4, A.v1, pos + sl);
  const float xyv2 = D4F(2) * A.v1[pos] + D4F(0) * A.v1[pos - 2 * sl] +
                     D4F(1) * A.v1[pos - sl] + D4F(3) * A.v1[pos + sl] - J12i * M.g_c[k] * acc;
  // xz
  float J13i = M.f_1[q] * M.g3[k];
  J13i = (float)(1.0 * 1.0 / (double)J13i);
  const float xzv1 = J13i * (D4F(1) * A.u1[pos] + D4F(0) * A.u1[pos - 1] +
                             D4F(2) * A.u1[pos + 1] + D4F(3) * A.u1[pos + 2]);
  acc = P4F(2) * M.f1_c[q] * zs_mid(pdhz4, A.w1, pos);
  acc = acc + P4F(0) * M.f1_c[q - 2 * fr] * zs_mid(pdhz4, A.w1, pos - 2 * sl);
  acc = acc + P4F(1) * M.f1_c[q - fr] * zs_mid(pdhz4, A.w1, pos - sl);
  acc = acc + P4F(3) * M.f1_c[q + fr] * zs_mid(pdhz4, A.w1, pos + sl);
  const float xzv2 = D4F(2) * A.w1[pos] + D4F(0) * A.w1[pos - 2 * sl] +
                     D4F(1) * A.w1[pos - sl] + D4F(3) * A.w1[pos + sl] - J13i * M.g[k] * acc;
  // yz
  float J23i = M.f_2[q] * M.g3[k];
  J23i = (float)(1.0 * 1.0 / (double)J23i);
  const float yzv1 = J23i * (D4F(1) * A.v1[pos] + D4F(0) * A.v1[pos - 1] +
                             D4F(2) * A.v1[pos + 1] + D4F(3) * A.v1[pos + 2]);
  acc = P4F(1) * M.f2_c[q] * zs_mid(pdhz4, A.w1, pos);
  acc = acc + P4F(0) * M.f2_c[q - 1] * zs_mid(pdhz4, A.w1, pos - yl);
  acc = acc + P4F(2) * M.f2_c[q + 1] * zs_mid(pdhz4, A.w1, pos + yl);
  acc = acc + P4F(3) * M.f2_c[q + 2] * zs_mid(pdhz4, A.w1, pos + 2 * yl);
  const float yzv2 = D4F(1) * A.w1[pos] + D4F(0) * A.w1[pos - yl] + D4F(2) * A.w1[pos + yl] +
                     D4F(3) * A.w1[pos + 2 * yl] - J23i * M.g[k] * acc;
  topo_str_update<ST>(A, pos, sl, yl, f_dcrj, vs1, vs2, vs3, xyv1, xyv2, xzv1, xzv2, yzv1, yzv2);
}

// Top six stress levels k = align+nz-1-kb, kb = 0..5 (dtopo_str_112).
template <class ST>
__global__ __launch_bounds__(192) void topo_str_top_kernel(const TopoStrArgs<ST> A,
                                                           const TGeom G) {
  const int kb = (int)threadIdx.x;
  const int j = G.bj + (int)(blockIdx.y * 32 + threadIdx.y);
  const int i = G.bi + (int)blockIdx.z;
  if (kb >= 6 || j >= G.ej || i >= G.ei) return;
  const int nz = G.nz;
  const int sl = G.sl, yl = G.yl, fr = G.frow;
  const int k = AWP_ALIGN + nz - 1 - kb;
  const int col = yl * j + sl * i + AWP_ALIGN;  // level 0 of the column
  const int pos = col + nz - 1 - kb;
  const int q = j + AWP_ALIGN + i * fr;
  const TMetrics& M = A.M;

  // column sums over levels nz-9 .. nz-1, table row kb, terms m = 8 .. 0
  auto zr = [&](const FTab<6, 9>& tab, const TField<ST>& fld, int o) {
    float acc = tab.v[kb][8] * fld[o + nz - 9];
#pragma unroll
    for (int m = 7; m >= 0; --m) acc = acc + tab.v[kb][m] * fld[o + nz - 1 - m];
    return acc;
  };
  const float f_dcrj = A.dcrjx[i] * A.dcrjy[j] * A.dcrjz[k];
  float acc;
  float Jii = M.f_c[q] * M.g3_c[k];
  Jii = (float)(1.0 * 1.0 / (double)Jii);
  acc = P4F(1) * M.f1_1[q] * zr(c_str_phdz4r, A.u1, col);
  acc = acc + P4F(0) * M.f1_1[q - fr] * zr(c_str_phdz4r, A.u1, col - sl);
  acc = acc + P4F(2) * M.f1_1[q + fr] * zr(c_str_phdz4r, A.u1, col + sl);
  acc = acc + P4F(3) * M.f1_1[q + 2 * fr] * zr(c_str_phdz4r, A.u1, col + 2 * sl);
  const float vs1 = D4F(1) * A.u1[pos] + D4F(0) * A.u1[pos - sl] + D4F(2) * A.u1[pos + sl] +
                    D4F(3) * A.u1[pos + 2 * sl] - Jii * M.g_c[k] * acc;
  acc = P4F(2) * M.f2_2[q] * zr(c_str_phdz4r, A.v1, col);
  acc = acc + P4F(0) * M.f2_2[q - 2] * zr(c_str_phdz4r, A.v1, col - 2 * yl);
  acc = acc + P4F(1) * M.f2_2[q - 1] * zr(c_str_phdz4r, A.v1, col - yl);
  acc = acc + P4F(3) * M.f2_2[q + 1] * zr(c_str_phdz4r, A.v1, col + yl);
  const float vs2 = D4F(2) * A.v1[pos] + D4F(0) * A.v1[pos - 2 * yl] + D4F(1) * A.v1[pos - yl] +
                    D4F(3) * A.v1[pos + yl] - Jii * M.g_c[k] * acc;
  float zs = c_str_dhz4r.v[kb][7] * A.w1[col + nz - 8];
#pragma unroll
  for (int m = 6; m >= 0; --m) zs = zs + c_str_dhz4r.v[kb][m] * A.w1[col + nz - 1 - m];
  const float vs3 = Jii * zs;
  float J12i = M.f[q] * M.g3_c[k];
  J12i = (float)(1.0 * 1.0 / (double)J12i);
  acc = P4F(1) * M.f2_1[q] * zr(c_str_phdz4r, A.u1, col);
  acc = acc + P4F(0) * M.f2_1[q - 1] * zr(c_str_phdz4r, A.u1, col - yl);
  acc = acc + P4F(2) * M.f2_1[q + 1] * zr(c_str_phdz4r, A.u1, col + yl);
  acc = acc + P4F(3) * M.f2_1[q + 2] * zr(c_str_phdz4r, A.u1, col + 2 * yl);
  const float xyv1 = D4F(1) * A.u1[pos] + D4F(0) * A.u1[pos - yl] + D4F(2) * A.u1[pos + yl] +
                     D4F(3) * A.u1[pos + 2 * yl] - J12i * M.g_c[k] * acc;
  acc = P4F(2) * M.f1_2[q] * zr(c_str_phdz4r, A.v1, col);
  acc = acc + P4F(0) * M.f1_2[q - 2 * fr] * zr(c_str_phdz4r, A.v1, col - 2 * sl);
  acc = acc + P4F(1) * M.f1_2[q - fr] * zr(c_str_phdz4r, A.v1, col - sl);
  acc = acc + P4F(3) * M.f1_2[q + fr] * zr(c_str_phdz4r, A.v1, col + sl);
  const float xyv2 = D4F(2) * A.v1[pos] + D4F(0) * A.v1[pos - 2 * sl] +
                     D4F(1) * A.v1[pos - sl] + D4F(3) * A.v1[pos + sl] - J12i * M.g_c[k] * acc;
  float J13i = M.f_1[q] * M.g3[k];
  J13i = (float)(1.0 * 1.0 / (double)J13i);
  zs = c_str_dz4r.v[kb][6] * A.u1[col + nz - 7];
#pragma unroll
  for (int m = 5; m >= 0; --m) zs = zs + c_str_dz4r.v[kb][m] * A.u1[col + nz - 1 - m];
  const float xzv1 = J13i * zs;
  acc = P4F(2) * M.f1_c[q] * zr(c_str_pdhz4r, A.w1, col);
  acc = acc + P4F(0) * M.f1_c[q - 2 * fr] * zr(c_str_pdhz4r, A.w1, col - 2 * sl);
  acc = acc + P4F(1) * M.f1_c[q - fr] * zr(c_str_pdhz4r, A.w1, col - sl);
  acc = acc + P4F(3) * M.f1_c[q + fr] * zr(c_str_pdhz4r, A.w1, col + sl);
  const float xzv2 = D4F(2) * A.w1[pos] + D4F(0) * A.w1[pos - 2 * sl] +
                     D4F(1) * A.w1[pos - sl] + D4F(3) * A.w1[pos + sl] - J13i * M.g[k] * acc;
  float J23i = M.f_2[q] * M.g3[k];
  J23i = (float)(1.0 * 1.0 / (double)J23i);
  zs = c_str_dz4r.v[kb][6] * A.v1[col + nz - 7];
#pragma unroll
  for (int m = 5; m >= 0; --m) zs = zs + c_str_dz4r.v[kb][m] * A.v1[col + nz - 1 - m];
  const float yzv1 = J23i * zs;
  acc = P4F(1) * M.f2_c[q] * zr(c_str_pdhz4r, A.w1, col);
  acc = acc + P4F(0) * M.f2_c[q - 1] * zr(c_str_pdhz4r, A.w1, col - yl);
  acc = acc + P4F(2) * M.f2_c[q + 1] * zr(c_str_pdhz4r, A.w1, col + yl);
  acc = acc + P4F(3) * M.f2_c[q + 2] * zr(c_str_pdhz4r, A.w1, col + 2 * yl);
  const float yzv2 = D4F(1) * A.w1[pos] + D4F(0) * A.w1[pos - yl] + D4F(2) * A.w1[pos + yl] +
                     D4F(3) * A.w1[pos + 2 * yl] - J23i * M.g[k] * acc;
  topo_str_update<ST>(A, pos, sl, yl, f_dcrj, vs1, vs2, vs3, xyv1, xyv2, xzv1, xzv2, yzv1, yzv2);
}

// ---------------------------------------------------------------------------
// Metrics (setup).  f* over topographic (a, b) in [-6, n+6)^2; f on [-8, n+8).
struct FPtrs {
  float *f, *f_1, *f_2, *f_c, *f1_1, *f1_2, *f1_c, *f2_1, *f2_2, *f2_c;
};

__global__ void geom_custom_kernel(float* f, int nx, int ny, int frow, float normalize) {
  const int b = -AWP_NGSL + (int)(blockIdx.x * 64 + threadIdx.x);
  const int a = -AWP_NGSL + (int)blockIdx.y;
  if (b >= ny + AWP_NGSL || a >= nx + AWP_NGSL) return;
  const int idx = (b + AWP_ALIGN + AWP_NGSL + 2) + (a + AWP_NGSL + 2) * frow;
  f[idx] = (float)(1.0 + (double)(f[idx] * normalize));
}

__global__ void metrics_interp_kernel(FPtrs F, int nx, int ny, int frow) {
  const int b = -6 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int a = -6 + (int)blockIdx.y;
  if (b >= ny + 6 || a >= nx + 6) return;
  const int q = (b + AWP_ALIGN + AWP_NGSL + 2) + (a + AWP_NGSL + 2) * frow;
  const float* f = F.f;
  F.f_1[q] = P4F(0) * f[q - 2] + P4F(1) * f[q - 1] + P4F(2) * f[q] + P4F(3) * f[q + 1];
  F.f_2[q] = P4F(0) * f[q - frow] + P4F(1) * f[q] + P4F(2) * f[q + frow] + P4F(3) * f[q + 2 * frow];
  float acc = 0.f;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int r = q - 2 + t;
    const float row = P4F(0) * f[r - frow] + P4F(1) * f[r] + P4F(2) * f[r + frow] +
                      P4F(3) * f[r + 2 * frow];
    acc = t == 0 ? P4F(0) * row : acc + P4F(t) * row;
  }
  F.f_c[q] = acc;
}

__global__ void metrics_diff_kernel(FPtrs F, int nx, int ny, int frow, float hi) {
  const int b = -6 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int a = -6 + (int)blockIdx.y;
  if (b >= ny + 6 || a >= nx + 6) return;
  const int q = (b + AWP_ALIGN + AWP_NGSL + 2) + (a + AWP_NGSL + 2) * frow;
  const int r = frow;
  const float *f = F.f, *f_1 = F.f_1, *f_2 = F.f_2, *f_c = F.f_c;
  F.f1_1[q] = hi * (D4F(0) * f_c[q - 2 * r] + D4F(1) * f_c[q - r] + D4F(2) * f_c[q] +
                    D4F(3) * f_c[q + r]);
  F.f1_2[q] = hi * (D4F(0) * f[q - r] + D4F(1) * f[q] + D4F(2) * f[q + r] + D4F(3) * f[q + 2 * r]);
  F.f1_c[q] = hi * (D4F(0) * f_1[q - r] + D4F(1) * f_1[q] + D4F(2) * f_1[q + r] +
                    D4F(3) * f_1[q + 2 * r]);
  F.f2_1[q] = hi * (D4F(0) * f[q - 2] + D4F(1) * f[q - 1] + D4F(2) * f[q] + D4F(3) * f[q + 1]);
  F.f2_2[q] = hi * (D4F(0) * f_c[q - 1] + D4F(1) * f_c[q] + D4F(2) * f_c[q + 1] +
                    D4F(3) * f_c[q + 2]);
  F.f2_c[q] = hi * (D4F(0) * f_2[q - 2] + D4F(1) * f_2[q - 1] + D4F(2) * f_2[q] +
                    D4F(3) * f_2[q + 1]);
}

// g without stretching (geom_no_grid_stretching): g(k) = h' k, 0 at the top
// ghost.
__global__ void g_fill_kernel(float* g0, int nz, float hp) {
  const int k = (int)(blockIdx.x * 64 + threadIdx.x);
  if (k >= nz) return;
  g0[AWP_ALIGN + k] = k == nz - 1 ? 0.0f : (float)((double)hp * ((double)k - 0.5 * 0 - 0));
}

// g_c, then g3 and g3_c; one workgroup, two passes.
__global__ __launch_bounds__(256) void metrics_g_kernel(const float* g0, float* gc0, float* g30,
                                                        float* g3c0, int nz, float hi) {
  const float* G = g0 + AWP_ALIGN;
  float* Gc = gc0 + AWP_ALIGN;
  float* G3 = g30 + AWP_ALIGN;
  float* G3c = g3c0 + AWP_ALIGN;
  for (int k = (int)threadIdx.x; k < nz; k += 256) {
    float acc;
    if (k < 6) {
      acc = (float)TM_PHZL[k][0] * G[0];
      for (int m = 1; m < 7; ++m) acc = acc + (float)TM_PHZL[k][m] * G[m];
    } else if (k >= nz - 6) {
      const int r = nz - 1 - k;
      acc = (float)TM_PHZR[r][7] * G[nz - 8];
      for (int m = 6; m >= 0; --m) acc = acc + (float)TM_PHZR[r][m] * G[nz - 1 - m];
    } else {
      acc = P4F(0) * G[k - 2] + P4F(1) * G[k - 1] + P4F(2) * G[k] + P4F(3) * G[k + 1];
    }
    Gc[k] = acc;
  }
  __syncthreads();
  for (int k = (int)threadIdx.x; k < nz; k += 256) {
    float acc, accc;
    if (k < 6) {
      acc = (float)TM_DZL[k][0] * Gc[0];
      for (int m = 1; m < 8; ++m) acc = acc + (float)TM_DZL[k][m] * Gc[m];
      accc = (float)TM_DHZL[k][0] * G[0];
      for (int m = 1; m < 7; ++m) accc = accc + (float)TM_DHZL[k][m] * G[m];
    } else if (k >= nz - 6) {
      const int r = nz - 1 - k;
      acc = (float)TM_DZR[r][6] * Gc[nz - 7];
      for (int m = 5; m >= 0; --m) acc = acc + (float)TM_DZR[r][m] * Gc[nz - 1 - m];
      accc = (float)TM_DHZR[r][7] * G[nz - 8];
      for (int m = 6; m >= 0; --m) accc = accc + (float)TM_DHZR[r][m] * G[nz - 1 - m];
    } else {
      acc = D4F(0) * Gc[k - 1] + D4F(1) * Gc[k] + D4F(2) * Gc[k + 1] + D4F(3) * Gc[k + 2];
      accc = D4F(0) * G[k - 2] + D4F(1) * G[k - 1] + D4F(2) * G[k] + D4F(3) * G[k + 1];
    }
    G3[k] = k == nz - 1 ? 1.0f : hi * acc;
    G3c[k] = hi * accc;
  }
}

// ---------------------------------------------------------------------------
// Host side.
bool topo_geom(const awp_topo_t* T, const char* where, TGeom* G) {
  if (!T) {
    awp_set_error(AWP_ERR_ARG, "%s: topo_t is NULL", where);
    return false;
  }
  if (T->nx < 8 || T->ny < 8 || T->nz < 18) {
    awp_set_error(AWP_ERR_ARG, "%s: grid %d x %d x %d too small (need nx, ny >= 8, nz >= 18)",
                  where, T->nx, T->ny, T->nz);
    return false;
  }
  const long long yl = T->nz + 2 * AWP_ALIGN;
  const long long sl = (long long)(T->ny + 4 + AWP_NGSL2) * yl;
  const long long n = sl * (T->nx + 4 + AWP_NGSL2);
  if (n * 2 >= (1ll << 31)) {
    awp_set_error(AWP_ERR_ARG, "%s: %lld cells per field exceed the 32-bit index range", where,
                  n);
    return false;
  }
  G->nx = T->nx;
  G->ny = T->ny;
  G->nz = T->nz;
  G->yl = (int)yl;
  G->sl = (int)sl;
  G->frow = 2 * AWP_ALIGN + 2 * AWP_NGSL + T->ny + 4;
  return true;
}

TMetrics topo_metrics(const awp_topo_t* T) {
  const awp_f_grid_t& f = T->metrics_f;
  const awp_g_grid_t& g = T->metrics_g;
  TMetrics M;
  M.f = f.d_f;
  M.f_1 = f.d_f_1;
  M.f_2 = f.d_f_2;
  M.f_c = f.d_f_c;
  M.f1_1 = f.d_f1_1;
  M.f1_2 = f.d_f1_2;
  M.f1_c = f.d_f1_c;
  M.f2_1 = f.d_f2_1;
  M.f2_2 = f.d_f2_2;
  M.f2_c = f.d_f2_c;
  M.g = g.d_g;
  M.g3 = g.d_g3;
  M.g3_c = g.d_g3_c;
  M.g_c = g.d_g_c;
  return M;
}

bool metrics_present(const TMetrics& M, const char* where) {
  const float* all[14] = {M.f, M.f_1, M.f_2, M.f_c, M.f1_1, M.f1_2, M.f1_c,
                          M.f2_1, M.f2_2, M.f2_c, M.g, M.g3, M.g3_c, M.g_c};
  for (const float* x : all)
    if (!x) {
      awp_set_error(AWP_ERR_ARG, "%s: a metric array is NULL", where);
      return false;
    }
  return true;
}

template <class ST>
void topo_velocity_launch(awp_topo_t* T, const WaveScales& W, const char* where) {
  TGeom G;
  if (!topo_geom(T, where, &G)) return;
  if (!T->use) return;
  TopoVelArgs<ST> A;
  using Q = typename ST::T;
  A.u1 = reinterpret_cast<Q*>(T->u1);
  A.u2 = reinterpret_cast<Q*>(T->v1);
  A.u3 = reinterpret_cast<Q*>(T->w1);
  A.s11 = TField<ST>{reinterpret_cast<const Q*>(T->xx), W.s[AWP_SXX]};
  A.s12 = TField<ST>{reinterpret_cast<const Q*>(T->xy), W.s[AWP_SXY]};
  A.s13 = TField<ST>{reinterpret_cast<const Q*>(T->xz), W.s[AWP_SXZ]};
  A.s22 = TField<ST>{reinterpret_cast<const Q*>(T->yy), W.s[AWP_SYY]};
  A.s23 = TField<ST>{reinterpret_cast<const Q*>(T->yz), W.s[AWP_SYZ]};
  A.s33 = TField<ST>{reinterpret_cast<const Q*>(T->zz), W.s[AWP_SZZ]};
  A.su1 = W.s[AWP_SU];
  A.su2 = W.s[AWP_SV];
  A.su3 = W.s[AWP_SW];
  A.iu1 = W.inv[AWP_SU];
  A.iu2 = W.inv[AWP_SV];
  A.iu3 = W.inv[AWP_SW];
  A.rho = T->rho;
  A.dcrjx = T->dcrjx;
  A.dcrjy = T->dcrjy;
  A.dcrjz = T->dcrjz;
  A.M = topo_metrics(T);
  A.a = T->timestep;
  A.nu = T->dth;
  if (!metrics_present(A.M, where)) return;
  if (!T->u1 || !T->v1 || !T->w1 || !T->xx || !T->yy || !T->zz || !T->xy || !T->xz ||
      !T->yz || !T->rho || !T->dcrjx || !T->dcrjy || !T->dcrjz) {
    awp_set_error(AWP_ERR_ARG, "%s: a field pointer is NULL", where);
    return;
  }
  // velocity bounds of one rank without neighbours: i in [0, nx), j in [0, ny)
  // (topography.c:309-335); the reference splits j into front / interior /
  // back launches of the same kernels.
  G.bi = 0;
  G.ei = T->nx;
  G.bj = 0;
  G.ej = T->ny;
  hipStream_t st = T->stream_i;
  {
    dim3 block(64, 4, 1);
    dim3 grid(awp_cdiv(T->nz - 6, 64), awp_cdiv(T->ny, 4), T->nx);
    hipLaunchKernelGGL(topo_vel_int_kernel<ST>, grid, block, 0, st, A, G);
    if (!awp_check_launch(where)) return;
  }
  {
    dim3 block(6, 32, 1);
    dim3 grid(1, awp_cdiv(T->ny, 32), T->nx);
    hipLaunchKernelGGL(topo_vel_top_kernel<ST>, grid, block, 0, st, A, G);
    awp_check_launch(where);
  }
}

template <class ST>
void topo_stress_launch(awp_topo_t* T, const WaveScales& W, const char* where) {
  TGeom G;
  if (!topo_geom(T, where, &G)) return;
  if (!T->use) return;
  if (!T->constants_set) {
    awp_set_error(AWP_ERR_STATE, "%s: topo_set_constants has not been called", where);
    return;
  }
  TopoStrArgs<ST> A;
  using Q = typename ST::T;
  A.xx = reinterpret_cast<Q*>(T->xx);
  A.yy = reinterpret_cast<Q*>(T->yy);
  A.zz = reinterpret_cast<Q*>(T->zz);
  A.xy = reinterpret_cast<Q*>(T->xy);
  A.xz = reinterpret_cast<Q*>(T->xz);
  A.yz = reinterpret_cast<Q*>(T->yz);
  A.r1 = reinterpret_cast<Q*>(T->r1);
  A.r2 = reinterpret_cast<Q*>(T->r2);
  A.r3 = reinterpret_cast<Q*>(T->r3);
  A.r4 = reinterpret_cast<Q*>(T->r4);
  A.r5 = reinterpret_cast<Q*>(T->r5);
  A.r6 = reinterpret_cast<Q*>(T->r6);
  A.u1 = TField<ST>{reinterpret_cast<const Q*>(T->u1), W.s[AWP_SU]};
  A.v1 = TField<ST>{reinterpret_cast<const Q*>(T->v1), W.s[AWP_SV]};
  A.w1 = TField<ST>{reinterpret_cast<const Q*>(T->w1), W.s[AWP_SW]};
  A.lam = T->lami;
  A.mu = T->mui;
  A.qp = T->qpi;
  A.coeff = T->coeff;
  A.qs = T->qsi;
  A.dcrjx = T->dcrjx;
  A.dcrjy = T->dcrjy;
  A.dcrjz = T->dcrjz;
  A.vx1 = T->vx1;
  A.vx2 = T->vx2;
  A.ww = T->ww;
  A.wwo = T->wwo;
  A.M = topo_metrics(T);
  A.W = W;
  A.DT = T->c_DT;
  A.dth = T->c_dth;
  A.dh1 = T->c_dh1;
  if (!metrics_present(A.M, where)) return;
  const void* need[] = {T->xx, T->yy, T->zz, T->xy, T->xz, T->yz, T->r1, T->r2, T->r3,
                        T->r4, T->r5, T->r6, T->u1, T->v1, T->w1, T->lami, T->mui,
                        T->qpi, T->coeff, T->qsi, T->dcrjx, T->dcrjy, T->dcrjz, T->vx1,
                        T->vx2, T->ww, T->wwo};
  for (const void* x : need)
    if (!x) {
      awp_set_error(AWP_ERR_ARG, "%s: a field pointer is NULL", where);
      return;
    }
  // stress bounds (topography.c:377-420) shifted by ngsl+2 into array
  // indices (stress_attenuation.cu:34, 76-79)
  const int half = AWP_NGSL / 2, shift = AWP_NGSL + 2;
  const int left1 = half;
  const int right0 = T->nx - half;
  const int y0 = T->y_rank_f < 0 ? half : -half;
  const int y1 = T->y_rank_b < 0 ? T->ny - half : T->ny + half;
  G.bi = left1 + shift;
  G.ei = right0 + shift;
  G.bj = y0 + shift;
  G.ej = y1 + shift;
  if (G.ei <= G.bi || G.ej <= G.bj) return;
  hipStream_t st = T->stream_i;
  {
    dim3 block(64, 4, 1);
    dim3 grid(awp_cdiv(AWP_ALIGN + T->nz - 6, 64), awp_cdiv(G.ej - G.bj, 4), G.ei - G.bi);
    hipLaunchKernelGGL(topo_str_int_kernel<ST>, grid, block, 0, st, A, G);
    if (!awp_check_launch(where)) return;
  }
  {
    dim3 block(6, 32, 1);
    dim3 grid(1, awp_cdiv(G.ej - G.bj, 32), G.ei - G.bi);
    hipLaunchKernelGGL(topo_str_top_kernel<ST>, grid, block, 0, st, A, G);
    awp_check_launch(where);
  }
}

bool neighbourless(int rank, const char* where) {
  if (rank >= 0) {
    awp_set_error(AWP_ERR_ARG,
                  "%s: a neighbouring rank (%d) needs the halo buffers of a multi-GPU run, "
                  "which this one-GPU library does not keep",
                  where, rank);
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

void topo_set_constants(awp_topo_t* T) {
  if (!T) {
    awp_set_error(AWP_ERR_ARG, "topo_set_constants: topo_t is NULL");
    return;
  }
  // set_constants(gridspacing, dth * gridspacing, ...) (stress_attenuation.cu:19-72)
  const float h = T->gridspacing;
  const float dt = T->dth * h;
  T->c_DT = dt;
  T->c_dth = dt / h;
  T->c_dh1 = (float)(1.0 / (double)h);
  T->constants_set = 1;
}

void topo_velocity_interior_H(awp_topo_t* T) {
  topo_velocity_launch<StF32>(T, WaveScales{}, "topo_velocity_interior_H");
}

void topo_velocity_front_H(awp_topo_t* T) {
  if (T && T->use) neighbourless(T->y_rank_f, "topo_velocity_front_H");
}

void topo_velocity_back_H(awp_topo_t* T) {
  if (T && T->use) neighbourless(T->y_rank_b, "topo_velocity_back_H");
}

void topo_stress_interior_H(awp_topo_t* T) {
  topo_stress_launch<StF32>(T, WaveScales{}, "topo_stress_interior_H");
}

void topo_stress_left_H(awp_topo_t* T) {
  if (T && T->use) neighbourless(T->x_rank_l, "topo_stress_left_H");
}

void topo_stress_right_H(awp_topo_t* T) {
  if (T && T->use) neighbourless(T->x_rank_r, "topo_stress_right_H");
}

void awp_topo_velocity_s16(awp_topo_t* T, int fmt, const float* scale) {
  WaveScales W;
  if (!awp_scales_from_host("awp_topo_velocity_s16", fmt, scale, &W)) return;
  if (fmt == AWP_FMT_BF16)
    topo_velocity_launch<StBF16>(T, W, "awp_topo_velocity_s16");
  else
    topo_velocity_launch<StF16>(T, W, "awp_topo_velocity_s16");
}

void awp_topo_stress_s16(awp_topo_t* T, int fmt, const float* scale) {
  WaveScales W;
  if (!awp_scales_from_host("awp_topo_stress_s16", fmt, scale, &W)) return;
  if (fmt == AWP_FMT_BF16)
    topo_stress_launch<StBF16>(T, W, "awp_topo_stress_s16");
  else
    topo_stress_launch<StF16>(T, W, "awp_topo_stress_s16");
}

void awp_topo_geom_custom(float* f, int nx, int ny, int nz, float h, hipStream_t St) {
  if (!f || nx < 1 || ny < 1 || nz < 3 || !(h > 0.0f)) {
    awp_set_error(AWP_ERR_ARG, "awp_topo_geom_custom: bad arguments");
    return;
  }
  const float normalize = (float)(1.0 / (double)h / (double)(nz - 2));
  const int frow = 2 * AWP_ALIGN + 2 * AWP_NGSL + ny + 4;
  dim3 grid(awp_cdiv(ny + 2 * AWP_NGSL, 64), nx + 2 * AWP_NGSL, 1);
  hipLaunchKernelGGL(geom_custom_kernel, grid, dim3(64), 0, St, f, nx, ny, frow, normalize);
  awp_check_launch("awp_topo_geom_custom");
}

void awp_topo_metrics_f(const awp_f_grid_t* mf, int nx, int ny, float h, hipStream_t St) {
  if (!mf || nx < 1 || ny < 1 || !(h > 0.0f)) {
    awp_set_error(AWP_ERR_ARG, "awp_topo_metrics_f: bad arguments");
    return;
  }
  FPtrs F{mf->d_f, mf->d_f_1, mf->d_f_2, mf->d_f_c, mf->d_f1_1,
          mf->d_f1_2, mf->d_f1_c, mf->d_f2_1, mf->d_f2_2, mf->d_f2_c};
  const float* all[10] = {F.f, F.f_1, F.f_2, F.f_c, F.f1_1, F.f1_2, F.f1_c, F.f2_1, F.f2_2,
                          F.f2_c};
  for (const float* x : all)
    if (!x) {
      awp_set_error(AWP_ERR_ARG, "awp_topo_metrics_f: a metric array is NULL");
      return;
    }
  const int frow = 2 * AWP_ALIGN + 2 * AWP_NGSL + ny + 4;
  const float hi = (float)(1.0 / (double)h);  // f_grid_t.hi (metrics.c:24)
  dim3 grid(awp_cdiv(ny + 12, 64), nx + 12, 1);
  hipLaunchKernelGGL(metrics_interp_kernel, grid, dim3(64), 0, St, F, nx, ny, frow);
  if (!awp_check_launch("awp_topo_metrics_f")) return;
  hipLaunchKernelGGL(metrics_diff_kernel, grid, dim3(64), 0, St, F, nx, ny, frow, hi);
  awp_check_launch("awp_topo_metrics_f");
}

void awp_topo_geom_no_stretching(float* g, int nz, float h, hipStream_t St) {
  if (!g || nz < 18 || !(h > 0.0f)) {
    awp_set_error(AWP_ERR_ARG, "awp_topo_geom_no_stretching: bad arguments (nz >= 18)");
    return;
  }
  const float hi = (float)(1.0 / (double)h);   // g_grid_t.hi (metrics.c:211)
  const float hp = (float)(1.0 / (double)hi);  // the z grid spacing (metrics.c:335)
  hipLaunchKernelGGL(g_fill_kernel, dim3(awp_cdiv(nz, 64)), dim3(64), 0, St, g, nz, hp);
  awp_check_launch("awp_topo_geom_no_stretching");
}

void awp_topo_metrics_g(const awp_g_grid_t* mg, int nz, float h, hipStream_t St) {
  if (!mg || !mg->d_g || !mg->d_g_c || !mg->d_g3 || !mg->d_g3_c || nz < 18 || !(h > 0.0f)) {
    awp_set_error(AWP_ERR_ARG, "awp_topo_metrics_g: bad arguments (nz >= 18, four arrays)");
    return;
  }
  const float hi = (float)(1.0 / (double)h);  // g_grid_t.hi (metrics.c:211)
  hipLaunchKernelGGL(metrics_g_kernel, dim3(1), dim3(256), 0, St, mg->d_g, mg->d_g_c, mg->d_g3,
                     mg->d_g3_c, nz, hi);
  awp_check_launch("awp_topo_metrics_g");
}

}  // extern "C"
