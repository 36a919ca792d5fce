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
#include "awp_tuning.h"
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

// the reference's nu * 1.0 / A with binary32 nu and A (x 1.0 is exact)
__device__ __forceinline__ float dbl_quot(float num, float den) { return awp_quot(num, den); }

// Orders one wave's LDS stores before its lanes read each other's values
// (release, then acquire, at wavefront scope).  Enough where a wave reads
// only the LDS it writes itself; a chunk of several waves sharing LDS needs
// the workgroup barrier.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Velocity: dtopo_vel_111_unroll (velocity_unroll.cu:1-461) on levels
// 3 <= k < nz-6 and dtopo_vel_112 (optimized_velocity.cu:1040-1627) on the top
// six, in one kernel.
//
// A wave owns one (j, 64-level) column segment and marches along i.  At every
// plane each lane forms, at its own level, the six horizontal 4-point
// interpolants whose z derivatives the metric terms need, and publishes them
// (with its own s13, s23, s33) in LDS; the 7-point (9-point on the top rows)
// z sums then read the neighbouring levels from LDS instead of recomputing
// 7 x 4 loads per term.  The x neighbours travel in register windows.  Lanes
// 3..60 of a segment produce output, so segments start every 58 levels; the
// last segment of a column ends at level nz-1 and also takes the one-sided top
// rows (table row kk = nz-1-k), as a divergent branch of its wave.
#define VT_BJ 4      // waves (j columns) per workgroup
#define VT_NL 9      // LDS arrays per wave
#define VT_MAXW 8    // waves per workgroup of the chunked stress kernel at most

// The k tiling of a launch.  The velocity kernel takes one-wave segments
// (the first four fields, 58 outputs each); the stress kernel takes chunks
// of nseg waves that share one LDS column (all fields, topo_tiles).
struct VelTiles {
  int nint;   // interior segments / chunks (58 or ch - 6 outputs each, from level 3)
  int ktop0;  // first level of the top one (nz - 64 or nz - ch, may be negative)
  int wtop0;  // first level the top one writes
  int ichunk; // planes per march
  int nseg;   // waves per chunk (64 levels each)
  int ch;     // levels per chunk, 64 * nseg
  int step;   // ch - 6
  int rj;     // j rows per workgroup (nseg * rj waves)
  int wsync;  // 1: a wave that reads only its own LDS syncs at wave scope, not the workgroup's
};

// COMP: 0 updates u1, u2 and u3; 1, 2, 3 update only u1, u2 or u3 (the
// other components' loads, LDS rows and arithmetic drop out), so three
// launches with fewer live registers each give the same results.  The
// one-component launches are asked for 8 waves per SIMD (64 VGPRs: 58-64
// from 71-75, no spills in binary32).
template <class ST, int COMP = 0>
__global__ __launch_bounds__(64 * VT_BJ) __attribute__((amdgpu_waves_per_eu(COMP == 0 ? 1 : 8)))
void topo_vel_kernel(const TopoVelArgs<ST> A,
                                                              const TGeom G, const VelTiles V) {
  constexpr bool DU = COMP == 0 || COMP == 1, DV = COMP == 0 || COMP == 2,
                 DW = COMP == 0 || COMP == 3;
  __shared__ float lds[2][VT_BJ][VT_NL][64];
  const int lane = (int)threadIdx.x;
  const int w = (int)threadIdx.y;
  const int t = (int)blockIdx.x;
  const bool top = t == V.nint;
  const int k0 = top ? V.ktop0 : 58 * t;
  const int k = k0 + lane;
  const int nz = G.nz;
  // levels this lane writes
  const int w0 = top ? V.wtop0 : 58 * t + 3;
  const int w1 = top ? nz : (58 * t + 61 < V.wtop0 ? 58 * t + 61 : V.wtop0);
  bool out = k >= w0 && k < w1;
  const bool bnd = k >= nz - 6;  // one-sided rows (top segment only)
  const int kk = nz - 1 - k;
  // every wave of the workgroup runs the same planes (uniform barrier count);
  // a wave past the last column works on a copy of it and stores nothing
  const int jw = G.bj + (int)(blockIdx.y * VT_BJ) + w;
  const bool jok = jw < G.ej;
  const int j = __builtin_amdgcn_readfirstlane(jok ? jw : G.ej - 1);
  const int ib = (int)blockIdx.z * V.ichunk;
  const int ie = ib + V.ichunk < G.ei ? ib + V.ichunk : G.ei;
  if (ib >= ie) return;  // uniform over the workgroup
  out = out && jok;
  const int sl = G.sl, yl = G.yl, fr = G.frow;
  // every lane reads at its own level; the top segment of a grid with
  // nz < 64 starts below level 0, and those lanes (never outputs, never read
  // by an output's stencil) read level 0 instead
  const int kr = k < 0 ? 0 : k;
  const int kg = kr + AWP_ALIGN;
  const int col0 = yl * (j + AWP_NGSL + 2) + AWP_ALIGN;   // level 0 of column (i = -10)
  const TMetrics& M = A.M;

  constexpr float dhpz4[7] = {(float)TV_DHPZ4[0], (float)TV_DHPZ4[1], (float)TV_DHPZ4[2],
                              (float)TV_DHPZ4[3], (float)TV_DHPZ4[4], (float)TV_DHPZ4[5],
                              (float)TV_DHPZ4[6]};
  constexpr float dphz4[7] = {(float)TV_DPHZ4[0], (float)TV_DPHZ4[1], (float)TV_DPHZ4[2],
                              (float)TV_DPHZ4[3], (float)TV_DPHZ4[4], (float)TV_DPHZ4[5],
                              (float)TV_DPHZ4[6]};
  // per-level weights of the z sums (interior rows); the top rows form theirs
  // from the tables inside their branch
  float cz[7], czp[7];
#pragma unroll
  for (int m = 0; m < 7; ++m) {
    const int kk2 = kg + m - 3;  // >= align - 3
    cz[m] = dhpz4[m] * M.g_c[kk2];
    czp[m] = dphz4[m] * M.g[kk2];
  }
  const float g3c = M.g3_c[kg], g3v = M.g3[kg];
  const float dcz = A.dcrjz[kg];
  const float dcy = A.dcrjy[j + AWP_NGSL + 2];

  // x windows at level k: s11(i-2..i+1), s12(i-1..i+2), s13(i-1..i+2)
  auto P = [&](int i) { return col0 + sl * (i + AWP_NGSL + 2) + kr; };
  float s11w[4], s12x[4], s13x[4];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    s11w[q + 1] = A.s11[P(ib - 2 + q)];
    s12x[q + 1] = A.s12[P(ib - 1 + q)];
    s13x[q + 1] = A.s13[P(ib - 1 + q)];
  }
  // density at plane i-1: (j, k-1), (j, k), (j-1, k)
  float rm_k1 = A.rho[P(ib - 1) - 1], rm_k = A.rho[P(ib - 1)], rm_j = A.rho[P(ib - 1) - yl];

  for (int i = ib; i < ie; ++i) {
    const int par = i & 1;
    const int p = P(i);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      s11w[q] = s11w[q + 1];
      s12x[q] = s12x[q + 1];
      s13x[q] = s13x[q + 1];
    }
    s11w[3] = A.s11[p + sl];
    s12x[3] = A.s12[p + 2 * sl];
    s13x[3] = A.s13[p + 2 * sl];
    const float s12m2 = A.s12[p - 2 * yl], s12m1 = A.s12[p - yl], s12p1 = A.s12[p + yl];
    const float s22m1 = A.s22[p - yl], s22o = A.s22[p], s22p1 = A.s22[p + yl],
                s22p2 = A.s22[p + 2 * yl];
    const float s23m2 = A.s23[p - 2 * yl], s23m1 = A.s23[p - yl], s23o = A.s23[p],
                s23p1 = A.s23[p + yl];
    const float s33o = A.s33[p];
    const float s12o = s12x[1], s13o = s13x[1];
    // interpolants at this level, operands in the reference's order
    const float I11 = P4F(2) * s11w[2] + P4F(0) * s11w[0] + P4F(1) * s11w[1] + P4F(3) * s11w[3];
    const float I12 = P4F(2) * s12o + P4F(0) * s12m2 + P4F(1) * s12m1 + P4F(3) * s12p1;
    const float I12x = P4F(1) * s12x[1] + P4F(0) * s12x[0] + P4F(2) * s12x[2] + P4F(3) * s12x[3];
    const float I22 = P4F(1) * s22o + P4F(0) * s22m1 + P4F(2) * s22p1 + P4F(3) * s22p2;
    const float I13 = P4F(1) * s13x[1] + P4F(0) * s13x[0] + P4F(2) * s13x[2] + P4F(3) * s13x[3];
    const float I23 = P4F(2) * s23o + P4F(0) * s23m2 + P4F(1) * s23m1 + P4F(3) * s23p1;
    float(*L)[64] = lds[par][w];
    if (DU) {
      L[0][lane] = I11;
      L[1][lane] = I12;
      L[6][lane] = s13o;
    }
    if (DV) {
      L[2][lane] = I12x;
      L[3][lane] = I22;
      L[7][lane] = s23o;
    }
    if (DW) {
      L[4][lane] = I13;
      L[5][lane] = I23;
      L[8][lane] = s33o;
    }
    // density at plane i
    const float r_k1 = A.rho[p - 1], r_k = A.rho[p], r_j = A.rho[p - yl],
                r_jk1 = A.rho[p - yl - 1];
    // each wave reads only its own LDS rows: with wsync the waves of a
    // workgroup do not wait for each other
    if (V.wsync)
      wave_lds_sync();
    else
      __syncthreads();

    if (out) {
      const int q = (j + AWP_ALIGN + AWP_NGSL + 2) + (i + AWP_NGSL + 2) * fr;
      float rho1, rho2, rho3;
      if (!bnd) {
        const float c = 0.25f;
        rho1 = c * (r_k1 + r_jk1) + c * (r_k + r_j);
        rho2 = c * (r_k1 + rm_k1) + c * (r_k + rm_k);
        rho3 = c * (r_k + rm_k) + c * (r_j + rm_j);
      } else {
        rho1 = (float)(0.25 * (double)(r_k1 + r_jk1) + 0.25 * (double)(r_k + r_j));
        rho2 = (float)(0.25 * (double)(r_k1 + rm_k1) + 0.25 * (double)(r_k + rm_k));
        rho3 = (float)(0.25 * (double)(r_k + rm_k) + 0.25 * (double)(r_j + rm_j));
      }
      float Ai1 = M.f_1[q] * g3c * rho1;
      Ai1 = dbl_quot(A.nu, Ai1);
      float Ai2 = M.f_2[q] * g3c * rho2;
      Ai2 = dbl_quot(A.nu, Ai2);
      float Ai3 = M.f_c[q] * g3v * rho3;
      Ai3 = dbl_quot(A.nu, Ai3);
      const float f_dcrj = A.dcrjx[i + AWP_NGSL + 2] * dcy * dcz;
      float z11, z12, z12x, z22, z13, z23;
      if (!bnd) {
        // order 3, 0, 1, 2, 4, 5, 6
#pragma unroll
        for (int tt = 0; tt < 7; ++tt) {
          const int m = tt == 0 ? 3 : (tt <= 3 ? tt - 1 : tt);
          const int o = lane + m - 3;
          const float t11 = cz[m] * L[0][o], t12 = cz[m] * L[1][o];
          const float t12x = cz[m] * L[2][o], t22 = cz[m] * L[3][o];
          const float t13 = czp[m] * L[4][o], t23 = czp[m] * L[5][o];
          if (tt == 0) {
            z11 = t11; z12 = t12; z12x = t12x; z22 = t22; z13 = t13; z23 = t23;
          } else {
            z11 = z11 + t11; z12 = z12 + t12; z12x = z12x + t12x; z22 = z22 + t22;
            z13 = z13 + t13; z23 = z23 + t23;
          }
        }
      } else {
        // levels nz-9 .. nz-1 live in lanes 55 .. 63 of the top segment
        const int lb = nz - 1 - k0;  // lane of level nz-1
#pragma unroll
        for (int m = 8; m >= 0; --m) {
          const int o = lb - m;
          const float cr = c_vel_dhpz4r.v[kk][m] * M.g_c[AWP_ALIGN + nz - 1 - m];
          const float crp = c_vel_dphz4r.v[kk][m] * M.g[AWP_ALIGN + nz - 1 - m];
          const float t11 = cr * L[0][o], t12 = cr * L[1][o];
          const float t12x = cr * L[2][o], t22 = cr * L[3][o];
          const float t13 = crp * L[4][o], t23 = crp * L[5][o];
          if (m == 8) {
            z11 = t11; z12 = t12; z12x = t12x; z22 = t22; z13 = t13; z23 = t23;
          } else {
            z11 = z11 + t11; z12 = z12 + t12; z12x = z12x + t12x; z22 = z22 + t22;
            z13 = z13 + t13; z23 = z23 + t23;
          }
        }
      }
      float s;
      // u1
      if (DU) {
      s = D4F(2) * M.f_c[q] * g3c * s11w[2];
      s = s + D4F(0) * M.f_c[q - 2 * fr] * g3c * s11w[0];
      s = s + D4F(1) * M.f_c[q - fr] * g3c * s11w[1];
      s = s + D4F(3) * M.f_c[q + fr] * g3c * s11w[3];
      s = s + D4F(2) * M.f[q] * g3c * s12o;
      s = s + D4F(0) * M.f[q - 2] * g3c * s12m2;
      s = s + D4F(1) * M.f[q - 1] * g3c * s12m1;
      s = s + D4F(3) * M.f[q + 1] * g3c * s12p1;
      if (!bnd) {
        s = s + D4F(2) * L[6][lane];
        s = s + D4F(0) * L[6][lane - 2];
        s = s + D4F(1) * L[6][lane - 1];
        s = s + D4F(3) * L[6][lane + 1];
      } else {
        const int lb = nz - 1 - k0;
#pragma unroll
        for (int m = 7; m >= 0; --m) s = s + c_vel_dhz4r.v[kk][m] * L[6][lb - m];
      }
      s = s - M.f1_1[q] * z11;
      s = s - M.f2_1[q] * z12;
      const float v1 = (A.a * ST::ld(A.u1, p, A.su1) + Ai1 * s) * f_dcrj;
      ST::st(A.u1, p, v1, A.iu1);
      }
      // u2
      if (DV) {
      if (!bnd) {
        s = D4F(2) * s23o;
        s = s + D4F(0) * L[7][lane - 2];
        s = s + D4F(1) * L[7][lane - 1];
        s = s + D4F(3) * L[7][lane + 1];
      } else {
        const int lb = nz - 1 - k0;
        s = c_vel_dhz4r.v[kk][7] * L[7][lb - 7];
#pragma unroll
        for (int m = 6; m >= 0; --m) s = s + c_vel_dhz4r.v[kk][m] * L[7][lb - m];
      }
      s = s + D4F(1) * M.f[q] * g3c * s12x[1];
      s = s + D4F(0) * M.f[q - fr] * g3c * s12x[0];
      s = s + D4F(2) * M.f[q + fr] * g3c * s12x[2];
      s = s + D4F(3) * M.f[q + 2 * fr] * g3c * s12x[3];
      s = s + D4F(1) * M.f_c[q] * g3c * s22o;
      s = s + D4F(0) * M.f_c[q - 1] * g3c * s22m1;
      s = s + D4F(2) * M.f_c[q + 1] * g3c * s22p1;
      s = s + D4F(3) * M.f_c[q + 2] * g3c * s22p2;
      s = s - M.f1_2[q] * z12x;
      s = s - M.f2_2[q] * z22;
      const float v2 = (A.a * ST::ld(A.u2, p, A.su2) + Ai2 * s) * f_dcrj;
      ST::st(A.u2, p, v2, A.iu2);
      }
      // u3
      if (DW) {
      s = D4F(2) * M.f_2[q] * g3v * s23o;
      s = s + D4F(0) * M.f_2[q - 2] * g3v * s23m2;
      s = s + D4F(1) * M.f_2[q - 1] * g3v * s23m1;
      s = s + D4F(3) * M.f_2[q + 1] * g3v * s23p1;
      s = s + D4F(1) * M.f_1[q] * g3v * s13x[1];
      s = s + D4F(0) * M.f_1[q - fr] * g3v * s13x[0];
      s = s + D4F(2) * M.f_1[q + fr] * g3v * s13x[2];
      s = s + D4F(3) * M.f_1[q + 2 * fr] * g3v * s13x[3];
      if (!bnd) {
        s = s + D4F(1) * s33o;
        s = s + D4F(0) * L[8][lane - 1];
        s = s + D4F(2) * L[8][lane + 1];
        s = s + D4F(3) * L[8][lane + 2];
      } else {
        const int lb = nz - 1 - k0;
#pragma unroll
        for (int m = 6; m >= 0; --m) s = s + c_vel_dz4r.v[kk][m] * L[8][lb - m];
      }
      s = s - M.f1_c[q] * z13;
      s = s - M.f2_c[q] * z23;
      const float v3 = (A.a * ST::ld(A.u3, p, A.su3) + Ai3 * s) * f_dcrj;
      ST::st(A.u3, p, v3, A.iu3);
      }
    }
    rm_k1 = r_k1;
    rm_k = r_k;
    rm_j = r_j;
  }
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
// f_vx1, f_vx2, f_wwo, f_ww: the cell's relaxation texture.
template <class ST>
__device__ __forceinline__ void topo_str_update(const TopoStrArgs<ST>& A, int pos, int sl, int yl,
                                                float f_dcrj, float nvs1, float nvs2,
                                                float nvs3, float xyv, float xzv, float yzv,
                                                float f_vx1, float f_vx2, float f_wwo,
                                                int f_ww) {
  const int pos_ip1 = pos + sl, pos_jm1 = pos - yl, pos_km1 = pos - 1;
  const int pos_jk1 = pos - yl - 1, pos_ik1 = pos + sl - 1, pos_ijk = pos + sl - yl,
            pos_ijk1 = pos + sl - yl - 1;
  const float* lam = A.lam;
  const float* mu = A.mu;
  const float* qp = A.qp;
  const float* qs = A.qs;
  float xl = 8.0f / (lam[pos] + lam[pos_ip1] + lam[pos_jm1] + lam[pos_ijk] + lam[pos_km1] +
                     lam[pos_ik1] + lam[pos_jk1] + lam[pos_ijk1]);
  float xm = 16.0f / (mu[pos] + mu[pos_ip1] + mu[pos_jm1] + mu[pos_ijk] + mu[pos_km1] +
                      mu[pos_ik1] + mu[pos_jk1] + mu[pos_ijk1]);
  float xmu1 = 2.0f / (mu[pos] + mu[pos_km1]);
  float xmu2 = awp_quot(2.0f, mu[pos] + mu[pos_jm1]);
  float xmu3 = awp_quot(2.0f, mu[pos] + mu[pos_ip1]);
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
  f_rtmp = h1 * xyv;
  sv = ST::ld(A.xy, pos, W.s[AWP_SXY]) + xmu1 * xyv + vx1 * f_r;
  ST::st(A.r4, pos, f_vx2 * f_r + f_wwo * f_rtmp, W.inv[AWP_SR4]);
  f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
  ST::st(A.xy, pos, (sv + DT * f_rtmp) * f_dcrj, W.inv[AWP_SXY]);

  f_r = ST::ld(A.r5, pos, W.s[AWP_SR5]);
  f_rtmp = h2 * xzv;
  sv = ST::ld(A.xz, pos, W.s[AWP_SXZ]) + xmu2 * xzv + vx1 * f_r;
  ST::st(A.r5, pos, f_vx2 * f_r + f_wwo * f_rtmp, W.inv[AWP_SR5]);
  f_rtmp = f_rtmp * (f_wwo - 1.0f) + f_vx2 * f_r * (1.0f - f_vx1);
  ST::st(A.xz, pos, (sv + DT * f_rtmp) * f_dcrj, W.inv[AWP_SXZ]);

  f_r = ST::ld(A.r6, pos, W.s[AWP_SR6]);
  f_rtmp = h3 * yzv;
  sv = ST::ld(A.yz, pos, W.s[AWP_SYZ]) + xmu3 * yzv + vx1 * f_r;
  ST::st(A.r6, pos, f_vx2 * f_r + f_wwo * f_rtmp, W.inv[AWP_SR6]);
  f_rtmp = f_rtmp * (f_wwo - 1.0f) + f_vx2 * f_r * (1.0f - f_vx1);
  ST::st(A.yz, pos, (sv + DT * f_rtmp) * f_dcrj, W.inv[AWP_SYZ]);
}

// Stress: dtopo_str_111 (stress_attenuation.cu:74-668) on levels
// 3 <= k < nz-6 and dtopo_str_112 (:670-1410) on the top six, in one kernel
// with the velocity kernel's lane map (a chunk of nseg waves = (j, 64*nseg
// levels), march along i, ch-6 output levels per chunk, the top rows in the
// last chunk).
//
// The metric terms need 7-point (top rows: 9-point table) z sums of
// velocity columns at x neighbours (i-2 .. i+2) and y neighbours (j-2 .. j+2).
// The x-neighbour sums ride in register windows: one new column's sum per
// field and plane.  The y-neighbour sums and the own column's z differences
// read velocity columns that each plane stages in LDS (15 columns per chunk
// and row).
#define ST_NC 15
#define TOPO_NRATES AWP_TOPO_NRATES
enum {
  C_UN, C_UO, C_UM1, C_UP1, C_UP2,   // u1 at (i+2, j), (i, j), (i, j-1), (i, j+1), (i, j+2)
  C_VN, C_VO, C_VM2, C_VM1, C_VP1,   // v1 at (i+1, j), (i, j), (i, j-2), (i, j-1), (i, j+1)
  C_WN, C_WO, C_WM1, C_WP1, C_WP2    // w1 at (i+1, j), (i, j), (i, j-1), (i, j+1), (i, j+2)
};
// Rate groups of the two-pass update: 0 all six rates (and the one-pass
// update), 1 the normal rates vs1, vs2, vs3, 2 the shear sums xy, xz, yz,
// 3 the xy sum, 4 the xz and yz sums.  A group stages only the columns its
// z sums read, in LDS slots 0.. .
constexpr unsigned ST_COLS_XY =
    (1u << C_UO) | (1u << C_UM1) | (1u << C_UP1) | (1u << C_UP2) | (1u << C_VN);
constexpr unsigned ST_COLS_XZYZ = (1u << C_UO) | (1u << C_VO) | (1u << C_WN) | (1u << C_WM1) |
                                  (1u << C_WP1) | (1u << C_WP2);
constexpr unsigned st_group_cols(int group) {
  return group == 1   ? (1u << C_UN) | (1u << C_VO) | (1u << C_VM2) | (1u << C_VM1) |
                            (1u << C_VP1) | (1u << C_WO)
         : group == 2 ? ST_COLS_XY | ST_COLS_XZYZ
         : group == 3 ? ST_COLS_XY
         : group == 4 ? ST_COLS_XZYZ
                      : (1u << ST_NC) - 1u;
}
constexpr int st_group_ncols(int group) { return __builtin_popcount(st_group_cols(group)); }
// LDS slot of column c in a group (the number of staged columns before it)
constexpr int st_slot(int group, int c) {
  return __builtin_popcount(st_group_cols(group) & ((1u << c) - 1u));
}
constexpr bool st_staged(int group, int c) { return (st_group_cols(group) >> c) & 1u; }

// RATES: store the six strain-rate terms of every output cell into R
// (TOPO_NRATES arrays of the field's size, rstride floats apart) for
// topo_str_q_kernel, instead of applying the material and memory-variable
// update here.
// GROUP (RATES only): the rate group this launch computes and stores
// (st_group_cols); launches of groups 1 and 2, or 1, 3 and 4, store what one
// of group 0 does, each with fewer live registers and LDS slots.
template <class ST, bool RATES, int GROUP = 0>
__global__ __launch_bounds__(64 * VT_MAXW) void topo_str_kernel(const TopoStrArgs<ST> A,
                                                              const TGeom G, const VelTiles V,
                                                              float* __restrict__ R,
                                                              long long rstride) {
  static_assert(GROUP == 0 || RATES, "rate groups apply to the two-pass update");
  // normal rates, the xy sum, the xz and yz sums
  constexpr bool GN = GROUP <= 1, GXY = GROUP != 1 && GROUP != 4,
                 GZ = GROUP != 1 && GROUP != 3;
  constexpr int NC = st_group_ncols(GROUP);
#define SC_(c) st_slot(GROUP, c)
  // [parity][row][slot][level of the chunk]: 2 * rj * NC * ch floats; the
  // column of array c at parity base L
#define LCOL(c) (L + SC_(c) * V.ch)
  extern __shared__ float st_lds[];
  const int lane = (int)threadIdx.x;
  const int w = (int)threadIdx.y;
  const int seg = w % V.nseg, row = w / V.nseg;
  const int t = (int)blockIdx.x;
  const bool top = t == V.nint;
  const int k0 = top ? V.ktop0 : V.step * t;
  const int kl = seg * 64 + lane;  // level within the chunk: the LDS index
  const int k = k0 + kl;  // topographic level
  const int nz = G.nz;
  const int w0 = top ? V.wtop0 : V.step * t + 3;
  const int w1 = top ? nz : (V.step * t + V.ch - 3 < V.wtop0 ? V.step * t + V.ch - 3 : V.wtop0);
  const bool bnd = k >= nz - 6;
  const int kb = nz - 1 - k;  // table row of a top level
  const int jw = G.bj + (int)blockIdx.y * V.rj + row;
  const bool jok = jw < G.ej;
  const int j = __builtin_amdgcn_readfirstlane(jok ? jw : G.ej - 1);
  const int ib = G.bi + (int)blockIdx.z * V.ichunk;
  const int ie = ib + V.ichunk < G.ei ? ib + V.ichunk : G.ei;
  if (ib >= ie) return;  // uniform over the workgroup
  const bool out = k >= w0 && k < w1 && jok;
  const int sl = G.sl, yl = G.yl, fr = G.frow;
  const int kr = k < 0 ? 0 : k;          // read level (see topo_vel_kernel)
  const int K = kr + AWP_ALIGN;          // array level
  const TMetrics& M = A.M;
  const int lb = nz - 1 - k0;            // LDS level of nz-1 (top chunk)

  constexpr float phdz4[7] = {(float)TS_PHDZ4[0], (float)TS_PHDZ4[1], (float)TS_PHDZ4[2],
                              (float)TS_PHDZ4[3], (float)TS_PHDZ4[4], (float)TS_PHDZ4[5],
                              (float)TS_PHDZ4[6]};
  constexpr float pdhz4[7] = {(float)TS_PDHZ4[0], (float)TS_PDHZ4[1], (float)TS_PDHZ4[2],
                              (float)TS_PDHZ4[3], (float)TS_PDHZ4[4], (float)TS_PDHZ4[5],
                              (float)TS_PDHZ4[6]};
  // z sums of an LDS column at this lane's level
  auto zasc = [&](const float* c) {  // phdz4, ascending (normal stresses)
    float acc = phdz4[0] * c[kl - 3];
#pragma unroll
    for (int m = 1; m < 7; ++m) acc = acc + phdz4[m] * c[kl + m - 3];
    return acc;
  };
  auto zmid = [&](const float(&cf)[7], const float* c) {  // order 3, 0, 1, 2, 4, 5, 6
    float acc = cf[3] * c[kl];
    acc = acc + cf[0] * c[kl - 3];
    acc = acc + cf[1] * c[kl - 2];
    acc = acc + cf[2] * c[kl - 1];
    acc = acc + cf[4] * c[kl + 1];
    acc = acc + cf[5] * c[kl + 2];
    acc = acc + cf[6] * c[kl + 3];
    return acc;
  };
  auto zrow = [&](const FTab<6, 9>& tab, const float* c) {  // top rows, m = 8 .. 0
    float acc = tab.v[kb][8] * c[lb - 8];
#pragma unroll
    for (int m = 7; m >= 0; --m) acc = acc + tab.v[kb][m] * c[lb - m];
    return acc;
  };
  // the three x-group sums of this lane for a staged column (non-output
  // lanes compute nothing: their neighbours may lie outside the segment)
  auto zu_x = [&](const float* c) { return !out ? 0.f : (bnd ? zrow(c_str_phdz4r, c) : zasc(c)); };
  auto zv_x = [&](const float* c) { return !out ? 0.f : (bnd ? zrow(c_str_phdz4r, c) : zmid(phdz4, c)); };
  auto zw = [&](const float* c) { return !out ? 0.f : (bnd ? zrow(c_str_pdhz4r, c) : zmid(pdhz4, c)); };

  const int colj = yl * j + K;  // level K of column (i = 0, j)
  // raw x windows at this level: u (i-1 .. i+2), v and w (i-2 .. i+1), and
  // the x-group z sums Zu (i-1 .. i+2), Zv, Zw (i-2 .. i+1)
  float ux[4], vx[4], wx[4], Zu[4], Zv[4], Zw[4];
  // prologue: stage the columns the first plane's windows need
  for (int q = 0; q < 3; ++q) {
    const int iu = ib - 1 + q, iv = ib - 2 + q;
    float* L = st_lds + ((q & 1) * V.rj + row) * NC * V.ch;
    ux[q + 1] = A.u1[colj + sl * iu];
    vx[q + 1] = A.v1[colj + sl * iv];
    wx[q + 1] = A.w1[colj + sl * iv];
    if constexpr (GN) LCOL(C_UN)[kl] = ux[q + 1];
    if constexpr (GXY) LCOL(C_VN)[kl] = vx[q + 1];
    if constexpr (GZ) LCOL(C_WN)[kl] = wx[q + 1];
    // one-wave chunks read only their own LDS rows (see topo_vel_kernel)
    if (V.wsync && V.nseg == 1)
      wave_lds_sync();
    else
      __syncthreads();
    if constexpr (GN) Zu[q + 1] = zu_x(LCOL(C_UN));
    if constexpr (GXY) Zv[q + 1] = zv_x(LCOL(C_VN));
    if constexpr (GZ) Zw[q + 1] = zw(LCOL(C_WN));
  }

  const float g_ck = M.g_c[K], g_k = M.g[K], g3c_k = M.g3_c[K], g3_k = M.g3[K];
  const float g3c_k6 = M.g3_c[bnd ? K : K + 6];  // dtopo_str_111 reads six levels up
  const float dcz = A.dcrjz[K];
  const float dcy = A.dcrjy[j];

  for (int i = ib; i < ie; ++i) {
    const int par = (i + 1) & 1;  // the prologue used parities 0, 1, 0
    float* L = st_lds + (par * V.rj + row) * NC * V.ch;
    const int pos = colj + sl * i;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      ux[q] = ux[q + 1];
      vx[q] = vx[q + 1];
      wx[q] = wx[q + 1];
      Zu[q] = Zu[q + 1];
      Zv[q] = Zv[q + 1];
      Zw[q] = Zw[q + 1];
    }
    ux[3] = A.u1[pos + 2 * sl];
    vx[3] = A.v1[pos + sl];
    wx[3] = A.w1[pos + sl];
    float um1 = 0.f, up1 = 0.f, up2 = 0.f, vm2 = 0.f, vm1 = 0.f, vp1 = 0.f, wm1 = 0.f,
          wp1 = 0.f, wp2 = 0.f;
    if constexpr (GXY) {
      um1 = A.u1[pos - yl];
      up1 = A.u1[pos + yl];
      up2 = A.u1[pos + 2 * yl];
    }
    if constexpr (GZ) {
      wm1 = A.w1[pos - yl];
      wp1 = A.w1[pos + yl];
      wp2 = A.w1[pos + 2 * yl];
    }
    if constexpr (GN) {
      vm2 = A.v1[pos - 2 * yl];
      vm1 = A.v1[pos - yl];
      vp1 = A.v1[pos + yl];
    }
    const float uo = ux[1], vo = vx[2], wo = wx[2];
    auto stage = [&](int c, float x) {
      // c is a literal column at every call, so the test folds away
      if (st_staged(GROUP, c)) L[st_slot(GROUP, c) * V.ch + kl] = x;
    };
    stage(C_UN, ux[3]);
    stage(C_UO, uo);
    stage(C_UM1, um1);
    stage(C_UP1, up1);
    stage(C_UP2, up2);
    stage(C_VN, vx[3]);
    stage(C_VO, vo);
    stage(C_VM2, vm2);
    stage(C_VM1, vm1);
    stage(C_VP1, vp1);
    stage(C_WN, wx[3]);
    stage(C_WO, wo);
    stage(C_WM1, wm1);
    stage(C_WP1, wp1);
    stage(C_WP2, wp2);
    if (V.wsync && V.nseg == 1)
      wave_lds_sync();
    else
      __syncthreads();
    if constexpr (GN) Zu[3] = zu_x(LCOL(C_UN));
    if constexpr (GXY) Zv[3] = zv_x(LCOL(C_VN));
    if constexpr (GZ) Zw[3] = zw(LCOL(C_WN));
    if (out) {
      const int q = j + AWP_ALIGN + i * fr;
      const float f_dcrj = A.dcrjx[i] * dcy * dcz;
      float vs1 = 0.f, vs2 = 0.f, vs3 = 0.f, xyv = 0.f, xzv = 0.f, yzv = 0.f;
      if constexpr (GN) {
        float acc;
        // xx, yy, zz
        float Jii = M.f_c[q] * g3c_k;
        Jii = awp_quot(1.0f, Jii);
        if (!bnd) {
          acc = P4F(0) * M.f1_1[q - fr] * Zu[0];
          acc = acc + P4F(1) * M.f1_1[q] * Zu[1];
        } else {
          acc = P4F(1) * M.f1_1[q] * Zu[1];
          acc = acc + P4F(0) * M.f1_1[q - fr] * Zu[0];
        }
        acc = acc + P4F(2) * M.f1_1[q + fr] * Zu[2];
        acc = acc + P4F(3) * M.f1_1[q + 2 * fr] * Zu[3];
        vs1 = D4F(1) * ux[1] + D4F(0) * ux[0] + D4F(2) * ux[2] + D4F(3) * ux[3] -
                          Jii * g_ck * acc;
        const float zv_o = bnd ? zrow(c_str_phdz4r, LCOL(C_VO)) : zasc(LCOL(C_VO));
        const float zv_m2 = bnd ? zrow(c_str_phdz4r, LCOL(C_VM2)) : zasc(LCOL(C_VM2));
        const float zv_m1 = bnd ? zrow(c_str_phdz4r, LCOL(C_VM1)) : zasc(LCOL(C_VM1));
        const float zv_p1 = bnd ? zrow(c_str_phdz4r, LCOL(C_VP1)) : zasc(LCOL(C_VP1));
        acc = P4F(2) * M.f2_2[q] * zv_o;
        acc = acc + P4F(0) * M.f2_2[q - 2] * zv_m2;
        acc = acc + P4F(1) * M.f2_2[q - 1] * zv_m1;
        acc = acc + P4F(3) * M.f2_2[q + 1] * zv_p1;
        vs2 = D4F(2) * vo + D4F(0) * vm2 + D4F(1) * vm1 + D4F(3) * vp1 -
                          Jii * g_ck * acc;
        if (!bnd) {
          const float* c = LCOL(C_WO);
          vs3 = Jii * (D4F(2) * c[kl] + D4F(0) * c[kl - 2] + D4F(1) * c[kl - 1] +
                       D4F(3) * c[kl + 1]);
        } else {
          const float* c = LCOL(C_WO);
          float zs = c_str_dhz4r.v[kb][7] * c[lb - 7];
  #pragma unroll
          for (int m = 6; m >= 0; --m) zs = zs + c_str_dhz4r.v[kb][m] * c[lb - m];
          vs3 = Jii * zs;
        }
      }
      if constexpr (GXY) {
        float acc;
        // xy
        float J12i = M.f[q] * g3c_k6;
        J12i = awp_quot(1.0f, J12i);
        const float zu_o = bnd ? zrow(c_str_phdz4r, LCOL(C_UO)) : zmid(phdz4, LCOL(C_UO));
        const float zu_m1 = bnd ? zrow(c_str_phdz4r, LCOL(C_UM1)) : zmid(phdz4, LCOL(C_UM1));
        const float zu_p1 = bnd ? zrow(c_str_phdz4r, LCOL(C_UP1)) : zmid(phdz4, LCOL(C_UP1));
        const float zu_p2 = bnd ? zrow(c_str_phdz4r, LCOL(C_UP2)) : zmid(phdz4, LCOL(C_UP2));
        acc = P4F(1) * M.f2_1[q] * zu_o;
        acc = acc + P4F(0) * M.f2_1[q - 1] * zu_m1;
        acc = acc + P4F(2) * M.f2_1[q + 1] * zu_p1;
        acc = acc + P4F(3) * M.f2_1[q + 2] * zu_p2;
        const float xyv1 = D4F(1) * uo + D4F(0) * um1 + D4F(2) * up1 + D4F(3) * up2 -
                           J12i * g_ck * acc;
        acc = P4F(2) * M.f1_2[q] * Zv[2];
        acc = acc + P4F(0) * M.f1_2[q - 2 * fr] * Zv[0];
        acc = acc + P4F(1) * M.f1_2[q - fr] * Zv[1];
        acc = acc + P4F(3) * M.f1_2[q + fr] * Zv[3];
        const float xyv2 = D4F(2) * vx[2] + D4F(0) * vx[0] + D4F(1) * vx[1] + D4F(3) * vx[3] -
                           J12i * g_ck * acc;
        // the update uses the shear terms only as the sums xyv1 + xyv2,
        // xzv1 + xzv2 and yzv1 + yzv2 (stress_attenuation.cu:306-327)
        xyv = xyv1 + xyv2;
      }
      if constexpr (GZ) {
        float acc;
        // xz
        float J13i = M.f_1[q] * g3_k;
        J13i = awp_quot(1.0f, J13i);
        float xzv1;
        if (!bnd) {
          const float* c = LCOL(C_UO);
          xzv1 = J13i * (D4F(1) * c[kl] + D4F(0) * c[kl - 1] + D4F(2) * c[kl + 1] +
                         D4F(3) * c[kl + 2]);
        } else {
          const float* c = LCOL(C_UO);
          float zs = c_str_dz4r.v[kb][6] * c[lb - 6];
  #pragma unroll
          for (int m = 5; m >= 0; --m) zs = zs + c_str_dz4r.v[kb][m] * c[lb - m];
          xzv1 = J13i * zs;
        }
        acc = P4F(2) * M.f1_c[q] * Zw[2];
        acc = acc + P4F(0) * M.f1_c[q - 2 * fr] * Zw[0];
        acc = acc + P4F(1) * M.f1_c[q - fr] * Zw[1];
        acc = acc + P4F(3) * M.f1_c[q + fr] * Zw[3];
        const float xzv2 = D4F(2) * wx[2] + D4F(0) * wx[0] + D4F(1) * wx[1] + D4F(3) * wx[3] -
                           J13i * g_k * acc;
        // yz
        float J23i = M.f_2[q] * g3_k;
        J23i = awp_quot(1.0f, J23i);
        float yzv1;
        if (!bnd) {
          const float* c = LCOL(C_VO);
          yzv1 = J23i * (D4F(1) * c[kl] + D4F(0) * c[kl - 1] + D4F(2) * c[kl + 1] +
                         D4F(3) * c[kl + 2]);
        } else {
          const float* c = LCOL(C_VO);
          float zs = c_str_dz4r.v[kb][6] * c[lb - 6];
  #pragma unroll
          for (int m = 5; m >= 0; --m) zs = zs + c_str_dz4r.v[kb][m] * c[lb - m];
          yzv1 = J23i * zs;
        }
        const float zw_m1 = zw(LCOL(C_WM1)), zw_p1 = zw(LCOL(C_WP1)), zw_p2 = zw(LCOL(C_WP2));
        acc = P4F(1) * M.f2_c[q] * Zw[2];
        acc = acc + P4F(0) * M.f2_c[q - 1] * zw_m1;
        acc = acc + P4F(2) * M.f2_c[q + 1] * zw_p1;
        acc = acc + P4F(3) * M.f2_c[q + 2] * zw_p2;
        const float yzv2 = D4F(1) * wo + D4F(0) * wm1 + D4F(2) * wp1 + D4F(3) * wp2 -
                           J23i * g_k * acc;
        xzv = xzv1 + xzv2;
        yzv = yzv1 + yzv2;
      }
      if (RATES) {
        if constexpr (GN) {
          R[pos] = vs1;
          R[rstride + pos] = vs2;
          R[2 * rstride + pos] = vs3;
        }
        if constexpr (GXY) R[3 * rstride + pos] = xyv;
        if constexpr (GZ) {
          R[4 * rstride + pos] = xzv;
          R[5 * rstride + pos] = yzv;
        }
      } else {
        topo_str_update<ST>(A, pos, sl, yl, f_dcrj, vs1, vs2, vs3, xyv, xzv, yzv, A.vx1[pos],
                            A.vx2[pos], A.wwo[pos], A.ww[pos]);
      }
    }
  }
#undef SC_
#undef LCOL
}

// The relaxation texture of the Q pass: the per-cell arrays (TOPO_TEX_ARRAYS),
// or vx1, vx2, wwo from the 8-entry table indexed by ww, with ww read
// (TOPO_TEX_TABLE) or evaluated from init_texture's parity rule
// (TOPO_TEX_RULE, mesh.c:1103-1137) over the stress region that starts at
// (xls, yls, align) and has nj columns per x plane and nzt levels.
enum { TOPO_TEX_ARRAYS = 0, TOPO_TEX_TABLE = 1, TOPO_TEX_RULE = 2 };
struct TopoTex {
  float t[24];  // vx1[8], vx2[8], wwo[8]
  int xls, yls, nj, nzt;
};

// The material averages, Q weights and memory-variable recursion of the
// stress update (topo_str_update) over the cells topo_str_kernel<ST, true>
// wrote rates for: levels 3 <= k < nz of the stress bounds.  One lane per
// level, 4 rows per workgroup, a march over a chunk of i; few registers, so
// the per-cell loads of this streaming pass stay in flight.
template <class ST, int TEX>
__global__ __launch_bounds__(256) void topo_str_q_kernel(const TopoStrArgs<ST> A, const TGeom G,
                                                         const float* __restrict__ R,
                                                         long long rstride, int ichunk,
                                                         const TopoTex TT) {
  __shared__ float s_tex[24];
  if (TEX != TOPO_TEX_ARRAYS) {
    const int tid = (int)(threadIdx.y * 64 + threadIdx.x);
    if (tid < 24) s_tex[tid] = TT.t[tid];
    __syncthreads();
  }
  const int k = 3 + (int)(blockIdx.x * 64 + threadIdx.x);  // topographic level
  const int j = G.bj + (int)(blockIdx.y * 4 + threadIdx.y);
  const int ib = G.bi + (int)blockIdx.z * ichunk;
  const int ie = ib + ichunk < G.ei ? ib + ichunk : G.ei;
  if (k >= G.nz || j >= G.ej) return;
  const int sl = G.sl, yl = G.yl;
  const int K = k + AWP_ALIGN;
  const float dcz = A.dcrjz[K];
  const float dcy = A.dcrjy[j];
  for (int i = ib; i < ie; ++i) {
    const int pos = yl * j + K + sl * i;
    const float f_dcrj = A.dcrjx[i] * dcy * dcz;
    float f_vx1, f_vx2, f_wwo;
    int f_ww;
    if (TEX == TOPO_TEX_ARRAYS) {
      f_vx1 = A.vx1[pos]; f_vx2 = A.vx2[pos]; f_wwo = A.wwo[pos];
      f_ww = A.ww[pos];
    } else {
      if (TEX == TOPO_TEX_TABLE) {
        f_ww = A.ww[pos];
      } else {
        const int ni = i - TT.xls, col = ni * TT.nj + (j - TT.yls);
        f_ww = 1 + 4 * ((ni + 1) & 1) + 2 * ((col + 1) & 1) + ((TT.nzt + col * TT.nzt + k) & 1);
      }
      const int t = (f_ww - 1) & 7;
      f_vx1 = s_tex[t]; f_vx2 = s_tex[8 + t]; f_wwo = s_tex[16 + t];
    }
    topo_str_update<ST>(A, pos, sl, yl, f_dcrj, R[pos], R[rstride + pos], R[2 * rstride + pos],
                        R[3 * rstride + pos], R[4 * rstride + pos], R[5 * rstride + pos],
                        f_vx1, f_vx2, f_wwo, f_ww);
  }
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
// The stress kernels' constants (set_constants, stress_attenuation.cu:39-72),
// process-wide like the reference's __constant__ block.
struct TopoConstants {
  float DT, dth, dh1;
  int set;
};
TopoConstants g_topo_constants = {0.0f, 0.0f, 0.0f, 0};

bool topo_geom(const topo_t* T, const char* where, TGeom* G) {
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

TMetrics topo_metrics(const topo_t* T) {
  const f_grid_t& f = T->metrics_f;
  const g_grid_t& g = T->metrics_g;
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

// Chunks of the curvilinear stress kernel's k lane map for a column of nz
// levels: nseg = min(topo_nseg, ceil(nz / 64)) waves of 64 levels per chunk,
// so a column of up to 64 * topo_nseg levels is one chunk (nz = 192: three
// waves, where one-wave segments of 58 outputs need four, one of them with
// 12).  The velocity kernel
// measured slower with chunks (1.06 against 0.99 ms at 384x384x192; its top
// wave's one-sided rows hold the chunk's other waves at the plane barrier)
// and keeps one-wave segments.
void topo_tiles(int nz, VelTiles* V) {
  int nseg = awp_cdiv(nz, 64);
  const int cap = awp_tuning().topo_nseg;
  if (nseg > cap) nseg = cap;
  if (nseg > VT_MAXW) nseg = VT_MAXW;
  if (nseg < 1) nseg = 1;
  V->nseg = nseg;
  V->ch = 64 * nseg;
  V->step = V->ch - 6;
  // one j row per workgroup for chunks, so each plane's barrier waits only
  // for the chunk's own waves (1.87 -> 1.78 ms per stress update at
  // 384x384x192 against two rows); four rows of one-wave segments
  V->rj = nseg == 1 ? 4 : 1;
  const int rjf = awp_tuning().topo_rj;  // > 0 forces the rows per workgroup
  if (rjf > 0) V->rj = rjf * nseg <= VT_MAXW ? rjf : VT_MAXW / nseg;
  if (V->rj < 1) V->rj = 1;
  V->ktop0 = nz - V->ch;
  V->wtop0 = V->ktop0 + 3 > 3 ? V->ktop0 + 3 : 3;
  V->nint = V->wtop0 > 3 ? awp_cdiv(V->wtop0 - 3, V->step) : 0;
}

template <class ST>
void topo_velocity_launch(topo_t* T, const WaveScales& W, const char* where) {
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
  // segments: interior ones every 58 levels from level 3, then the top one
  // (levels nz-64 .. nz-1) writing from nz-61 (or 3) up
  VelTiles V;
  V.ktop0 = T->nz - 64;
  V.wtop0 = V.ktop0 + 3 > 3 ? V.ktop0 + 3 : 3;
  V.nint = V.wtop0 > 3 ? awp_cdiv(V.wtop0 - 3, 58) : 0;
  V.ichunk = awp_tuning().topo_vel_ichunk;
  V.wsync = awp_tuning().topo_wsync;
  dim3 block(64, VT_BJ, 1);
  dim3 grid(V.nint + 1, awp_cdiv(T->ny, VT_BJ), awp_cdiv(T->nx, V.ichunk));
  if (awp_tuning().topo_vel_split) {
    hipLaunchKernelGGL((topo_vel_kernel<ST, 1>), grid, block, 0, st, A, G, V);
    hipLaunchKernelGGL((topo_vel_kernel<ST, 2>), grid, block, 0, st, A, G, V);
    hipLaunchKernelGGL((topo_vel_kernel<ST, 3>), grid, block, 0, st, A, G, V);
  } else {
    hipLaunchKernelGGL((topo_vel_kernel<ST, 0>), grid, block, 0, st, A, G, V);
  }
  awp_check_launch(where);
}

// tex, texmode: the Q pass's relaxation texture (TOPO_TEX_*; tex holds
// vx1[8], vx2[8], wwo[8] unless texmode is TOPO_TEX_ARRAYS, which the
// one-pass update always uses).
template <class ST>
void topo_stress_launch(topo_t* T, const WaveScales& W, const char* where, float* rates,
                        const float* tex = nullptr, int texmode = TOPO_TEX_ARRAYS) {
  TGeom G;
  if (!topo_geom(T, where, &G)) return;
  if (!T->use) return;
  if (!g_topo_constants.set) {
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
  A.DT = g_topo_constants.DT;
  A.dth = g_topo_constants.dth;
  A.dh1 = g_topo_constants.dh1;
  if (!metrics_present(A.M, where)) return;
  const void* need[] = {T->xx, T->yy, T->zz, T->xy, T->xz, T->yz, T->r1, T->r2, T->r3,
                        T->r4, T->r5, T->r6, T->u1, T->v1, T->w1, T->lami, T->mui,
                        T->qpi, T->coeff, T->qsi, T->dcrjx, T->dcrjy, T->dcrjz};
  for (const void* x : need)
    if (!x) {
      awp_set_error(AWP_ERR_ARG, "%s: a field pointer is NULL", where);
      return;
    }
  if (texmode != TOPO_TEX_ARRAYS && (!rates || !tex)) {
    awp_set_error(AWP_ERR_ARG, "%s: the texture table applies to the two-pass update", where);
    return;
  }
  if ((texmode == TOPO_TEX_ARRAYS && (!T->vx1 || !T->vx2 || !T->ww || !T->wwo)) ||
      (texmode == TOPO_TEX_TABLE && !T->ww)) {
    awp_set_error(AWP_ERR_ARG, "%s: a relaxation-texture pointer is NULL", where);
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
  VelTiles V;
  topo_tiles(T->nz, &V);
  V.ichunk = awp_tuning().topo_str_ichunk;
  V.wsync = awp_tuning().topo_wsync;
  dim3 block(64, V.nseg * V.rj, 1);
  dim3 grid(V.nint + 1, awp_cdiv(G.ej - G.bj, V.rj), awp_cdiv(G.ei - G.bi, V.ichunk));
  // LDS of a launch of rate group g: 2 parities x rj rows x its columns
  auto lds = [&](int g) {
    return (size_t)2 * V.rj * st_group_ncols(g) * V.ch * sizeof(float);
  };
  if (!rates) {
    hipLaunchKernelGGL((topo_str_kernel<ST, false>), grid, block, lds(0), st, A, G, V, nullptr,
                       0ll);
    awp_check_launch(where);
    return;
  }
  // two passes: strain rates into rates, then the material and memory-
  // variable update streaming them back
  const long long rstride = (long long)G.sl * (T->nx + 4 + AWP_NGSL2);
  if (awp_tuning().topo_str_split == 2) {
    hipLaunchKernelGGL((topo_str_kernel<ST, true, 1>), grid, block, lds(1), st, A, G, V, rates,
                       rstride);
    hipLaunchKernelGGL((topo_str_kernel<ST, true, 3>), grid, block, lds(3), st, A, G, V, rates,
                       rstride);
    hipLaunchKernelGGL((topo_str_kernel<ST, true, 4>), grid, block, lds(4), st, A, G, V, rates,
                       rstride);
  } else if (awp_tuning().topo_str_split) {
    hipLaunchKernelGGL((topo_str_kernel<ST, true, 1>), grid, block, lds(1), st, A, G, V, rates,
                       rstride);
    hipLaunchKernelGGL((topo_str_kernel<ST, true, 2>), grid, block, lds(2), st, A, G, V, rates,
                       rstride);
  } else {
    hipLaunchKernelGGL((topo_str_kernel<ST, true, 0>), grid, block, lds(0), st, A, G, V, rates,
                       rstride);
  }
  const int qchunk = awp_tuning().topo_q_ichunk;
  dim3 gq(awp_cdiv(T->nz - 3, 64), awp_cdiv(G.ej - G.bj, 4), awp_cdiv(G.ei - G.bi, qchunk));
  TopoTex TT = {};
  if (texmode != TOPO_TEX_ARRAYS) {
    for (int m = 0; m < 24; ++m) TT.t[m] = tex[m];
    // the stress region init_texture walks (mesh.c:1103-1137)
    TT.xls = 2 + AWP_NGSL;
    TT.yls = 2 + AWP_NGSL;
    TT.nj = T->ny;
    TT.nzt = T->nz;
  }
  if (texmode == TOPO_TEX_RULE)
    hipLaunchKernelGGL((topo_str_q_kernel<ST, TOPO_TEX_RULE>), gq, dim3(64, 4), 0, st, A, G,
                       rates, rstride, qchunk, TT);
  else if (texmode == TOPO_TEX_TABLE)
    hipLaunchKernelGGL((topo_str_q_kernel<ST, TOPO_TEX_TABLE>), gq, dim3(64, 4), 0, st, A, G,
                       rates, rstride, qchunk, TT);
  else
    hipLaunchKernelGGL((topo_str_q_kernel<ST, TOPO_TEX_ARRAYS>), gq, dim3(64, 4), 0, st, A, G,
                       rates, rstride, qchunk, TT);
  awp_check_launch(where);
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

void topo_set_constants(topo_t* T) {
  if (!T) {
    awp_set_error(AWP_ERR_ARG, "topo_set_constants: topo_t is NULL");
    return;
  }
  // set_constants(gridspacing, dth * gridspacing, ...) (stress_attenuation.cu:19-72)
  const float h = T->gridspacing;
  const float dt = T->dth * h;
  g_topo_constants.DT = dt;
  g_topo_constants.dth = dt / h;
  g_topo_constants.dh1 = (float)(1.0 / (double)h);
  g_topo_constants.set = 1;
}

void topo_velocity_interior_H(topo_t* T) {
  topo_velocity_launch<StF32>(T, WaveScales{}, "topo_velocity_interior_H");
}

void topo_velocity_front_H(topo_t* T) {
  if (T && T->use) neighbourless(T->y_rank_f, "topo_velocity_front_H");
}

void topo_velocity_back_H(topo_t* T) {
  if (T && T->use) neighbourless(T->y_rank_b, "topo_velocity_back_H");
}

void topo_stress_interior_H(topo_t* T) {
  topo_stress_launch<StF32>(T, WaveScales{}, "topo_stress_interior_H", nullptr);
}

void awp_topo_stress_interior_rates_H(topo_t* T, float* rates) {
  if (!rates) {
    awp_set_error(AWP_ERR_ARG, "awp_topo_stress_interior_rates_H: rates is NULL");
    return;
  }
  topo_stress_launch<StF32>(T, WaveScales{}, "awp_topo_stress_interior_rates_H", rates);
}

void awp_topo_stress_interior_tex_H(topo_t* T, float* rates, const float* tex, int rule) {
  const char* W = "awp_topo_stress_interior_tex_H";
  if (!rates || !tex) {
    awp_set_error(AWP_ERR_ARG, "%s: rates and tex must not be NULL", W);
    return;
  }
  topo_stress_launch<StF32>(T, WaveScales{}, W, rates, tex,
                            rule ? TOPO_TEX_RULE : TOPO_TEX_TABLE);
}

void topo_stress_left_H(topo_t* T) {
  if (T && T->use) neighbourless(T->x_rank_l, "topo_stress_left_H");
}

void topo_stress_right_H(topo_t* T) {
  if (T && T->use) neighbourless(T->x_rank_r, "topo_stress_right_H");
}

void awp_topo_velocity_s16(topo_t* T, int fmt, const float* scale) {
  WaveScales W;
  if (!awp_scales_from_host("awp_topo_velocity_s16", fmt, scale, &W)) return;
  if (fmt == AWP_FMT_BF16)
    topo_velocity_launch<StBF16>(T, W, "awp_topo_velocity_s16");
  else
    topo_velocity_launch<StF16>(T, W, "awp_topo_velocity_s16");
}

void awp_topo_stress_s16(topo_t* T, int fmt, const float* scale, float* rates) {
  WaveScales W;
  if (!awp_scales_from_host("awp_topo_stress_s16", fmt, scale, &W)) return;
  if (fmt == AWP_FMT_BF16)
    topo_stress_launch<StBF16>(T, W, "awp_topo_stress_s16", rates);
  else
    topo_stress_launch<StF16>(T, W, "awp_topo_stress_s16", rates);
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

void awp_topo_metrics_f(const f_grid_t* mf, int nx, int ny, float h, hipStream_t St) {
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

void awp_topo_metrics_g(const g_grid_t* mg, int nz, float h, hipStream_t St) {
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
