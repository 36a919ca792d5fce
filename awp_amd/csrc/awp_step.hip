// One NVE = 1 time step of a single Cartesian grid in one sweep: the velocity
// update (reference dvelcx_opt, src/awp/kernel.cu:125-268), the free-surface
// velocity ghosts (:1087-1112) and the stress + memory-variable update
// (dstrqc, :880-1197), with the relaxation texture as a table
// (init_texture, src/awp/mesh.c:1103-1137), gfx950.
//
// The separate launchers move 52 B (velocity) + 140 B (stress) per cell: the
// stresses and velocities are each read twice.  Here the state before the
// step is read once and the state after it written once (156 B/cell; 144 B
// with the texture table, 140 B with init_texture's rule): the step reads
// state A (u v w and the six stresses) and writes state B, and the memory
// variables are updated in place.  A and B must be distinct; B must hold A's
// values in every cell the step does not write (initialise both from one
// state: they then stay equal there).
//
// Mapping: a workgroup of 16 waves owns 12 j rows x 64 k (one wave64 per
// row) and marches along i through a chunk of planes.  The stress at plane i
// needs the new velocities at planes i-2 .. i+2 of its own column (kept in
// registers) and at plane i in a cross around it: j-2 .. j+2 and k-2 .. k+2.
// Those are computed in the workgroup, the rows outside the tile by four
// halo waves and the levels above and below the tile by the same halo waves
// (lanes 0..11 of a second item), so no workgroup waits for another.  Plane
// i+2's velocities go to a 4-slot LDS ring (68 levels x 16 rows x 3
// components per slot); one barrier per plane orders the ring.
//
// Every value is computed with the expression trees of dvelcx_H_opt,
// dstrqc_H's free-surface ghosts and dstrqc_H (built with -ffp-contract=off),
// so state B equals what dvelcx_H_opt + dstrqc_H leave in place, bit for bit.
#include <type_traits>

#include "awp_common.h"
#include "awp_tuning.h"
#include "../../include/awp_hip.h"

namespace {

constexpr int FBJ = 12;          // core rows per workgroup
constexpr int FNW = FBJ + 4;     // waves per workgroup
constexpr int FR = FBJ + 4;      // LDS rows: j0-2 .. j0+FBJ+1
constexpr int FC = 68;           // LDS levels: K0-2 .. K0+65
constexpr int FSL = 4;           // LDS ring slots (planes)
constexpr int FSLOT = 3 * FR * FC;

struct StepArgs {
  const float* a[9];  // u1 v1 w1 xx yy zz xy xz yz before the step
  float* b[9];        // the same after the step
  float* r[6];        // r1..r6, in place
  const float *d1, *lam, *mu, *qp, *qs, *coeff, *dcrjx, *dcrjy, *dcrjz, *lam_mu;
  const int* ww;      // null: init_texture's parity rule
};

struct StepLaunch {
  int xls, xre, yls, yre, top;  // update region: i, j in [ls, re], k in [align, top]
  int nkt, njt, ichunk, ni, nblocks, remap, jfast;
  int yline, lam_mu_ld, nxt;
  long long slice;
  float c1, c2, dth, dh1, DT;
  float vx1[8], vx2[8], wwo[8];  // texture table, by ww - 1
  int tex_nj, tex_nzt;           // parity rule: interior rows and levels
};

// Production Q weight (kernel.cu:978-1045), as dstrqc_H.
__device__ __forceinline__ float qw(float q, float ca, float cb, float wwo) {
  float w;
  const double q2 = (double)q * 2.0;
  if (1. / q2 <= 200.0)
    w = (float)((double)ca * (2. * (double)q) * (2. * (double)q) + (double)cb * (2. * (double)q));
  else
    w = 2.0f * wwo * q;
  return w / wwo;
}

// x window of the velocity update of one (j, k) column at plane m:
// xx at m-2..m, xy and xz at m-1..m+1, density corners (j,k) (j-1,k) (j,k-1)
// (j-1,k-1) at plane m.
struct VelWin {
  float xx_m2, xx_m1, xx_0, xy_m1, xy_0, xy_p1, xz_m1, xz_0, xz_p1, dc, djm, dkm, djk;
};

struct Fields {
  const float *__restrict__ u, *__restrict__ v, *__restrict__ w, *__restrict__ xx,
      *__restrict__ yy, *__restrict__ zz, *__restrict__ xy, *__restrict__ xz, *__restrict__ yz,
      *__restrict__ d1;
};

// Field access at (plane m + di, row j + dj, level k + dk) as a 32-bit byte
// offset from the field's base: global_load v, v_off, s[base] with the base
// a kernel argument in scalar registers.  The offsets of the stencil (di, dj)
// shifts are the same for every field, so they are formed once per plane;
// dk shifts go into the instruction's immediate.  The launcher checks that a
// field spans less than 4 GiB.
template <bool UNI>
struct Acc {
  unsigned b;       // byte offset of (m, j, k)
  unsigned S4, Y4;  // plane and row strides in bytes
  __device__ __forceinline__ Acc(int m, int j, int k, long long S, int Y)
      : b((unsigned)(((long long)m * S + (long long)j * Y + k) * 4)), S4((unsigned)(S * 4)),
        Y4((unsigned)Y * 4u) {}
  __device__ __forceinline__ unsigned off(int di, int dj, int dk) const {
    return b + (unsigned)di * S4 + (unsigned)dj * Y4 + (unsigned)dk * 4u;
  }
  __device__ __forceinline__ float operator()(const float* __restrict__ a, int di, int dj,
                                              int dk) const {
    return awp_ld(a, off(di, dj, dk));
  }
};

template <bool UNI>
__device__ __forceinline__ void vw_prime(VelWin& W, const Fields& F, const Acc<UNI>& X) {
  W.xx_m2 = X(F.xx, -2, 0, 0); W.xx_m1 = X(F.xx, -1, 0, 0); W.xx_0 = X(F.xx, 0, 0, 0);
  W.xy_m1 = X(F.xy, -1, 0, 0); W.xy_0 = X(F.xy, 0, 0, 0); W.xy_p1 = X(F.xy, 1, 0, 0);
  W.xz_m1 = X(F.xz, -1, 0, 0); W.xz_0 = X(F.xz, 0, 0, 0); W.xz_p1 = X(F.xz, 1, 0, 0);
  W.dc = X(F.d1, 0, 0, 0); W.djm = X(F.d1, 0, -1, 0); W.dkm = X(F.d1, 0, 0, -1);
  W.djk = X(F.d1, 0, -1, -1);
}

// Velocities at plane m (X's plane) when upd, else the stored (unchanging)
// values; the window advances to plane m+1 either way.
template <bool UNI>
__device__ __forceinline__ void vw_step(VelWin& W, const Fields& F, const Acc<UNI>& X, bool upd,
                                        float c1, float c2, float dth, float dcrj, float& u,
                                        float& v, float& w) {
  const float xx_p1 = X(F.xx, 1, 0, 0);
  const float xy_p2 = X(F.xy, 2, 0, 0);
  const float xz_p2 = X(F.xz, 2, 0, 0);
  const float ec = X(F.d1, 1, 0, 0), ejm = X(F.d1, 1, -1, 0), ekm = X(F.d1, 1, 0, -1),
              ejk = X(F.d1, 1, -1, -1);
  if (upd) {
    const float xy_jm1 = X(F.xy, 0, -1, 0), xy_jp1 = X(F.xy, 0, 1, 0), xy_jm2 = X(F.xy, 0, -2, 0);
    const float xz_km1 = X(F.xz, 0, 0, -1), xz_kp1 = X(F.xz, 0, 0, 1), xz_km2 = X(F.xz, 0, 0, -2);
    const float yy_0 = X(F.yy, 0, 0, 0), yy_jp1 = X(F.yy, 0, 1, 0), yy_jp2 = X(F.yy, 0, 2, 0),
                yy_jm1 = X(F.yy, 0, -1, 0);
    const float yz_0 = X(F.yz, 0, 0, 0), yz_km1 = X(F.yz, 0, 0, -1), yz_kp1 = X(F.yz, 0, 0, 1),
                yz_km2 = X(F.yz, 0, 0, -2);
    const float yz_jm1 = X(F.yz, 0, -1, 0), yz_jp1 = X(F.yz, 0, 1, 0), yz_jm2 = X(F.yz, 0, -2, 0);
    const float zz_0 = X(F.zz, 0, 0, 0), zz_kp1 = X(F.zz, 0, 0, 1), zz_kp2 = X(F.zz, 0, 0, 2),
                zz_km1 = X(F.zz, 0, 0, -1);
    float f_d1 = 0.25f * (W.dc + W.djm + W.dkm + W.djk);
    float f_d2 = 0.25f * (W.dc + ec + W.dkm + ekm);
    float f_d3 = 0.25f * (W.dc + ec + W.djm + ejm);
    f_d1 = dth / f_d1;
    f_d2 = dth / f_d2;
    f_d3 = dth / f_d3;
    u = (X(F.u, 0, 0, 0) + f_d1 * (c1 * (W.xx_0 - W.xx_m1) + c2 * (xx_p1 - W.xx_m2) +
                                   c1 * (W.xy_0 - xy_jm1) + c2 * (xy_jp1 - xy_jm2) +
                                   c1 * (W.xz_0 - xz_km1) + c2 * (xz_kp1 - xz_km2))) * dcrj;
    v = (X(F.v, 0, 0, 0) + f_d2 * (c1 * (W.xy_p1 - W.xy_0) + c2 * (xy_p2 - W.xy_m1) +
                                   c1 * (yy_jp1 - yy_0) + c2 * (yy_jp2 - yy_jm1) +
                                   c1 * (yz_0 - yz_km1) + c2 * (yz_kp1 - yz_km2))) * dcrj;
    w = (X(F.w, 0, 0, 0) + f_d3 * (c1 * (W.xz_p1 - W.xz_0) + c2 * (xz_p2 - W.xz_m1) +
                                   c1 * (yz_0 - yz_jm1) + c2 * (yz_jp1 - yz_jm2) +
                                   c1 * (zz_kp1 - zz_0) + c2 * (zz_kp2 - zz_km1))) * dcrj;
  } else {
    u = X(F.u, 0, 0, 0);
    v = X(F.v, 0, 0, 0);
    w = X(F.w, 0, 0, 0);
  }
  W.xx_m2 = W.xx_m1; W.xx_m1 = W.xx_0; W.xx_0 = xx_p1;
  W.xy_m1 = W.xy_0; W.xy_0 = W.xy_p1; W.xy_p1 = xy_p2;
  W.xz_m1 = W.xz_0; W.xz_0 = W.xz_p1; W.xz_p1 = xz_p2;
  W.dc = ec; W.djm = ejm; W.dkm = ekm; W.djk = ejk;
}

__device__ __forceinline__ int lidx(int slot, int comp, int r, int c) {
  return ((slot * 3 + comp) * FR + r) * FC + c;
}

template <bool RULE>
__global__ __launch_bounds__(64 * FNW) void step_kernel(const StepArgs A, const StepLaunch L) {
  __shared__ float s_coeff[16];
  __shared__ float s_tex[24];
  __shared__ float vl[FSL * FSLOT];

  const int lane = (int)threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const int tid = wave * 64 + lane;
  if (tid < 16) s_coeff[tid] = A.coeff[tid];
  if (tid >= 32 && tid < 56) {
    const int t = tid - 32;
    s_tex[t] = t < 8 ? L.vx1[t] : t < 16 ? L.vx2[t - 8] : L.wwo[t - 16];
  }

  const int b = L.remap ? awp_xcd_remap(blockIdx.x, L.nblocks) : (int)blockIdx.x;
  int kt, jt, ic;
  awp_tile3(b, L.nkt, L.njt, L.jfast, kt, jt, ic);
  const int K0 = AWP_ALIGN + 64 * kt;
  const int j0 = L.yls + jt * FBJ;
  const int ia = L.xls + ic * L.ichunk;
  int nsteps = L.ni - ic * L.ichunk;
  if (nsteps > L.ichunk) nsteps = L.ichunk;
  const int ib = ia + nsteps - 1;

  const int Y = L.yline;
  const long long S = L.slice;
  const float c1 = L.c1, c2 = L.c2, dth = L.dth, dh1 = L.dh1, DT = L.DT;
  const Fields F = {A.a[0], A.a[1], A.a[2], A.a[3], A.a[4], A.a[5], A.a[6], A.a[7], A.a[8], A.d1};
  const int top = L.top;

  const bool core = wave < FBJ;
  // item 0: every wave's row item (core row, or one of the four halo rows)
  int r0;
  if (core) r0 = wave + 2;
  else {
    const int h = wave - FBJ;
    r0 = h < 2 ? h : FBJ + h;
  }
  const int j = j0 - 2 + r0;
  const int c0 = lane + 2;
  const int k = K0 + lane;
  // item 1 (halo waves, lanes 0..FBJ-1): level column hc of core row 2+lane
  const int hc = core ? 0 : ((wave - FBJ) < 2 ? (wave - FBJ) : (wave - FBJ) + 64);
  const bool has1 = !core && lane < FBJ;
  const int r1 = 2 + lane;
  const int j1 = j0 + lane;
  const int k1 = K0 - 2 + hc;

  // which items exist and which are updated (the velocity region)
  const bool need0 = j <= L.yre + 2 && k <= top + 2;
  const bool sp0 = j >= L.yls && j <= L.yre && k >= AWP_ALIGN && k <= top;
  const bool need1 = has1 && j1 <= L.yre + 2 && k1 <= top + 2;
  const bool sp1 = has1 && j1 >= L.yls && j1 <= L.yre && k1 >= AWP_ALIGN && k1 <= top;
  const long long o0 = (long long)j * Y + k;
  const long long o1 = (long long)j1 * Y + k1;
  const float dy0 = sp0 ? A.dcrjy[j] : 0.0f, dz0 = sp0 ? A.dcrjz[k] : 0.0f;
  const float dy1 = sp1 ? A.dcrjy[j1] : 0.0f, dz1 = sp1 ? A.dcrjz[k1] : 0.0f;

  // stress lanes: core rows in the stress region
  const bool sact = core && j >= L.yls && j <= L.yre && k >= AWP_ALIGN + 3 && k <= top;
  const bool surf = sact && k == top;

  __syncthreads();  // s_coeff, s_tex

  // One velocity item at plane m: the update where the item is in the
  // velocity region (and m in [xls, xre]), else the stored value; into the
  // LDS ring at (row r, level c), and returned.
  auto vel_item = [&](auto uni, VelWin& Wn, bool need, bool sp, int jj, int kk, float dy,
                      float dz, int r, int c, int m, float& u, float& v, float& w) {
    constexpr bool UNI = decltype(uni)::value;
    if (!need) return;
    const Acc<UNI> X(m, jj, kk, S, Y);
    if (sp) {
      const bool in_x = m >= L.xls && m <= L.xre;
      const float dcrj = A.dcrjx[m] * dy * dz;
      vw_step<UNI>(Wn, F, X, in_x, c1, c2, dth, dcrj, u, v, w);
    } else {
      u = X(F.u, 0, 0, 0); v = X(F.v, 0, 0, 0); w = X(F.w, 0, 0, 0);
    }
    const int slot = m & (FSL - 1);
    vl[lidx(slot, 0, r, c)] = u;
    vl[lidx(slot, 1, r, c)] = v;
    vl[lidx(slot, 2, r, c)] = w;
  };
  using UniT = std::integral_constant<bool, true>;
  using LaneT = std::integral_constant<bool, false>;

  if (!core) {
    // halo waves: their row item and (lanes 0..FBJ-1) one level column of
    // the core rows, planes ia .. ib+2; the same barriers as the core waves
    VelWin W0, W1;
    if (sp0) vw_prime<true>(W0, F, Acc<true>(ia, j, k, S, Y));
    if (sp1) vw_prime<false>(W1, F, Acc<false>(ia, j1, k1, S, Y));
    float u, v, w;
    for (int m = ia; m <= ia + 1; ++m) {
      vel_item(UniT(), W0, need0, sp0, j, k, dy0, dz0, r0, c0, m, u, v, w);
      vel_item(LaneT(), W1, need1, sp1, j1, k1, dy1, dz1, r1, hc, m, u, v, w);
    }
    __syncthreads();
    for (int s = 0; s < nsteps; ++s) {
      const int m = __builtin_amdgcn_readfirstlane(ia + s + 2);
      vel_item(UniT(), W0, need0, sp0, j, k, dy0, dz0, r0, c0, m, u, v, w);
      vel_item(LaneT(), W1, need1, sp1, j1, k1, dy1, dz1, r1, hc, m, u, v, w);
      __syncthreads();
    }
    return;
  }

  // core waves: velocity of the own column, planes ia-2 .. ib+2, and the
  // stress of planes ia .. ib
  VelWin W0;
  if (sp0) vw_prime<true>(W0, F, Acc<true>(ia - 2, j, k, S, Y));
  // the core row's velocities at plane m; owned cells are written to B
  auto vel_core = [&](int m, float& u, float& v, float& w) {
    vel_item(UniT(), W0, need0, sp0, j, k, dy0, dz0, r0, c0, m, u, v, w);
    if (sp0 && m >= ia && m <= ib) {
      const unsigned ov = Acc<true>(m, j, k, S, Y).off(0, 0, 0);
      awp_st(A.b[0], ov, u); awp_st(A.b[1], ov, v); awp_st(A.b[2], ov, w);
    }
  };

  // register windows of the stress: u at i-1..i+2, v and w at i-2..i+1
  float u_m1 = 0.f, u_0 = 0.f, u_p1 = 0.f;
  float v_m2 = 0.f, v_m1 = 0.f, v_0 = 0.f, v_p1 = 0.f;
  float w_m2 = 0.f, w_m1 = 0.f, w_0 = 0.f, w_p1 = 0.f;
  {
    float u = 0.f, v = 0.f, w = 0.f;
    vel_core(ia - 2, u, v, w);
    v_m2 = v; w_m2 = w;
    vel_core(ia - 1, u, v, w);
    u_m1 = u; v_m1 = v; w_m1 = w;
    vel_core(ia, u, v, w);
    u_0 = u; v_0 = v; w_0 = w;
    vel_core(ia + 1, u, v, w);
    u_p1 = u; v_p1 = v; w_p1 = w;
  }
  __syncthreads();

  // material corners at plane ia
  const long long so = o0;
  float lc = 0.f, ljm = 0.f, lkm = 0.f, ljk = 0.f, mc = 0.f, mjm = 0.f, mkm = 0.f, mjk = 0.f;
  float pc = 0.f, pjm = 0.f, pkm = 0.f, pjk = 0.f, sc = 0.f, sjm = 0.f, skm = 0.f, sjk = 0.f;
  if (sact) {
    const Acc<true> X(ia, j, k, S, Y);
    lc = X(A.lam, 0, 0, 0); ljm = X(A.lam, 0, -1, 0); lkm = X(A.lam, 0, 0, -1);
    ljk = X(A.lam, 0, -1, -1);
    mc = X(A.mu, 0, 0, 0); mjm = X(A.mu, 0, -1, 0); mkm = X(A.mu, 0, 0, -1);
    mjk = X(A.mu, 0, -1, -1);
    pc = X(A.qp, 0, 0, 0); pjm = X(A.qp, 0, -1, 0); pkm = X(A.qp, 0, 0, -1);
    pjk = X(A.qp, 0, -1, -1);
    sc = X(A.qs, 0, 0, 0); sjm = X(A.qs, 0, -1, 0); skm = X(A.qs, 0, 0, -1);
    sjk = X(A.qs, 0, -1, -1);
  }

  for (int s = 0; s < nsteps; ++s) {
    const int i = __builtin_amdgcn_readfirstlane(ia + s);
    // the stress operands that do not depend on the new velocities are
    // loaded before the velocity update, so both loads overlap
    float o_s[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, o_r[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (sact) {
      const Acc<true> X(i, j, k, S, Y);
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        o_s[c] = X(A.a[3 + c], 0, 0, 0);
        o_r[c] = X(A.r[c], 0, 0, 0);
      }
    }
    float u_p2 = 0.f, v_n = 0.f, w_n = 0.f;
    vel_core(i + 2, u_p2, v_n, w_n);
    __syncthreads();
    const int slot = i & (FSL - 1);
    const long long pl = (long long)i * S;
    if (surf) {
      // free-surface ghosts at (i, j, top+1) (fs_ghost: dstrqc_H's pass)
      const float uc = u_0, vc = v_0, wc = w_0;
      const float ug = uc - (wc - w_m1);
      const float vg = vc - (vl[lidx(slot, 2, r0 + 1, c0)] - wc);
      float vs1, vs2;
      if (i - AWP_NGSL - 1 < L.nxt)
        vs1 = u_p1 - (w_p1 - wc);
      else
        vs1 = 0.0f;
      const float v_jm = vl[lidx(slot, 1, r0 - 1, c0)];
      if (j - AWP_NGSL - 1 > 1)
        vs2 = v_jm - (wc - vl[lidx(slot, 2, r0 - 1, c0)]);
      else
        vs2 = 0.0f;
      const float wg = vl[lidx(slot, 2, r0, c0 - 1)] -
                       A.lam_mu[(long long)i * L.lam_mu_ld + j] *
                           ((vs1 - ug) + (u_p1 - uc) + (vg - vs2) + (vc - v_jm));
      vl[lidx(slot, 0, r0, c0 + 1)] = ug;
      vl[lidx(slot, 1, r0, c0 + 1)] = vg;
      vl[lidx(slot, 2, r0, c0 + 1)] = wg;
      const unsigned og = Acc<true>(i, j, k + 1, S, Y).off(0, 0, 0);
      awp_st(A.b[0], og, ug); awp_st(A.b[1], og, vg); awp_st(A.b[2], og, wg);
    }
    if (sact) {
      const Acc<true> X(i, j, k, S, Y);
      const unsigned ob = X.off(0, 0, 0);
      const float Lc = X(A.lam, 1, 0, 0), Ljm = X(A.lam, 1, -1, 0), Lkm = X(A.lam, 1, 0, -1),
                  Ljk = X(A.lam, 1, -1, -1);
      const float Mc = X(A.mu, 1, 0, 0), Mjm = X(A.mu, 1, -1, 0), Mkm = X(A.mu, 1, 0, -1),
                  Mjk = X(A.mu, 1, -1, -1);
      const float Pc = X(A.qp, 1, 0, 0), Pjm = X(A.qp, 1, -1, 0), Pkm = X(A.qp, 1, 0, -1),
                  Pjk = X(A.qp, 1, -1, -1);
      const float Sc = X(A.qs, 1, 0, 0), Sjm = X(A.qs, 1, -1, 0), Skm = X(A.qs, 1, 0, -1),
                  Sjk = X(A.qs, 1, -1, -1);
      // in-plane velocity neighbours from the ring
      const float u_jp1 = vl[lidx(slot, 0, r0 + 1, c0)], u_jp2 = vl[lidx(slot, 0, r0 + 2, c0)],
                  u_jm1 = vl[lidx(slot, 0, r0 - 1, c0)];
      const float u_kp1 = vl[lidx(slot, 0, r0, c0 + 1)], u_kp2 = vl[lidx(slot, 0, r0, c0 + 2)],
                  u_km1 = vl[lidx(slot, 0, r0, c0 - 1)];
      const float v_jm1 = vl[lidx(slot, 1, r0 - 1, c0)], v_jp1 = vl[lidx(slot, 1, r0 + 1, c0)],
                  v_jm2 = vl[lidx(slot, 1, r0 - 2, c0)];
      const float v_kp1 = vl[lidx(slot, 1, r0, c0 + 1)], v_kp2 = vl[lidx(slot, 1, r0, c0 + 2)],
                  v_km1 = vl[lidx(slot, 1, r0, c0 - 1)];
      const float w_km1 = vl[lidx(slot, 2, r0, c0 - 1)], w_kp1 = vl[lidx(slot, 2, r0, c0 + 1)],
                  w_km2 = vl[lidx(slot, 2, r0, c0 - 2)];
      const float w_jp1 = vl[lidx(slot, 2, r0 + 1, c0)], w_jp2 = vl[lidx(slot, 2, r0 + 2, c0)],
                  w_jm1 = vl[lidx(slot, 2, r0 - 1, c0)];

      int f_ww;
      if (RULE) {
        const int ni = i - L.xls, col = ni * L.tex_nj + (j - L.yls);
        f_ww = 1 + 4 * ((ni + 1) & 1) + 2 * ((col + 1) & 1) +
               ((L.tex_nzt + col * L.tex_nzt + (k - AWP_ALIGN)) & 1);
      } else {
        f_ww = awp_ldi(A.ww, ob);
      }
      const int tt = (f_ww - 1) & 7;
      const float f_vx1 = s_tex[tt], f_vx2 = s_tex[8 + tt], f_wwo = s_tex[16 + tt];
      const float ca = s_coeff[(f_ww * 2 - 2) & 15], cb = s_coeff[(f_ww * 2 - 1) & 15];
      const float f_dcrj = A.dcrjx[i] * dy0 * dz0;

      const float sl8 = lc + Lc + ljm + Ljm + lkm + Lkm + ljk + Ljk;
      const float sm8 = mc + Mc + mjm + Mjm + mkm + Mkm + mjk + Mjk;
      const float sp8 = pc + Pc + pjm + Pjm + pkm + Pkm + pjk + Pjk;
      const float ss8 = sc + Sc + sjm + Sjm + skm + Skm + sjk + Sjk;
      float xl = awp_quot(8.0f, sl8);
      float xm = awp_quot(16.0f, sm8);
      float xmu1 = awp_quot(2.0f, mc + mkm);
      float xmu2 = awp_quot(2.0f, mc + mjm);
      float xmu3 = awp_quot(2.0f, mc + Mc);
      xl = xl + xm;
      float qpa = 0.0625f * sp8;
      const float qpaw = qw(qpa, ca, cb, f_wwo);
      float h = 0.0625f * ss8;
      const float hw = qw(h, ca, cb, f_wwo);
      float h1 = 0.250f * (sc + skm);
      const float h1w = qw(h1, ca, cb, f_wwo);
      float h2 = 0.250f * (sc + sjm);
      const float h2w = qw(h2, ca, cb, f_wwo);
      float h3 = 0.250f * (sc + Sc);
      const float h3w = qw(h3, ca, cb, f_wwo);
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
      const float vx1p = DT * (1 + f_vx2 * f_vx1);

      float vs1 = c1 * (u_p1 - u_0) + c2 * (u_p2 - u_m1);
      float vs2 = c1 * (v_0 - v_jm1) + c2 * (v_jp1 - v_jm2);
      float vs3 = c1 * (w_0 - w_km1) + c2 * (w_kp1 - w_km2);
      float tmp = xl * (vs1 + vs2 + vs3);
      const float a1 = qpa * (vs1 + vs2 + vs3);
      tmp = tmp + DT * a1;

      float f_r, f_rtmp, t;
      const float one_m_vx1 = 1 - f_vx1;
      const float wwo_m1 = f_wwo - 1;

      f_r = o_r[0];
      f_rtmp = -h * (vs2 + vs3) + a1;
      t = o_s[0] + tmp - xm * (vs2 + vs3) + vx1p * f_r;
      awp_st(A.r[0], ob, f_vx2 * f_r + f_wwo * f_rtmp);
      f_rtmp = f_rtmp * wwo_m1 + f_vx2 * f_r * one_m_vx1;
      awp_st(A.b[3], ob, (t + DT * f_rtmp) * f_dcrj);

      f_r = o_r[1];
      f_rtmp = -h * (vs1 + vs3) + a1;
      t = (o_s[1] + tmp - xm * (vs1 + vs3) + vx1p * f_r) * f_dcrj;
      awp_st(A.r[1], ob, f_vx2 * f_r + f_wwo * f_rtmp);
      f_rtmp = f_rtmp * wwo_m1 + f_vx2 * f_r * one_m_vx1;
      awp_st(A.b[4], ob, (t + DT * f_rtmp) * f_dcrj);

      f_r = o_r[2];
      f_rtmp = -h * (vs1 + vs2) + a1;
      t = (o_s[2] + tmp - xm * (vs1 + vs2) + vx1p * f_r) * f_dcrj;
      awp_st(A.r[2], ob, f_vx2 * f_r + f_wwo * f_rtmp);
      f_rtmp = f_rtmp * wwo_m1 + f_vx2 * f_r * one_m_vx1;
      awp_st(A.b[5], ob, (t + DT * f_rtmp) * f_dcrj);

      vs1 = c1 * (u_jp1 - u_0) + c2 * (u_jp2 - u_jm1);
      vs2 = c1 * (v_0 - v_m1) + c2 * (v_p1 - v_m2);
      f_r = o_r[3];
      f_rtmp = h1 * (vs1 + vs2);
      t = o_s[3] + xmu1 * (vs1 + vs2) + vx1p * f_r;
      awp_st(A.r[3], ob, f_vx2 * f_r + f_wwo * f_rtmp);
      f_rtmp = f_rtmp * wwo_m1 + f_vx2 * f_r * one_m_vx1;
      awp_st(A.b[6], ob, (t + DT * f_rtmp) * f_dcrj);

      vs1 = c1 * (u_kp1 - u_0) + c2 * (u_kp2 - u_km1);
      vs2 = c1 * (w_0 - w_m1) + c2 * (w_p1 - w_m2);
      f_r = o_r[4];
      f_rtmp = h2 * (vs1 + vs2);
      t = o_s[4] + xmu2 * (vs1 + vs2) + vx1p * f_r;
      awp_st(A.r[4], ob, f_vx2 * f_r + f_wwo * f_rtmp);
      f_rtmp = f_rtmp * wwo_m1 + f_vx2 * f_r * one_m_vx1;
      awp_st(A.b[7], ob, (t + DT * f_rtmp) * f_dcrj);

      vs1 = c1 * (v_kp1 - v_0) + c2 * (v_kp2 - v_km1);
      vs2 = c1 * (w_jp1 - w_0) + c2 * (w_jp2 - w_jm1);
      f_r = o_r[5];
      f_rtmp = h3 * (vs1 + vs2);
      t = o_s[5] + xmu3 * (vs1 + vs2) + vx1p * f_r;
      awp_st(A.r[5], ob, f_vx2 * f_r + f_wwo * f_rtmp);
      f_rtmp = f_rtmp * wwo_m1 + f_vx2 * f_r * one_m_vx1;
      awp_st(A.b[8], ob, (t + DT * f_rtmp) * f_dcrj);

      lc = Lc; ljm = Ljm; lkm = Lkm; ljk = Ljk;
      mc = Mc; mjm = Mjm; mkm = Mkm; mjk = Mjk;
      pc = Pc; pjm = Pjm; pkm = Pkm; pjk = Pjk;
      sc = Sc; sjm = Sjm; skm = Skm; sjk = Sjk;
    }
    u_m1 = u_0; u_0 = u_p1; u_p1 = u_p2;
    v_m2 = v_m1; v_m1 = v_0; v_0 = v_p1; v_p1 = v_n;
    w_m2 = w_m1; w_m1 = w_0; w_0 = w_p1; w_p1 = w_n;
  }
}

}  // namespace

extern "C" {

void awp_step(const float* const* a9, float* const* b9, float* const* r6,
              const float* const* med, const float* coeff, const float* tex, const int* ww,
              const float* dcrjx, const float* dcrjy, const float* dcrjz, const float* lam_mu,
              int nxt, int nyt, int nzt, hipStream_t St) {
  const char* W = "awp_step";
  if (!awp_grid_ok(0, W)) return;
  const AwpState& st = awp_state();
  if (st.ngrids != 1) {
    awp_set_error(AWP_ERR_STATE, "%s: one grid only (ngrids=%d)", W, st.ngrids);
    return;
  }
  const AwpGrid& g = st.grid[0];
  if (nxt != g.nxt || nyt != g.nyt || nzt != g.nzt) {
    awp_set_error(AWP_ERR_ARG, "%s: nxt=%d nyt=%d nzt=%d disagree with grid 0 (%d, %d, %d)", W,
                  nxt, nyt, nzt, g.nxt, g.nyt, g.nzt);
    return;
  }
  if (!a9 || !b9 || !r6 || !med || !coeff || !tex || !dcrjx || !dcrjy || !dcrjz || !lam_mu) {
    awp_set_error(AWP_ERR_ARG, "%s: a pointer argument is NULL", W);
    return;
  }
  StepArgs A;
  for (int f = 0; f < 9; ++f) {
    A.a[f] = a9[f];
    A.b[f] = b9[f];
    if (!A.a[f] || !A.b[f] || (const float*)A.b[f] == A.a[f]) {
      awp_set_error(AWP_ERR_ARG, "%s: field %d: the states before and after the step must be "
                    "distinct, non-NULL arrays", W, f);
      return;
    }
  }
  for (int f = 0; f < 6; ++f) A.r[f] = r6[f];
  A.d1 = med[0]; A.lam = med[1]; A.mu = med[2]; A.qp = med[3]; A.qs = med[4];
  A.coeff = coeff; A.dcrjx = dcrjx; A.dcrjy = dcrjy; A.dcrjz = dcrjz; A.lam_mu = lam_mu;
  A.ww = ww;

  // 32-bit byte offsets inside a field
  if ((long long)(nxt + 4 + AWP_NGSL2) * g.slice * 4 >= (1ll << 32)) {
    awp_set_error(AWP_ERR_ARG, "%s: a field of %d x %d x %d cells spans 4 GiB or more", W, nxt,
                  nyt, nzt);
    return;
  }
  StepLaunch L;
  L.xls = 2 + AWP_NGSL; L.xre = nxt + AWP_NGSL + 1;
  L.yls = 2 + AWP_NGSL; L.yre = nyt + AWP_NGSL + 1;
  L.top = nzt + AWP_ALIGN - 1;
  L.yline = g.yline; L.slice = g.slice;
  L.lam_mu_ld = nyt + 4 + AWP_NGSL2;
  L.nxt = nxt;
  L.c1 = g.c1; L.c2 = g.c2; L.dth = g.dth; L.dh1 = g.dh1; L.DT = g.DT;
  for (int m = 0; m < 8; ++m) {
    L.vx1[m] = tex[m];
    L.vx2[m] = tex[8 + m];
    L.wwo[m] = tex[16 + m];
  }
  L.tex_nj = nyt;
  L.tex_nzt = nzt;
  L.remap = awp_tuning().xcd_remap;
  L.jfast = awp_tuning().step_jfast;
  L.ni = L.xre - L.xls + 1;
  L.nkt = awp_cdiv(L.top - AWP_ALIGN + 1, 64);
  L.njt = awp_cdiv(L.yre - L.yls + 1, FBJ);
  int nic, ichunk;
  awp_split_i(L.ni, L.nkt * L.njt, awp_tuning().step_ichunk, awp_tuning().step_target_wgs, &nic,
              &ichunk);
  L.ichunk = ichunk;
  L.nblocks = L.nkt * L.njt * nic;
  if (ww)
    hipLaunchKernelGGL(step_kernel<false>, dim3(L.nblocks), dim3(64, FNW), 0, St, A, L);
  else
    hipLaunchKernelGGL(step_kernel<true>, dim3(L.nblocks), dim3(64, FNW), 0, St, A, L);
  awp_check_launch(W);
}

}  // extern "C"
