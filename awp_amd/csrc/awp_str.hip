// Stress + anelastic memory-variable update (reference dstrqc,
// src/awp/kernel.cu:880-1197, and dstrqc_new :343-859), free-surface
// velocity ghosts (:1087-1112) and stress imaging fstr (:1616-1647), gfx950.
//
// Mapping as in the velocity kernel: wave64 along k, BJ waves along j, an
// i-march with register pipelines.  All four 8-cell material means
// (lam, mu, qp, qs) keep their plane-(i+1) corners in registers, so every
// per-cell array is fetched once per march; u/v/w keep a 4-plane x window.
// The 16 Q coefficients sit in LDS and are indexed per lane by ww.
//
// Operation order and literal types follow the reference (double-literal
// sub-expressions are evaluated in double), built with -ffp-contract=off.
#include "awp_common.h"
#include "awp_store.h"
#include "awp_tuning.h"
#include "../../include/awp_hip.h"

namespace {

struct StrLaunch {
  int kb, k0, k1, j0, j1, i0;  // kb: k of lane 0 in tile 0; [k0, k1] active
  int nkt, njt, ichunk, ni, nblocks, remap, jfast;
  int yline;
  long long slice;
  float c1, c2, dth, dh1, DT;
};

// Production Q weight (kernel.cu:978-1045).
__device__ __forceinline__ float qw_prod(float q, float ca, float cb, float wwo) {
  float w;
  const double q2 = (double)q * 2.0;
  if (1. / q2 <= 200.0)
    w = (float)((double)ca * (2. * (double)q) * (2. * (double)q) + (double)cb * (2. * (double)q));
  else
    w = 2.0f * wwo * q;
  return w / wwo;
}

// dstrqc_new's Q weights (kernel.cu:505-604): float test; qpa keeps a double
// quadratic, the others are float.
__device__ __forceinline__ float qw_new(float q, float ca, float cb, float wwo, bool dbl) {
  float w;
  if (1.0f / (q * 2.0f) <= 200.0f) {
    if (dbl)
      w = (float)((double)ca * (2. * (double)q) * (2. * (double)q) + (double)cb * (2. * (double)q));
    else
      w = ca * (2.0f * q) * (2.0f * q) + cb * (2.0f * q);
  } else {
    w = 2.0f * wwo * q;
  }
  return w / wwo;
}

// OOP: the stresses before the step are read from s0 (xx yy zz xy xz yz) and
// the updated ones written to xx..yz (distinct arrays); otherwise xx..yz are
// updated in place and s0 is unused.  The memory variables are always
// updated in place.
struct CStress6 {
  const void* f[6];
};

// Relaxation texture.  init_texture (reference src/awp/mesh.c:1103-1137)
// gives every cell of the stress region one of 8 mechanisms, ww = 1..8, and
// vx1, vx2, wwo are functions of the mechanism alone.  TEX selects where the
// stress kernel takes them from:
//   TEX_ARRAYS  the per-cell vx1, vx2, wwo and ww arrays (dstrqc_H);
//   TEX_TABLE   the per-cell ww array and an 8-entry table (12 B/cell less);
//   TEX_RULE    ww from init_texture's parity rule and the table (16 B/cell
//               less): ni = i - xls, nj = j - yls, nk = k - align count
//               from the region origin, and
//               itx = (ni+1)&1, ity = (ni*NJ+nj+1)&1,
//               itz = (nzt + (ni*NJ+nj)*nzt + nk)&1, ww = 1 + 4itx + 2ity + itz.
enum { TEX_ARRAYS = 0, TEX_TABLE = 1, TEX_RULE = 2 };
struct TexTable {
  float vx1[8], vx2[8], wwo[8];  // indexed by ww - 1
  int xls, yls, nj, nzt;         // TEX_RULE only
};

// KW: k-run width (awp_pick_kw); with KW = 32 a wave covers two j rows.
template <int BJ, bool NEWLIT, bool OOP, class ST, int TEX = TEX_ARRAYS, int KW = 64>
__global__ __launch_bounds__(64 * BJ) void dstrqc_kernel(
    typename ST::T* __restrict__ xx, typename ST::T* __restrict__ yy,
    typename ST::T* __restrict__ zz, typename ST::T* __restrict__ xy,
    typename ST::T* __restrict__ xz, typename ST::T* __restrict__ yz, const CStress6 S0,
    typename ST::T* __restrict__ r1, typename ST::T* __restrict__ r2,
    typename ST::T* __restrict__ r3, typename ST::T* __restrict__ r4,
    typename ST::T* __restrict__ r5, typename ST::T* __restrict__ r6,
    const typename ST::T* __restrict__ u1, const typename ST::T* __restrict__ v1,
    const typename ST::T* __restrict__ w1,
    const float* __restrict__ lam, const float* __restrict__ mu, const float* __restrict__ qp,
    const float* __restrict__ coeff, const float* __restrict__ qs,
    const float* __restrict__ dcrjx, const float* __restrict__ dcrjy,
    const float* __restrict__ dcrjz, const float* __restrict__ vx1,
    const float* __restrict__ vx2, const int* __restrict__ ww, const float* __restrict__ wwo,
    const StrLaunch L, const WaveScales SC, const TexTable TT) {
  __shared__ float s_coeff[16];
  __shared__ float s_tex[24];
  const int tid = threadIdx.y * 64 + threadIdx.x;
  if (tid < 16) s_coeff[tid] = coeff[tid];
  if (TEX != TEX_ARRAYS && tid < 24)
    s_tex[tid] = tid < 8 ? TT.vx1[tid] : tid < 16 ? TT.vx2[tid - 8] : TT.wwo[tid - 16];
  __syncthreads();

  const int b = L.remap ? awp_xcd_remap(blockIdx.x, L.nblocks) : (int)blockIdx.x;
  int kt, jt, ic;
  awp_tile3(b, L.nkt, L.njt, L.jfast, kt, jt, ic);
  constexpr int RW = 64 / KW;  // j rows per wave
  const int lane = (int)threadIdx.x;
  const int k = L.kb + kt * KW + (KW == 64 ? lane : lane % KW);
  const int j = L.j0 + (jt * BJ + (int)threadIdx.y) * RW + (KW == 64 ? 0 : lane / KW);
  if (k < L.k0 || k > L.k1 || j > L.j1) return;
  const int ia = L.i0 + ic * L.ichunk;
  int nsteps = L.ni - ic * L.ichunk;
  if (nsteps > L.ichunk) nsteps = L.ichunk;

  const int Y = L.yline;
  const long long S = L.slice;
  const int o = j * Y + k;
  const float c1 = L.c1, c2 = L.c2, dth = L.dth, dh1 = L.dh1, DT = L.DT;
  const float dy = dcrjy[j];
  const float dz = dcrjz[k];

  typedef typename ST::T T;
  const T* const xx_in = OOP ? (const T*)S0.f[0] : xx;
  const T* const yy_in = OOP ? (const T*)S0.f[1] : yy;
  const T* const zz_in = OOP ? (const T*)S0.f[2] : zz;
  const T* const xy_in = OOP ? (const T*)S0.f[3] : xy;
  const T* const xz_in = OOP ? (const T*)S0.f[4] : xz;
  const T* const yz_in = OOP ? (const T*)S0.f[5] : yz;

  const float su = SC.s[AWP_SU], sv = SC.s[AWP_SV], sw = SC.s[AWP_SW];
  const long long p0 = (long long)ia * S + o;
  // x windows: u needs i-1..i+2, v and w need i-2..i+1
  float u_m1 = ST::ld(u1, p0 - S, su), u_0 = ST::ld(u1, p0, su), u_p1 = ST::ld(u1, p0 + S, su);
  float v_m2 = ST::ld(v1, p0 - 2 * S, sv), v_m1 = ST::ld(v1, p0 - S, sv), v_0 = ST::ld(v1, p0, sv);
  float w_m2 = ST::ld(w1, p0 - 2 * S, sw), w_m1 = ST::ld(w1, p0 - S, sw), w_0 = ST::ld(w1, p0, sw);
  // material corners at plane i: (j,k), (j-1,k), (j,k-1), (j-1,k-1)
  float lc = lam[p0], ljm = lam[p0 - Y], lkm = lam[p0 - 1], ljk = lam[p0 - Y - 1];
  float mc = mu[p0], mjm = mu[p0 - Y], mkm = mu[p0 - 1], mjk = mu[p0 - Y - 1];
  float pc = qp[p0], pjm = qp[p0 - Y], pkm = qp[p0 - 1], pjk = qp[p0 - Y - 1];
  float sc = qs[p0], sjm = qs[p0 - Y], skm = qs[p0 - 1], sjk = qs[p0 - Y - 1];

  for (int s = 0; s < nsteps; ++s) {
    const int i = ia + s;
    const long long pl = (long long)i * S;
    const typename ST::T* up = u1 + pl;
    const typename ST::T* vp = v1 + pl;
    const typename ST::T* wp = w1 + pl;
    const long long pn = pl + S;  // plane i+1

    // leads
    const float u_p2 = ST::ld(up, 2 * S + o, su);
    const float v_p1 = ST::ld(vp, S + o, sv);
    const float w_p1 = ST::ld(wp, S + o, sw);
    const float Lc = lam[pn + o], Ljm = lam[pn + o - Y], Lkm = lam[pn + o - 1], Ljk = lam[pn + o - Y - 1];
    const float Mc = mu[pn + o], Mjm = mu[pn + o - Y], Mkm = mu[pn + o - 1], Mjk = mu[pn + o - Y - 1];
    const float Pc = qp[pn + o], Pjm = qp[pn + o - Y], Pkm = qp[pn + o - 1], Pjk = qp[pn + o - Y - 1];
    const float Sc = qs[pn + o], Sjm = qs[pn + o - Y], Skm = qs[pn + o - 1], Sjk = qs[pn + o - Y - 1];

    // in-plane stencils
    const float u_jp1 = ST::ld(up, o + Y, su), u_jp2 = ST::ld(up, o + 2 * Y, su), u_jm1 = ST::ld(up, o - Y, su);
    const float u_kp1 = ST::ld(up, o + 1, su), u_kp2 = ST::ld(up, o + 2, su), u_km1 = ST::ld(up, o - 1, su);
    const float v_jm1 = ST::ld(vp, o - Y, sv), v_jp1 = ST::ld(vp, o + Y, sv), v_jm2 = ST::ld(vp, o - 2 * Y, sv);
    const float v_kp1 = ST::ld(vp, o + 1, sv), v_kp2 = ST::ld(vp, o + 2, sv), v_km1 = ST::ld(vp, o - 1, sv);
    const float w_km1 = ST::ld(wp, o - 1, sw), w_kp1 = ST::ld(wp, o + 1, sw), w_km2 = ST::ld(wp, o - 2, sw);
    const float w_jp1 = ST::ld(wp, o + Y, sw), w_jp2 = ST::ld(wp, o + 2 * Y, sw), w_jm1 = ST::ld(wp, o - Y, sw);

    const long long p = pl + o;
    float f_vx1, f_vx2, f_wwo;
    int f_ww;
    if (TEX == TEX_ARRAYS) {
      f_vx1 = vx1[p]; f_vx2 = vx2[p]; f_wwo = wwo[p];
      f_ww = ww[p];
    } else {
      if (TEX == TEX_TABLE) {
        f_ww = ww[p];
      } else {
        const int ni = i - TT.xls, col = ni * TT.nj + (j - TT.yls);
        f_ww = 1 + 4 * ((ni + 1) & 1) + 2 * ((col + 1) & 1) +
               ((TT.nzt + col * TT.nzt + (k - AWP_ALIGN)) & 1);
      }
      const int t = (f_ww - 1) & 7;
      f_vx1 = s_tex[t]; f_vx2 = s_tex[8 + t]; f_wwo = s_tex[16 + t];
    }
    const float ca = s_coeff[(f_ww * 2 - 2) & 15], cb = s_coeff[(f_ww * 2 - 1) & 15];
    const float f_dcrj = dcrjx[i] * dy * dz;

    const float sl8 = lc + Lc + ljm + Ljm + lkm + Lkm + ljk + Ljk;
    const float sm8 = mc + Mc + mjm + Mjm + mkm + Mkm + mjk + Mjk;
    const float sp8 = pc + Pc + pjm + Pjm + pkm + Pkm + pjk + Pjk;
    const float ss8 = sc + Sc + sjm + Sjm + skm + Skm + sjk + Sjk;
    float xl, xm, xmu1, xmu2, xmu3, qpa, h, h1, h2, h3, qpaw, hw, h1w, h2w, h3w;
    if (!NEWLIT) {
      xl = awp_quot(8.0f, sl8);
      xm = awp_quot(16.0f, sm8);
      xmu1 = awp_quot(2.0f, mc + mkm);
      xmu2 = awp_quot(2.0f, mc + mjm);
      xmu3 = awp_quot(2.0f, mc + Mc);
      xl = xl + xm;
      qpa = 0.0625f * sp8;
      qpaw = qw_prod(qpa, ca, cb, f_wwo);
      h = 0.0625f * ss8;
      hw = qw_prod(h, ca, cb, f_wwo);
      h1 = 0.250f * (sc + skm);
      h1w = qw_prod(h1, ca, cb, f_wwo);
      h2 = 0.250f * (sc + sjm);
      h2w = qw_prod(h2, ca, cb, f_wwo);
      h3 = 0.250f * (sc + Sc);
      h3w = qw_prod(h3, ca, cb, f_wwo);
    } else {
      xl = 8.0f / sl8;
      xm = 16.0f / sm8;
      xmu1 = 2.0f / (mc + mkm);
      xmu2 = awp_quot(2.0f, mc + mjm);
      xmu3 = awp_quot(2.0f, mc + Mc);
      xl = xl + xm;
      qpa = 0.0625f * sp8;
      qpaw = qw_new(qpa, ca, cb, f_wwo, true);
      h = 0.0625f * ss8;
      hw = qw_new(h, ca, cb, f_wwo, false);
      h1 = 0.250f * (sc + skm);
      h1w = qw_new(h1, ca, cb, f_wwo, false);
      h2 = 0.250f * (sc + sjm);
      h2w = qw_new(h2, ca, cb, f_wwo, false);
      h3 = 0.250f * (sc + Sc);
      h3w = qw_new(h3, ca, cb, f_wwo, false);
    }
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

    f_r = ST::ld(r1, p, SC.s[AWP_SR1]);
    f_rtmp = -h * (vs2 + vs3) + a1;
    t = ST::ld(xx_in, p, SC.s[AWP_SXX]) + tmp - xm * (vs2 + vs3) + vx1p * f_r;
    ST::st(r1, p, f_vx2 * f_r + f_wwo * f_rtmp, SC.inv[AWP_SR1]);
    f_rtmp = f_rtmp * wwo_m1 + f_vx2 * f_r * one_m_vx1;
    ST::st(xx, p, (t + DT * f_rtmp) * f_dcrj, SC.inv[AWP_SXX]);

    f_r = ST::ld(r2, p, SC.s[AWP_SR2]);
    f_rtmp = -h * (vs1 + vs3) + a1;
    t = (ST::ld(yy_in, p, SC.s[AWP_SYY]) + tmp - xm * (vs1 + vs3) + vx1p * f_r) * f_dcrj;
    ST::st(r2, p, f_vx2 * f_r + f_wwo * f_rtmp, SC.inv[AWP_SR2]);
    f_rtmp = f_rtmp * wwo_m1 + f_vx2 * f_r * one_m_vx1;
    ST::st(yy, p, (t + DT * f_rtmp) * f_dcrj, SC.inv[AWP_SYY]);

    f_r = ST::ld(r3, p, SC.s[AWP_SR3]);
    f_rtmp = -h * (vs1 + vs2) + a1;
    t = (ST::ld(zz_in, p, SC.s[AWP_SZZ]) + tmp - xm * (vs1 + vs2) + vx1p * f_r) * f_dcrj;
    ST::st(r3, p, f_vx2 * f_r + f_wwo * f_rtmp, SC.inv[AWP_SR3]);
    f_rtmp = f_rtmp * wwo_m1 + f_vx2 * f_r * one_m_vx1;
    ST::st(zz, p, (t + DT * f_rtmp) * f_dcrj, SC.inv[AWP_SZZ]);

    vs1 = c1 * (u_jp1 - u_0) + c2 * (u_jp2 - u_jm1);
    vs2 = c1 * (v_0 - v_m1) + c2 * (v_p1 - v_m2);
    f_r = ST::ld(r4, p, SC.s[AWP_SR4]);
    f_rtmp = h1 * (vs1 + vs2);
    t = ST::ld(xy_in, p, SC.s[AWP_SXY]) + xmu1 * (vs1 + vs2) + vx1p * f_r;
    ST::st(r4, p, f_vx2 * f_r + f_wwo * f_rtmp, SC.inv[AWP_SR4]);
    f_rtmp = f_rtmp * wwo_m1 + f_vx2 * f_r * one_m_vx1;
    ST::st(xy, p, (t + DT * f_rtmp) * f_dcrj, SC.inv[AWP_SXY]);

    vs1 = c1 * (u_kp1 - u_0) + c2 * (u_kp2 - u_km1);
    vs2 = c1 * (w_0 - w_m1) + c2 * (w_p1 - w_m2);
    f_r = ST::ld(r5, p, SC.s[AWP_SR5]);
    f_rtmp = h2 * (vs1 + vs2);
    t = ST::ld(xz_in, p, SC.s[AWP_SXZ]) + xmu2 * (vs1 + vs2) + vx1p * f_r;
    ST::st(r5, p, f_vx2 * f_r + f_wwo * f_rtmp, SC.inv[AWP_SR5]);
    f_rtmp = f_rtmp * wwo_m1 + f_vx2 * f_r * one_m_vx1;
    ST::st(xz, p, (t + DT * f_rtmp) * f_dcrj, SC.inv[AWP_SXZ]);

    vs1 = c1 * (v_kp1 - v_0) + c2 * (v_kp2 - v_km1);
    vs2 = c1 * (w_jp1 - w_0) + c2 * (w_jp2 - w_jm1);
    f_r = ST::ld(r6, p, SC.s[AWP_SR6]);
    f_rtmp = h3 * (vs1 + vs2);
    t = ST::ld(yz_in, p, SC.s[AWP_SYZ]) + xmu3 * (vs1 + vs2) + vx1p * f_r;
    ST::st(r6, p, f_vx2 * f_r + f_wwo * f_rtmp, SC.inv[AWP_SR6]);
    f_rtmp = f_rtmp * wwo_m1 + f_vx2 * f_r * one_m_vx1;
    ST::st(yz, p, (t + DT * f_rtmp) * f_dcrj, SC.inv[AWP_SYZ]);

    // advance x windows and material corners
    u_m1 = u_0; u_0 = u_p1; u_p1 = u_p2;
    v_m2 = v_m1; v_m1 = v_0; v_0 = v_p1;
    w_m2 = w_m1; w_m1 = w_0; w_0 = w_p1;
    lc = Lc; ljm = Ljm; lkm = Lkm; ljk = Ljk;
    mc = Mc; mjm = Mjm; mkm = Mkm; mjk = Mjk;
    pc = Pc; pjm = Pjm; pkm = Pkm; pjk = Pjk;
    sc = Sc; sjm = Sjm; skm = Skm; sjk = Sjk;
  }
}

// Free-surface velocity ghosts at k = top+1 (kernel.cu:1087-1112 for dstrqc,
// 660-687 for dstrqc_new).  In the reference the k=top lane of the stress
// kernel writes these and reads them back itself, and the k=top-1 lane
// rewrites u1, v1 at top+1 with the same formula; every value depends only
// on interior (k <= top) velocities, so computing them in a pass of their
// own before the stress sweep leaves the same memory state and the same
// stresses.  One lane per (i, j) column, lanes along j.
struct FsLaunch {
  int i0, i1, j0, j1, ktop, yline, lam_mu_ld, nxt, nyt, NX, NPC, rankx, ranky;
  long long slice;
};

template <bool NEWLIT, class ST>
__global__ __launch_bounds__(256) void fs_ghost_kernel(typename ST::T* __restrict__ u1,
                                                       typename ST::T* __restrict__ v1,
                                                       typename ST::T* __restrict__ w1,
                                                       const float* __restrict__ lam_mu,
                                                       const FsLaunch F, const WaveScales SC) {
  const int j = F.j0 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int i = F.i0 + (int)(blockIdx.y * 4 + threadIdx.y);
  if (j > F.j1 || i > F.i1) return;
  const long long S = F.slice;
  const int Y = F.yline;
  const long long p = (long long)i * S + (long long)j * Y + F.ktop;
  const float su = SC.s[AWP_SU], sv = SC.s[AWP_SV], sw = SC.s[AWP_SW];
  const float uc = ST::ld(u1, p, su), vc = ST::ld(v1, p, sv), wc = ST::ld(w1, p, sw);
  const float ug = uc - (wc - ST::ld(w1, p - S, sw));
  const float vg = vc - (ST::ld(w1, p + Y, sw) - wc);
  ST::st(u1, p + 1, ug, SC.inv[AWP_SU]);
  ST::st(v1, p + 1, vg, SC.inv[AWP_SV]);
  float vs1, vs2;
  int gi = F.nxt * F.rankx + i - AWP_NGSL - 1;
  if (gi < F.NX || (NEWLIT && F.NPC == 2))
    vs1 = ST::ld(u1, p + S, su) - (ST::ld(w1, p + S, sw) - wc);
  else
    vs1 = 0.0f;
  gi = F.nyt * F.ranky + j - AWP_NGSL - 1;
  if (gi > 1 || F.NPC == 2)
    vs2 = ST::ld(v1, p - Y, sv) - (wc - ST::ld(w1, p - Y, sw));
  else
    vs2 = 0.0f;
  ST::st(w1, p + 1,
         ST::ld(w1, p - 1, sw) - lam_mu[(long long)i * F.lam_mu_ld + j] *
                                     ((vs1 - ug) + (ST::ld(u1, p + S, su) - uc) + (vg - vs2) +
                                      (vc - ST::ld(v1, p - Y, sv))),
         SC.inv[AWP_SW]);
}

// f15: xx yy zz xy xz yz r1..r6 u1 v1 w1 in the storage type of ST; with OOP,
// S0 holds the stresses before the step and f15's stresses receive the update.
template <bool NEWLIT, bool OOP, class ST, int TEX = TEX_ARRAYS>
void dstrqc_common(const char* W, void* const* f15, const CStress6& S0, float* lam, float* mu,
                   float* qp,
                   float* coeff, float* qs, float* dcrjx, float* dcrjy, float* dcrjz, int nyt,
                   int nzt, hipStream_t St, float* lam_mu, float* vx1, float* vx2, int* ww,
                   float* wwo, int NX, int NPC, int rankx, int ranky, int s_i, int e_i, int s_j,
                   int e_j, int d_i, const WaveScales& SC, const TexTable& TT = TexTable{}) {
  typedef typename ST::T T;
  T* const* A = reinterpret_cast<T* const*>(f15);
  if (!awp_grid_ok(d_i, W)) return;
  const AwpGrid& g = awp_state().grid[d_i];
  if (nyt != g.nyt || nzt != g.nzt) {
    awp_set_error(AWP_ERR_ARG, "%s: nyt=%d nzt=%d disagree with grid %d (%d, %d)", W, nyt, nzt,
                  d_i, g.nyt, g.nzt);
    return;
  }
  if (s_i > e_i || s_j > e_j) return;
  if (s_i < 2 || e_i + 2 > g.nxt + 3 + AWP_NGSL2 || s_j < 2 || e_j + 2 > g.nyt + 3 + AWP_NGSL2) {
    awp_set_error(AWP_ERR_ARG, "%s: range i[%d,%d] j[%d,%d] leaves the padded grid (%d x %d)", W,
                  s_i, e_i, s_j, e_j, g.nxt, g.nyt);
    return;
  }
  if (d_i == 0) {
    FsLaunch F;
    F.i0 = s_i; F.i1 = e_i; F.j0 = s_j; F.j1 = e_j;
    F.ktop = nzt + AWP_ALIGN - 1;
    F.yline = g.yline; F.slice = g.slice;
    F.lam_mu_ld = nyt + 4 + AWP_NGSL2;
    F.nxt = g.nxt; F.nyt = g.nyt; F.NX = NX; F.NPC = NPC; F.rankx = rankx; F.ranky = ranky;
    dim3 grid(awp_cdiv(e_j - s_j + 1, 64), awp_cdiv(e_i - s_i + 1, 4));
    hipLaunchKernelGGL((fs_ghost_kernel<NEWLIT, ST>), grid, dim3(64, 4), 0, St, A[12], A[13],
                       A[14], lam_mu, F, SC);
    if (!awp_check_launch(W)) return;
  }
  StrLaunch L;
  L.k0 = AWP_ALIGN + 3;  // kernel.cu:901 (mink)
  // Tiles start on a 128-byte line (k = 32): whole-line loads and stores;
  // the three lanes below mink idle.
  L.kb = awp_tuning().kalign ? (L.k0 & ~31) : L.k0;
  L.k1 = (d_i == 0) ? nzt + AWP_ALIGN - 1 : nzt + AWP_ALIGN - 3;
  L.j0 = s_j;
  L.j1 = e_j;
  L.i0 = s_i;
  L.ni = e_i - s_i + 1;
  L.yline = g.yline;
  L.slice = g.slice;
  L.c1 = g.c1; L.c2 = g.c2; L.dth = g.dth; L.dh1 = g.dh1; L.DT = g.DT;
  L.remap = awp_tuning().xcd_remap;
  L.jfast = awp_tuning().str_jfast;
  if (L.k1 < L.k0) return;
  const int bj = awp_tuning().str_bj;
  const int kw = awp_pick_kw(L.k1 - L.kb + 1, awp_tuning().kw);
  L.nkt = awp_cdiv(L.k1 - L.kb + 1, kw);
  L.njt = awp_cdiv(L.j1 - L.j0 + 1, bj * (64 / kw));
  int nic, ichunk;
  awp_split_i(L.ni, L.nkt * L.njt, awp_tuning().str_ichunk, awp_tuning().str_target_wgs, &nic, &ichunk);
  L.ichunk = ichunk;
  L.nblocks = L.nkt * L.njt * nic;
#define AWP_STR_ARGS                                                                          \
  A[0], A[1], A[2], A[3], A[4], A[5], S0, A[6], A[7], A[8], A[9], A[10], A[11], A[12], A[13], \
      A[14], lam, mu, qp, coeff, qs, dcrjx, dcrjy, dcrjz, vx1, vx2, ww, wwo, L, SC, TT
#define AWP_STR_LAUNCH(B, KW)                                                               \
  hipLaunchKernelGGL((dstrqc_kernel<B, NEWLIT, OOP, ST, TEX, KW>), dim3(L.nblocks), dim3(64, B), \
                     0, St, AWP_STR_ARGS)
  if (kw == 32) {
    switch (bj) {
      case 1: AWP_STR_LAUNCH(1, 32); break;
      case 2: AWP_STR_LAUNCH(2, 32); break;
      case 4: AWP_STR_LAUNCH(4, 32); break;
      case 8: AWP_STR_LAUNCH(8, 32); break;
      default: AWP_STR_LAUNCH(16, 32); break;
    }
  } else {
    switch (bj) {
      case 1: AWP_STR_LAUNCH(1, 64); break;
      case 2: AWP_STR_LAUNCH(2, 64); break;
      case 4: AWP_STR_LAUNCH(4, 64); break;
      case 8: AWP_STR_LAUNCH(8, 64); break;
      default: AWP_STR_LAUNCH(16, 64); break;
    }
  }
#undef AWP_STR_LAUNCH
#undef AWP_STR_ARGS
  awp_check_launch(W);
}

// fstr (kernel.cu:1616-1647): one lane per (i, j) column of grid 0.
struct FstrLaunch {
  int i0, i1, j0, j1, ktop, yline;
  long long slice;
};

// Every store is a negated copy or a zero, so in 16-bit storage it acts on
// the stored bits exactly as on the decoded values.
template <class ST>
__global__ __launch_bounds__(256) void fstr_kernel(typename ST::T* __restrict__ zz,
                                                   typename ST::T* __restrict__ xz,
                                                   typename ST::T* __restrict__ yz,
                                                   const FstrLaunch F, const WaveScales SC) {
  const int j = F.j0 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int i = F.i0 + (int)(blockIdx.y * 4 + threadIdx.y);
  if (j > F.j1 || i > F.i1) return;
  const long long p = (long long)i * F.slice + (long long)j * F.yline + F.ktop;
  const float sz = SC.s[AWP_SZZ], sx = SC.s[AWP_SXZ], sy = SC.s[AWP_SYZ];
  const float iz = SC.inv[AWP_SZZ], ix = SC.inv[AWP_SXZ], iy = SC.inv[AWP_SYZ];
  const float zc = ST::ld(zz, p, sz), zm1 = ST::ld(zz, p - 1, sz);
  const float xm1 = ST::ld(xz, p - 1, sx), xm2 = ST::ld(xz, p - 2, sx);
  const float ym1 = ST::ld(yz, p - 1, sy), ym2 = ST::ld(yz, p - 2, sy);
  ST::st(zz, p + 1, -zc, iz);
  ST::st(zz, p + 2, -zm1, iz);
  ST::st(xz, p + 1, -xm1, ix);
  ST::st(xz, p + 2, -xm2, ix);
  ST::st(yz, p + 1, -ym1, iy);
  ST::st(yz, p + 2, -ym2, iy);
  ST::st(xz, p, 0.0f, ix);
  ST::st(yz, p, 0.0f, iy);
}

template <class ST>
void fstr_common(const char* W, void* zz, void* xz, void* yz, hipStream_t St, int s_i, int e_i,
                 int s_j, int e_j, const WaveScales& SC) {
  typedef typename ST::T T;
  if (!awp_grid_ok(0, W)) return;
  const AwpGrid& g = awp_state().grid[0];
  if (s_i > e_i || s_j > e_j) return;
  if (s_i < 0 || e_i >= g.nxt + 4 + AWP_NGSL2 || s_j < 0 || e_j >= g.nyt + 4 + AWP_NGSL2) {
    awp_set_error(AWP_ERR_ARG, "%s: range i[%d,%d] j[%d,%d] leaves the padded grid", W, s_i, e_i,
                  s_j, e_j);
    return;
  }
  FstrLaunch F;
  F.i0 = s_i; F.i1 = e_i; F.j0 = s_j; F.j1 = e_j;
  F.ktop = g.nzt + AWP_ALIGN - 1;
  F.yline = g.yline;
  F.slice = g.slice;
  dim3 grid(awp_cdiv(e_j - s_j + 1, 64), awp_cdiv(e_i - s_i + 1, 4));
  hipLaunchKernelGGL((fstr_kernel<ST>), grid, dim3(64, 4), 0, St, (T*)zz, (T*)xz, (T*)yz, F, SC);
  awp_check_launch(W);
}

}  // namespace

extern "C" {

void dstrqc_H(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz, float* r1,
              float* r2, float* r3, float* r4, float* r5, float* r6, float* u1, float* v1,
              float* w1, float* lam, float* mu, float* qp, float* coeff, float* qs,
              float* dcrjx, float* dcrjy, float* dcrjz, int nyt, int nzt, hipStream_t St,
              float* lam_mu, float* vx1, float* vx2, int* ww, float* wwo, int NX, int NPC,
              int rankx, int ranky, int s_i, int e_i, int s_j, int e_j, int d_i) {
  void* f15[15] = {xx, yy, zz, xy, xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1};
  WaveScales SC = {};
  dstrqc_common<false, false, StF32>("dstrqc_H", f15, CStress6{}, lam, mu, qp, coeff, qs, dcrjx,
                                     dcrjy, dcrjz, nyt, nzt, St, lam_mu, vx1, vx2, ww, wwo, NX,
                                     NPC, rankx, ranky, s_i, e_i, s_j, e_j, d_i, SC);
}

void dstrqc_H_new(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz,
                  float* r1, float* r2, float* r3, float* r4, float* r5, float* r6,
                  float* u1, float* v1, float* w1, float* lam, float* mu, float* qp,
                  float* coeff, float* qs, float* dcrjx, float* dcrjy, float* dcrjz, int nyt,
                  int nzt, hipStream_t St, float* lam_mu, float* vx1, float* vx2, int* ww,
                  float* wwo, int NX, int NPC, int rankx, int ranky, int s_i, int e_i,
                  int s_j, int e_j, int d_i) {
  void* f15[15] = {xx, yy, zz, xy, xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1};
  WaveScales SC = {};
  dstrqc_common<true, false, StF32>("dstrqc_H_new", f15, CStress6{}, lam, mu, qp, coeff, qs,
                                    dcrjx, dcrjy, dcrjz, nyt, nzt, St, lam_mu, vx1, vx2, ww, wwo,
                                    NX, NPC, rankx, ranky, s_i, e_i, s_j, e_j, d_i, SC);
}

void awp_dstrqc_s16(uint16_t* xx, uint16_t* yy, uint16_t* zz, uint16_t* xy, uint16_t* xz,
                    uint16_t* yz, uint16_t* r1, uint16_t* r2, uint16_t* r3, uint16_t* r4,
                    uint16_t* r5, uint16_t* r6, uint16_t* u1, uint16_t* v1, uint16_t* w1,
                    float* lam, float* mu, float* qp, float* coeff, float* qs, float* dcrjx,
                    float* dcrjy, float* dcrjz, int nyt, int nzt, hipStream_t St, float* lam_mu,
                    float* vx1, float* vx2, int* ww, float* wwo, int NX, int NPC, int rankx,
                    int ranky, int s_i, int e_i, int s_j, int e_j, int d_i, int fmt,
                    const float* scale) {
  const char* W = "awp_dstrqc_s16";
  WaveScales SC;
  if (!awp_scales_from_host(W, fmt, scale, &SC)) return;
  void* f15[15] = {xx, yy, zz, xy, xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1};
  if (fmt == AWP_FMT_BF16)
    dstrqc_common<false, false, StBF16>(W, f15, CStress6{}, lam, mu, qp, coeff, qs, dcrjx, dcrjy,
                                        dcrjz, nyt, nzt, St, lam_mu, vx1, vx2, ww, wwo, NX, NPC,
                                        rankx, ranky, s_i, e_i, s_j, e_j, d_i, SC);
  else
    dstrqc_common<false, false, StF16>(W, f15, CStress6{}, lam, mu, qp, coeff, qs, dcrjx, dcrjy,
                                       dcrjz, nyt, nzt, St, lam_mu, vx1, vx2, ww, wwo, NX, NPC,
                                       rankx, ranky, s_i, e_i, s_j, e_j, d_i, SC);
}

void awp_dstrqc_tex_s16(uint16_t* xx, uint16_t* yy, uint16_t* zz, uint16_t* xy, uint16_t* xz,
                        uint16_t* yz, uint16_t* r1, uint16_t* r2, uint16_t* r3, uint16_t* r4,
                        uint16_t* r5, uint16_t* r6, uint16_t* u1, uint16_t* v1, uint16_t* w1,
                        float* lam, float* mu, float* qp, float* coeff, float* qs, float* dcrjx,
                        float* dcrjy, float* dcrjz, int nyt, int nzt, hipStream_t St,
                        float* lam_mu, const float* tex, int* ww, int NX, int NPC, int rankx,
                        int ranky, int s_i, int e_i, int s_j, int e_j, int d_i, int fmt,
                        const float* scale) {
  const char* W = "awp_dstrqc_tex_s16";
  if (!tex) {
    awp_set_error(AWP_ERR_ARG, "%s: the texture table is NULL", W);
    return;
  }
  if (!awp_grid_ok(d_i, W)) return;
  WaveScales SC;
  if (!awp_scales_from_host(W, fmt, scale, &SC)) return;
  const AwpGrid& g = awp_state().grid[d_i];
  TexTable TT;
  for (int m = 0; m < 8; ++m) {
    TT.vx1[m] = tex[m];
    TT.vx2[m] = tex[8 + m];
    TT.wwo[m] = tex[16 + m];
  }
  // init_texture's region on one rank: the interior columns, all levels
  TT.xls = 2 + AWP_NGSL;
  TT.yls = 2 + AWP_NGSL;
  TT.nj = g.nyt;
  TT.nzt = g.nzt;
  void* f15[15] = {xx, yy, zz, xy, xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1};
#define AWP_TEX16(STT, TEXM, WWP)                                                             \
  dstrqc_common<false, false, STT, TEXM>(W, f15, CStress6{}, lam, mu, qp, coeff, qs, dcrjx,   \
                                         dcrjy, dcrjz, nyt, nzt, St, lam_mu, nullptr, nullptr, \
                                         WWP, nullptr, NX, NPC, rankx, ranky, s_i, e_i, s_j,   \
                                         e_j, d_i, SC, TT)
  if (fmt == AWP_FMT_BF16) {
    if (ww)
      AWP_TEX16(StBF16, TEX_TABLE, ww);
    else
      AWP_TEX16(StBF16, TEX_RULE, nullptr);
  } else {
    if (ww)
      AWP_TEX16(StF16, TEX_TABLE, ww);
    else
      AWP_TEX16(StF16, TEX_RULE, nullptr);
  }
#undef AWP_TEX16
}

void awp_dstrqc_oop(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz,
                    const float* const* s0, float* r1, float* r2, float* r3, float* r4,
                    float* r5, float* r6, float* u1, float* v1, float* w1, float* lam, float* mu,
                    float* qp, float* coeff, float* qs, float* dcrjx, float* dcrjy,
                    float* dcrjz, int nyt, int nzt, hipStream_t St, float* lam_mu, float* vx1,
                    float* vx2, int* ww, float* wwo, int NX, int NPC, int rankx, int ranky,
                    int s_i, int e_i, int s_j, int e_j, int d_i) {
  const char* W = "awp_dstrqc_oop";
  void* f15[15] = {xx, yy, zz, xy, xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1};
  CStress6 S0;
  for (int c = 0; c < 6; ++c) {
    S0.f[c] = s0 ? s0[c] : nullptr;
    if (!S0.f[c] || S0.f[c] == f15[c]) {
      awp_set_error(AWP_ERR_ARG, "%s: the stresses before and after the step must be distinct "
                    "arrays (dstrqc_H updates in place)", W);
      return;
    }
  }
  WaveScales SC = {};
  dstrqc_common<false, true, StF32>(W, f15, S0, lam, mu, qp, coeff, qs, dcrjx, dcrjy, dcrjz, nyt,
                                    nzt, St, lam_mu, vx1, vx2, ww, wwo, NX, NPC, rankx, ranky,
                                    s_i, e_i, s_j, e_j, d_i, SC);
}

void awp_dstrqc_tex(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz,
                    float* r1, float* r2, float* r3, float* r4, float* r5, float* r6, float* u1,
                    float* v1, float* w1, float* lam, float* mu, float* qp, float* coeff,
                    float* qs, float* dcrjx, float* dcrjy, float* dcrjz, int nyt, int nzt,
                    hipStream_t St, float* lam_mu, const float* tex, int* ww, int NX, int NPC,
                    int rankx, int ranky, int s_i, int e_i, int s_j, int e_j, int d_i) {
  const char* W = "awp_dstrqc_tex";
  if (!tex) {
    awp_set_error(AWP_ERR_ARG, "%s: the texture table is NULL", W);
    return;
  }
  if (!awp_grid_ok(d_i, W)) return;
  const AwpGrid& g = awp_state().grid[d_i];
  TexTable TT;
  for (int m = 0; m < 8; ++m) {
    TT.vx1[m] = tex[m];
    TT.vx2[m] = tex[8 + m];
    TT.wwo[m] = tex[16 + m];
  }
  // init_texture's region on one rank: the interior columns, all levels
  TT.xls = 2 + AWP_NGSL;
  TT.yls = 2 + AWP_NGSL;
  TT.nj = g.nyt;
  TT.nzt = g.nzt;
  void* f15[15] = {xx, yy, zz, xy, xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1};
  WaveScales SC = {};
  if (ww)
    dstrqc_common<false, false, StF32, TEX_TABLE>(W, f15, CStress6{}, lam, mu, qp, coeff, qs,
                                                  dcrjx, dcrjy, dcrjz, nyt, nzt, St, lam_mu,
                                                  nullptr, nullptr, ww, nullptr, NX, NPC, rankx,
                                                  ranky, s_i, e_i, s_j, e_j, d_i, SC, TT);
  else
    dstrqc_common<false, false, StF32, TEX_RULE>(W, f15, CStress6{}, lam, mu, qp, coeff, qs,
                                                 dcrjx, dcrjy, dcrjz, nyt, nzt, St, lam_mu,
                                                 nullptr, nullptr, nullptr, nullptr, NX, NPC,
                                                 rankx, ranky, s_i, e_i, s_j, e_j, d_i, SC, TT);
}

void fstr_H(float* zz, float* xz, float* yz, hipStream_t St, int s_i, int e_i, int s_j,
            int e_j) {
  WaveScales SC = {};
  fstr_common<StF32>("fstr_H", zz, xz, yz, St, s_i, e_i, s_j, e_j, SC);
}

void awp_fstr_s16(uint16_t* zz, uint16_t* xz, uint16_t* yz, hipStream_t St, int s_i, int e_i,
                  int s_j, int e_j, int fmt, const float* scale) {
  const char* W = "awp_fstr_s16";
  WaveScales SC;
  if (!awp_scales_from_host(W, fmt, scale, &SC)) return;
  if (fmt == AWP_FMT_BF16)
    fstr_common<StBF16>(W, zz, xz, yz, St, s_i, e_i, s_j, e_j, SC);
  else
    fstr_common<StF16>(W, zz, xz, yz, St, s_i, e_i, s_j, e_j, SC);
}

}  // extern "C"
