// Discrete adjoint of one grid's NVE = 1 time step and the material gradient.
//
// The forward step (velocity, free-surface ghosts, stress + memory variables,
// source, stress imaging) is linear in the wavefield for a fixed medium, so
// its adjoint is the transpose, applied in reverse order:
//   fstr^T -> stress^T (local part, then the strain -> velocity gather)
//   -> ghost^T -> velocity^T (local part, then the velocity -> stress gather).
// Every transposed stencil is written as a gather, so each output cell is
// owned by one lane and no atomics are needed.  Adjoint fields are binary32.
//
// Gradient: the step is also linear in per-cell coefficients of the medium --
// f_d1..3 = dth / (4-point density means) of the velocity update, and
// xl = 8/sum8(lam) + xm, xm = 16/sum8(mu), xmu1..3 = 2/(2-point mu sums) of
// the stress update, and lam_mu of the ghosts.  Their gradients are summed
// over the steps (binary32 inside a checkpoint segment, added into binary64
// at its end) and mapped to the arrays lam (1/lambda), mu (1/mu), d1 (rho)
// once, in binary64.
#include <mutex>

#include "awp_common.h"
#include "awp_tuning.h"
#include "../../include/awp_hip.h"

namespace {

struct AdjGrid {
  int nx, ny, nz, top, yline, lam_mu_ld;
  int jgroup;  // block order: 0 plane by plane, else groups of jgroup j tiles over all i
  long long slice;
  float c1, c2, dth, dh1, DT;
};

// Block (x, y, z) of a (k tile, j tile, i plane) grid after the XCD-aware
// renumbering (awp_xcd_remap): consecutive logical blocks share one XCD's
// L2, so the j and i neighbours the gathers read come from L2.  Logical
// order: with G.jgroup = 0 the k and j tiles of one plane, then of the next;
// with G.jgroup = JG > 0 the k tiles and JG j tiles of one plane, then of
// the next plane, through all planes, then the next JG j tiles (the last
// group takes the remaining tiles), so the blocks an XCD holds at once span
// a few planes of a narrow j band.
struct Blk3 {
  int x, y, z;
};
__device__ __forceinline__ Blk3 adj_block(const AdjGrid& G) {
  const int nx = (int)gridDim.x, ny = (int)gridDim.y, nz = (int)gridDim.z;
  const int nb = nx * ny * nz;
  int b = awp_xcd_remap((int)(blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z)), nb);
  if (G.jgroup <= 0 || G.jgroup >= ny) return Blk3{b % nx, (b / nx) % ny, b / (nx * ny)};
  const int x = b % nx;
  b /= nx;  // (j group, i, j tile in group), j tile fastest
  const int JG = G.jgroup, full = ny / JG, per = JG * nz;
  int y0, jg;
  if (b < full * per) {
    y0 = (b / per) * JG;
    jg = JG;
    b %= per;
  } else {
    y0 = full * JG;
    jg = ny - y0;
    b -= full * per;
  }
  return Blk3{x, y0 + b % jg, b / jg};
}

__device__ __forceinline__ long long aidx(const AdjGrid& G, int i, int j, int k) {
  return (long long)i * G.slice + (long long)j * G.yline + k;
}

// Cell of a launch over the free-surface band, levels [k0, k1] of the
// interior columns: the band is a few levels deep, so the 256 lanes of a
// block take nk = k1-k0+1 consecutive levels of 256/nk consecutive columns
// (k fastest, then j) of one i plane, instead of one column per wave with
// nk of its 64 lanes active.  Grid: (cdiv(nk * ny, 256), 1, nx).
struct BandCell {
  int i, j, k;
  bool ok;
};
__device__ __forceinline__ BandCell band_cell(const AdjGrid& G, int k0, int k1) {
  const Blk3 B = adj_block(G);
  const int nk = k1 - k0 + 1;
  const int t = B.x * 256 + (int)(threadIdx.y * 64 + threadIdx.x);
  BandCell c;
  c.j = 2 + AWP_NGSL + t / nk;
  c.k = k0 + t % nk;
  c.i = 2 + AWP_NGSL + B.z;
  c.ok = c.j < G.ny + 2 + AWP_NGSL;
  return c;
}

// Production Q weight (as the forward dstrqc).
__device__ __forceinline__ float adj_qw(float q, float ca, float cb, float wwo) {
  float w;
  if (1. / ((double)q * 2.0) <= 200.0)
    w = (float)((double)ca * (2. * (double)q) * (2. * (double)q) + (double)cb * (2. * (double)q));
  else
    w = 2.0f * wwo * q;
  return w / wwo;
}

// fstr^T: the imaging copies move adjoint values from above the surface back
// to the levels they were copied from; the zeroed and overwritten cells lose
// their adjoint.
__global__ __launch_bounds__(256) void adj_fstr_kernel(float* __restrict__ zz,
                                                       float* __restrict__ xz,
                                                       float* __restrict__ yz, const AdjGrid G) {
  const int j = 2 + AWP_NGSL + (int)(blockIdx.x * 64 + threadIdx.x);
  const int i = 2 + AWP_NGSL + (int)(blockIdx.y * 4 + threadIdx.y);
  if (j >= G.ny + 2 + AWP_NGSL || i >= G.nx + 2 + AWP_NGSL) return;
  const long long p = aidx(G, i, j, G.top);
  zz[p] -= zz[p + 1];
  zz[p - 1] -= zz[p + 2];
  zz[p + 1] = 0.0f;
  zz[p + 2] = 0.0f;
  xz[p - 1] -= xz[p + 1];
  xz[p - 2] -= xz[p + 2];
  xz[p] = 0.0f;
  xz[p + 1] = 0.0f;
  xz[p + 2] = 0.0f;
  yz[p - 1] -= yz[p + 1];
  yz[p - 2] -= yz[p + 2];
  yz[p] = 0.0f;
  yz[p + 1] = 0.0f;
  yz[p + 2] = 0.0f;
}

struct Adj15 {
  float* f[15];  // xx yy zz xy xz yz r1..r6 u1 v1 w1 (adjoint)
};
struct Ptr6 {
  float* f[6];
};
struct CPtr3 {
  const float* f[3];
};
struct CPtr6 {
  const float* f[6];
};
struct Grad8 {
  float* f[8];  // d1 d2 d3 xl xm xmu1 xmu2 xmu3
};
struct Medium {
  const float *d1, *lam, *mu, *qp, *qs, *vx1, *vx2, *wwo, *lam_mu, *coeff, *dcrjx, *dcrjy, *dcrjz;
  const int* ww;
};

// The velocity^T gather at stress cell q: the six stress adjoints'
// increments from the weighted velocity adjoints Aw (u, v, w).
__device__ __forceinline__ void vel_gather6(const CPtr6& Aw, long long q, const AdjGrid& G,
                                            float* g) {
  const long long S = G.slice, Y = G.yline;
  const float c1 = G.c1, c2 = G.c2;
  const float *Au = Aw.f[0], *Av = Aw.f[1], *Aw_ = Aw.f[2];
  g[0] = c1 * (Au[q] - Au[q + S]) + c2 * (Au[q - S] - Au[q + 2 * S]);
  g[1] = c1 * (Av[q - Y] - Av[q]) + c2 * (Av[q - 2 * Y] - Av[q + Y]);
  g[2] = c1 * (Aw_[q - 1] - Aw_[q]) + c2 * (Aw_[q - 2] - Aw_[q + 1]);
  g[3] = c1 * (Au[q] - Au[q + Y]) + c2 * (Au[q - Y] - Au[q + 2 * Y]) +
         c1 * (Av[q - S] - Av[q]) + c2 * (Av[q - 2 * S] - Av[q + S]);
  g[4] = c1 * (Au[q] - Au[q + 1]) + c2 * (Au[q - 1] - Au[q + 2]) +
         c1 * (Aw_[q - S] - Aw_[q]) + c2 * (Aw_[q - 2 * S] - Aw_[q + S]);
  g[5] = c1 * (Av[q] - Av[q + 1]) + c2 * (Av[q - 1] - Av[q + 2]) +
         c1 * (Aw_[q] - Aw_[q + Y]) + c2 * (Aw_[q - Y] - Aw_[q + 2 * Y]);
}

// stress^T, local part.  Per stress cell: the adjoint of the old stress and
// memory variables, the adjoints of the six strain combinations (N1..N3 for
// the normal strains, E4..E6 for the shear pairs) and the coefficient
// gradients (strains of the forward velocity after the step).
// Below the free-surface band (k <= top-3) the previous step's pending
// velocity^T gather (weights Awp) is added to the stress adjoints first.
__global__ __launch_bounds__(256) void adj_stress_local_kernel(Adj15 A, Ptr6 Es, CPtr3 V,
                                                               Medium M, Grad8 Gr,
                                                               const AdjGrid G, CPtr6 Awp) {
  // the Q coefficients by mechanism, from LDS (the lookup by ww would
  // otherwise be a second dependent round trip to memory)
  __shared__ float s_coeff[16];
  if (threadIdx.y == 0 && threadIdx.x < 16) s_coeff[threadIdx.x] = M.coeff[threadIdx.x];
  __syncthreads();
  const Blk3 B = adj_block(G);
  const int k = AWP_ALIGN + 3 + (int)(B.x * 64 + threadIdx.x);
  const int j = 2 + AWP_NGSL + (int)(B.y * 4 + threadIdx.y);
  const int i = 2 + AWP_NGSL + B.z;
  if (k > G.top || j >= G.ny + 2 + AWP_NGSL) return;
  const long long S = G.slice, Y = G.yline;
  const long long p = aidx(G, i, j, k);
  const long long ip1 = p + S, jm1 = p - Y, ijk = p + S - Y, km1 = p - 1, ik1 = p + S - 1,
                  jk1 = p - Y - 1, ijk1 = p + S - Y - 1;
  const bool has_pend = k <= G.top - 3;

  // Every load whose address is known up front, issued together before any
  // arithmetic (the scheduling barrier keeps the compiler from sinking them
  // to their uses), so a wave waits for one round trip instead of three.
  const float f_vx1 = M.vx1[p], f_vx2 = M.vx2[p], f_wwo = M.wwo[p];
  const int f_ww = M.ww[p];
  const float l0 = M.lam[p], l1 = M.lam[ip1], l2 = M.lam[jm1], l3 = M.lam[ijk], l4 = M.lam[km1],
              l5 = M.lam[ik1], l6 = M.lam[jk1], l7 = M.lam[ijk1];
  const float m0 = M.mu[p], m1 = M.mu[ip1], m2 = M.mu[jm1], m3 = M.mu[ijk], m4 = M.mu[km1],
              m5 = M.mu[ik1], m6 = M.mu[jk1], m7 = M.mu[ijk1];
  const float p0 = M.qp[p], p1 = M.qp[ip1], p2 = M.qp[jm1], p3 = M.qp[ijk], p4 = M.qp[km1],
              p5 = M.qp[ik1], p6 = M.qp[jk1], p7 = M.qp[ijk1];
  const float q0 = M.qs[p], q1 = M.qs[ip1], q2 = M.qs[jm1], q3 = M.qs[ijk], q4 = M.qs[km1],
              q5 = M.qs[ik1], q6 = M.qs[jk1], q7 = M.qs[ijk1];
  const float dcx = M.dcrjx[i], dcy = M.dcrjy[j], dcz = M.dcrjz[k];
  const float* u = V.f[0];
  const float* v = V.f[1];
  const float* w = V.f[2];
  const float u_0 = u[p], u_is = u[p + S], u_i2 = u[p + 2 * S], u_im = u[p - S];
  const float u_jp = u[p + Y], u_j2 = u[p + 2 * Y], u_jm = u[p - Y];
  const float u_kp = u[p + 1], u_k2 = u[p + 2], u_km = u[p - 1];
  const float v_0 = v[p], v_jm = v[p - Y], v_jp = v[p + Y], v_j2m = v[p - 2 * Y];
  const float v_im = v[p - S], v_ip = v[p + S], v_i2m = v[p - 2 * S];
  const float v_kp = v[p + 1], v_k2 = v[p + 2], v_km = v[p - 1];
  const float w_0 = w[p], w_km = w[p - 1], w_kp = w[p + 1], w_k2m = w[p - 2];
  const float w_im = w[p - S], w_ip = w[p + S], w_i2m = w[p - 2 * S];
  const float w_jp = w[p + Y], w_j2 = w[p + 2 * Y], w_jm = w[p - Y];
  float sbl[6], rbl[6], pend[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    sbl[c] = A.f[c][p];
    rbl[c] = A.f[6 + c][p];
  }
  // the pending velocity^T gather (vel_gather6's operands)
  float au0 = 0.f, au_is = 0.f, au_im = 0.f, au_i2 = 0.f, au_jp = 0.f, au_jm = 0.f,
        au_j2 = 0.f, au_kp = 0.f, au_km = 0.f, au_k2 = 0.f;
  float av0 = 0.f, av_jm = 0.f, av_j2m = 0.f, av_jp = 0.f, av_im = 0.f, av_i2m = 0.f,
        av_ip = 0.f, av_kp = 0.f, av_km = 0.f, av_k2 = 0.f;
  float aw0 = 0.f, aw_km = 0.f, aw_k2m = 0.f, aw_kp = 0.f, aw_im = 0.f, aw_i2m = 0.f,
        aw_ip = 0.f, aw_jp = 0.f, aw_jm = 0.f, aw_j2 = 0.f;
  if (has_pend) {
    const float *Au = Awp.f[0], *Av = Awp.f[1], *Aw_ = Awp.f[2];
    au0 = Au[p]; au_is = Au[p + S]; au_im = Au[p - S]; au_i2 = Au[p + 2 * S];
    au_jp = Au[p + Y]; au_jm = Au[p - Y]; au_j2 = Au[p + 2 * Y];
    au_kp = Au[p + 1]; au_km = Au[p - 1]; au_k2 = Au[p + 2];
    av0 = Av[p]; av_jm = Av[p - Y]; av_j2m = Av[p - 2 * Y]; av_jp = Av[p + Y];
    av_im = Av[p - S]; av_i2m = Av[p - 2 * S]; av_ip = Av[p + S];
    av_kp = Av[p + 1]; av_km = Av[p - 1]; av_k2 = Av[p + 2];
    aw0 = Aw_[p]; aw_km = Aw_[p - 1]; aw_k2m = Aw_[p - 2]; aw_kp = Aw_[p + 1];
    aw_im = Aw_[p - S]; aw_i2m = Aw_[p - 2 * S]; aw_ip = Aw_[p + S];
    aw_jp = Aw_[p + Y]; aw_jm = Aw_[p - Y]; aw_j2 = Aw_[p + 2 * Y];
  }
  float gold[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) gold[c] = Gr.f[3 + c][p];
  __builtin_amdgcn_sched_barrier(0);

  const float c1 = G.c1, c2 = G.c2;
  if (has_pend) {
    // vel_gather6's sums, term for term
    pend[0] = c1 * (au0 - au_is) + c2 * (au_im - au_i2);
    pend[1] = c1 * (av_jm - av0) + c2 * (av_j2m - av_jp);
    pend[2] = c1 * (aw_km - aw0) + c2 * (aw_k2m - aw_kp);
    pend[3] = c1 * (au0 - au_jp) + c2 * (au_jm - au_j2) + c1 * (av_im - av0) +
              c2 * (av_i2m - av_ip);
    pend[4] = c1 * (au0 - au_kp) + c2 * (au_km - au_k2) + c1 * (aw_im - aw0) +
              c2 * (aw_i2m - aw_ip);
    pend[5] = c1 * (av0 - av_kp) + c2 * (av_km - av_k2) + c1 * (aw0 - aw_jp) +
              c2 * (aw_jm - aw_j2);
  }
  const float ca = s_coeff[(f_ww * 2 - 2) & 15], cb = s_coeff[(f_ww * 2 - 1) & 15];
  const float dcrj = dcx * dcy * dcz;
  const float sl8 = l0 + l1 + l2 + l3 + l4 + l5 + l6 + l7;
  const float sm8 = m0 + m1 + m2 + m3 + m4 + m5 + m6 + m7;
  const float sp8 = p0 + p1 + p2 + p3 + p4 + p5 + p6 + p7;
  const float ss8 = q0 + q1 + q2 + q3 + q4 + q5 + q6 + q7;
  const float xm = awp_quot(16.0f, sm8);
  const float xl = awp_quot(8.0f, sl8) + xm;
  const float xmu1 = awp_quot(2.0f, m0 + m4);
  const float xmu2 = awp_quot(2.0f, m0 + m2);
  const float xmu3 = awp_quot(2.0f, m0 + m1);
  const float qpaw = adj_qw(0.0625f * sp8, ca, cb, f_wwo);
  const float hw = adj_qw(0.0625f * ss8, ca, cb, f_wwo);
  const float h1w = adj_qw(0.250f * (q0 + q4), ca, cb, f_wwo);
  const float h2w = adj_qw(0.250f * (q0 + q2), ca, cb, f_wwo);
  const float h3w = adj_qw(0.250f * (q0 + q1), ca, cb, f_wwo);
  const float dth = G.dth, dh1 = G.dh1, DT = G.DT;
  // coefficients per unit of xl, xm, xmu_k (the step is linear in them)
  const float qf = -qpaw * dh1 * f_vx1;            // qpa = xl * qf
  const float hf = -hw * dh1 * f_vx1;              // h = xm * hf
  const float h1f = -h1w * dh1 * f_vx1, h2f = -h2w * dh1 * f_vx1, h3f = -h3w * dh1 * f_vx1;
  const float xf = dth + DT * hf;                  // XM = xm * xf
  const float x1f = dth + DT * h1f, x2f = dth + DT * h2f, x3f = dth + DT * h3f;
  const float XL = xl * dth, QPA = xl * qf, XM = xm * xf, H = xm * hf;
  const float XMU[3] = {xmu1 * x1f, xmu2 * x2f, xmu3 * x3f};
  const float HS[3] = {xmu1 * h1f, xmu2 * h2f, xmu3 * h3f};
  const float HSF[3] = {h1f, h2f, h3f};
  const float XSF[3] = {x1f, x2f, x3f};
  const float vx1p = DT * (1 + f_vx2 * f_vx1);
  const float keep = f_vx2 * (1 - f_vx1);

  // forward strains after the step (for the coefficient gradients)
  const float N1 = c1 * (u_is - u_0) + c2 * (u_i2 - u_im);
  const float N2 = c1 * (v_0 - v_jm) + c2 * (v_jp - v_j2m);
  const float N3 = c1 * (w_0 - w_km) + c2 * (w_kp - w_k2m);
  const float Esh[3] = {
      c1 * (u_jp - u_0) + c2 * (u_j2 - u_jm) + c1 * (v_0 - v_im) + c2 * (v_ip - v_i2m),
      c1 * (u_kp - u_0) + c2 * (u_k2 - u_km) + c1 * (w_0 - w_im) + c2 * (w_ip - w_i2m),
      c1 * (v_kp - v_0) + c2 * (v_k2 - v_km) + c1 * (w_jp - w_0) + c2 * (w_j2 - w_jm)};
  const float Ssum = N1 + N2 + N3;
  const float O[3] = {N2 + N3, N1 + N3, N1 + N2};

  // normal components: sigma' = P (sigma + (XL + DT QPA) S - XM O + vx1p r)
  //                              + D DT ((wwo-1) frt + keep r),  P = D (xx), D^2 (yy, zz)
  //                     r' = vx2 r + wwo frt,  frt = -H O + QPA S
  float Sbar = 0.0f, Obar[3], gxl = 0.0f, gxm = 0.0f;
  for (int c = 0; c < 3; ++c) {
    const float P = (c == 0) ? dcrj : dcrj * dcrj;
    const float sb = has_pend ? sbl[c] + pend[c] : sbl[c];
    const float rb = rbl[c];
    const float frtb = dcrj * DT * (f_wwo - 1) * sb + f_wwo * rb;
    A.f[c][p] = P * sb;
    A.f[6 + c][p] = P * vx1p * sb + dcrj * DT * keep * sb + f_vx2 * rb;
    Sbar += P * (XL + DT * QPA) * sb + QPA * frtb;
    Obar[c] = -P * XM * sb - H * frtb;
    gxl += sb * P * (dth + DT * qf) * Ssum + frtb * qf * Ssum;
    gxm += -sb * P * xf * O[c] - frtb * hf * O[c];
  }
  float gmu[3];
  for (int c = 0; c < 3; ++c) {
    const float sb = has_pend ? sbl[3 + c] + pend[3 + c] : sbl[3 + c];
    const float rb = rbl[3 + c];
    const float frtb = dcrj * DT * (f_wwo - 1) * sb + f_wwo * rb;
    A.f[3 + c][p] = dcrj * sb;
    A.f[9 + c][p] = dcrj * vx1p * sb + dcrj * DT * keep * sb + f_vx2 * rb;
    Es.f[3 + c][p] = dcrj * XMU[c] * sb + HS[c] * frtb;
    gmu[c] = sb * dcrj * XSF[c] * Esh[c] + frtb * HSF[c] * Esh[c];
  }
  Es.f[0][p] = Sbar + Obar[1] + Obar[2];
  Es.f[1][p] = Sbar + Obar[0] + Obar[2];
  Es.f[2][p] = Sbar + Obar[0] + Obar[1];
  Gr.f[3][p] = gold[0] + gxl;
  Gr.f[4][p] = gold[1] + gxm;
  Gr.f[5][p] = gold[2] + gmu[0];
  Gr.f[6][p] = gold[3] + gmu[1];
  Gr.f[7][p] = gold[4] + gmu[2];
}

// stress^T, gather part: velocity adjoints from the strain adjoints of the
// stress cells whose stencils read them (interior columns, k up to top+1).
// The strain -> velocity gather at cell c (the velocity adjoints' increments).
struct Vec3 {
  float u, v, w;
};
__device__ __forceinline__ Vec3 strain_gather(const CPtr6& E, long long c, long long S,
                                              long long Y, float c1, float c2) {
  const float *N1 = E.f[0], *N2 = E.f[1], *N3 = E.f[2], *E4 = E.f[3], *E5 = E.f[4],
              *E6 = E.f[5];
  Vec3 r;
  r.u = c1 * (N1[c - S] - N1[c]) + c2 * (N1[c - 2 * S] - N1[c + S]) +
        c1 * (E4[c - Y] - E4[c]) + c2 * (E4[c - 2 * Y] - E4[c + Y]) +
        c1 * (E5[c - 1] - E5[c]) + c2 * (E5[c - 2] - E5[c + 1]);
  r.v = c1 * (N2[c] - N2[c + Y]) + c2 * (N2[c - Y] - N2[c + 2 * Y]) +
        c1 * (E4[c] - E4[c + S]) + c2 * (E4[c - S] - E4[c + 2 * S]) +
        c1 * (E6[c - 1] - E6[c]) + c2 * (E6[c - 2] - E6[c + 1]);
  r.w = c1 * (N3[c] - N3[c + 1]) + c2 * (N3[c - 1] - N3[c + 2]) +
        c1 * (E5[c] - E5[c + S]) + c2 * (E5[c - S] - E5[c + 2 * S]) +
        c1 * (E6[c - Y] - E6[c]) + c2 * (E6[c - 2 * Y] - E6[c + Y]);
  return r;
}

// ghost^T, part 1 (column (i, j)): the adjoints of ug, vg and of the bracket
// of the w ghost, and the lam_mu gradient, from the ghost-level adjoints
// ubg, vbg, wbg (ub, vb, wb at top+1).
__device__ __forceinline__ void ghost1_column(int i, int j, float ubg, float vbg, float wbg,
                                              const CPtr3& V,
                                              const float* __restrict__ lam_mu,
                                              float* __restrict__ col,
                                              float* __restrict__ glm, const AdjGrid& G) {
  const long long S = G.slice, Y = G.yline;
  const long long p = aidx(G, i, j, G.top);
  const long long q2 = (long long)i * G.lam_mu_ld + j;
  const float* u = V.f[0];
  const float* v = V.f[1];
  const float* w = V.f[2];
  const bool cx = (i - AWP_NGSL - 1) < G.nx, cy = (j - AWP_NGSL - 1) > 1;
  const float ug = u[p] - (w[p] - w[p - S]);
  const float vg = v[p] - (w[p + Y] - w[p]);
  const float vs1 = cx ? u[p + S] - (w[p + S] - w[p]) : 0.0f;
  const float vs2 = cy ? v[p - Y] - (w[p] - w[p - Y]) : 0.0f;
  const float Bk = (vs1 - ug) + (u[p + S] - u[p]) + (vg - vs2) + (v[p] - v[p - Y]);
  const float wgb = wbg;
  const float s = -lam_mu[q2] * wgb;
  glm[q2] += -wgb * Bk;
  const long long n2 = (long long)(G.nx + 4 + AWP_NGSL2) * G.lam_mu_ld;
  col[q2] = ubg - s;       // adjoint of ug
  col[n2 + q2] = vbg + s;  // adjoint of vg
  col[2 * n2 + q2] = s;    // adjoint of the bracket
}

// ghost^T, part 2 (column (i, j)): gather the ghost adjoints onto the cells
// the ghosts were computed from, then clear the ghost adjoints.  The w
// ghost's own term, wb[top-1] += wb[top+1], is left to the caller (the lane
// of level top-1).
__device__ __forceinline__ void ghost2_column(int i, int j, float* __restrict__ ub,
                                              float* __restrict__ vb, float* __restrict__ wb,
                                              const float* __restrict__ col,
                                              const AdjGrid& G) {
  const long long p = aidx(G, i, j, G.top);
  const int L = G.lam_mu_ld;
  const long long n2 = (long long)(G.nx + 4 + AWP_NGSL2) * L;
  const int i0 = 2 + AWP_NGSL, i1 = G.nx + 1 + AWP_NGSL, j0 = 2 + AWP_NGSL, j1 = G.ny + 1 + AWP_NGSL;
  auto AU = [&](int ii, int jj) { return col[(long long)ii * L + jj]; };
  auto AV = [&](int ii, int jj) { return col[n2 + (long long)ii * L + jj]; };
  auto SB = [&](int ii, int jj) { return col[2 * n2 + (long long)ii * L + jj]; };
  auto CX = [&](int ii) { return (ii - AWP_NGSL - 1) < G.nx; };
  auto CY = [&](int jj) { return (jj - AWP_NGSL - 1) > 1; };
  const float s = SB(i, j);
  float du = AU(i, j) - s;
  float dv = AV(i, j) + s;
  float dw = -AU(i, j) + AV(i, j) + (CX(i) ? s : 0.0f) + (CY(j) ? s : 0.0f);
  if (i - 1 >= i0) {
    const float sl = SB(i - 1, j);
    du += sl + (CX(i - 1) ? sl : 0.0f);   // u[top+x] of the left column
    dw += CX(i - 1) ? -sl : 0.0f;         // w[top+x] in its vs1
  }
  if (j + 1 <= j1) {
    const float sf = SB(i, j + 1);
    dv += -sf + (CY(j + 1) ? -sf : 0.0f);  // v[top-y] of the next column
    dw += CY(j + 1) ? -sf : 0.0f;          // w[top-y] in its vs2
    dw += 0.0f;
  }
  if (i + 1 <= i1) dw += AU(i + 1, j);     // w[top-x] of the right column's ug
  if (j - 1 >= j0) dw += -AV(i, j - 1);    // w[top+y] of the previous column's vg
  ub[p] += du;
  vb[p] += dv;
  wb[p] += dw;
  ub[p + 1] = 0.0f;
  vb[p + 1] = 0.0f;
  wb[p + 1] = 0.0f;
}

// velocity^T, local part at cell (i, j, k): A = dcrj * f_d * adjoint (the
// weights the stress adjoints gather), the velocity adjoint scaled by dcrj,
// and the f_d gradients with the stress differences of the state before the
// step.  GATHER: the strain -> velocity gather of the same cell first (levels
// that the free-surface ghosts do not touch), so the velocity adjoints make
// one trip through HBM instead of two.
template <bool GATHER>
__device__ __forceinline__ void vel_local_cell(int i, int j, int k, float* __restrict__ ub,
                                               float* __restrict__ vb, float* __restrict__ wb,
                                               const Ptr6& Aw, const CPtr6& Sg, const Medium& M,
                                               const Grad8& Gr, const AdjGrid& G,
                                               const CPtr6& E) {
  const long long S = G.slice, Y = G.yline;
  const long long p = aidx(G, i, j, k);
  const float* d1 = M.d1;
  const float dcrj = M.dcrjx[i] * M.dcrjy[j] * M.dcrjz[k];
  const float f1 = G.dth / (float)(0.25 * (double)(d1[p] + d1[p - Y] + d1[p - 1] + d1[p - Y - 1]));
  const float f2 = G.dth / (float)(0.25 * (double)(d1[p] + d1[p + S] + d1[p - 1] + d1[p + S - 1]));
  const float f3 = G.dth / (float)(0.25 * (double)(d1[p] + d1[p + S] + d1[p - Y] + d1[p + S - Y]));
  const float c1 = G.c1, c2 = G.c2;
  const float *xx = Sg.f[0], *yy = Sg.f[1], *zz = Sg.f[2], *xy = Sg.f[3], *xz = Sg.f[4],
              *yz = Sg.f[5];
  const float Xu = c1 * (xx[p] - xx[p - S]) + c2 * (xx[p + S] - xx[p - 2 * S]) +
                   c1 * (xy[p] - xy[p - Y]) + c2 * (xy[p + Y] - xy[p - 2 * Y]) +
                   c1 * (xz[p] - xz[p - 1]) + c2 * (xz[p + 1] - xz[p - 2]);
  const float Xv = c1 * (xy[p + S] - xy[p]) + c2 * (xy[p + 2 * S] - xy[p - S]) +
                   c1 * (yy[p + Y] - yy[p]) + c2 * (yy[p + 2 * Y] - yy[p - Y]) +
                   c1 * (yz[p] - yz[p - 1]) + c2 * (yz[p + 1] - yz[p - 2]);
  const float Xw = c1 * (xz[p + S] - xz[p]) + c2 * (xz[p + 2 * S] - xz[p - S]) +
                   c1 * (yz[p] - yz[p - Y]) + c2 * (yz[p + Y] - yz[p - 2 * Y]) +
                   c1 * (zz[p + 1] - zz[p]) + c2 * (zz[p + 2] - zz[p - 1]);
  float a = ub[p], b = vb[p], c = wb[p];
  if (GATHER) {
    const Vec3 g = strain_gather(E, p, S, Y, c1, c2);
    a += g.u;
    b += g.v;
    c += g.w;
  }
  Gr.f[0][p] += a * dcrj * Xu;
  Gr.f[1][p] += b * dcrj * Xv;
  Gr.f[2][p] += c * dcrj * Xw;
  Aw.f[0][p] = dcrj * f1 * a;
  Aw.f[1][p] = dcrj * f2 * b;
  Aw.f[2][p] = dcrj * f3 * c;
  ub[p] = dcrj * a;
  vb[p] = dcrj * b;
  wb[p] = dcrj * c;
}

// The strain gather and velocity^T below the free-surface band, levels
// [k0, k1], one column's levels per wave; asked for 8 waves per SIMD (63
// VGPRs from 72, no spills).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
void adj_vel_local_kernel(float* __restrict__ ub,
                                                            float* __restrict__ vb,
                                                            float* __restrict__ wb, Ptr6 Aw,
                                                            CPtr6 Sg, Medium M, Grad8 Gr,
                                                            const AdjGrid G, CPtr6 E, int k0,
                                                            int k1) {
  const Blk3 B = adj_block(G);
  const int k = k0 + (int)(B.x * 64 + threadIdx.x);
  const int j = 2 + AWP_NGSL + (int)(B.y * 4 + threadIdx.y);
  const int i = 2 + AWP_NGSL + B.z;
  if (k > k1 || j >= G.ny + 2 + AWP_NGSL) return;
  vel_local_cell<true>(i, j, k, ub, vb, wb, Aw, Sg, M, Gr, G, E);
}

// The free-surface band, first part (levels top-1 .. top+1, band_cell's lane
// map): the strain -> velocity gather of each cell, then, on the ghost level
// top+1, ghost^T part 1 of the cell's column from the values just gathered.
// Part 1 reads only its own column's ghost level, so it needs no barrier
// after the gathers of the other columns.
__global__ __launch_bounds__(256) void adj_band_gather_kernel(
    float* __restrict__ ub, float* __restrict__ vb, float* __restrict__ wb, CPtr6 E, CPtr3 V,
    const float* __restrict__ lam_mu, float* __restrict__ col, float* __restrict__ glm,
    const AdjGrid G) {
  const BandCell b = band_cell(G, G.top - 1, G.top + 1);
  if (!b.ok) return;
  const long long S = G.slice, Y = G.yline;
  const long long c = aidx(G, b.i, b.j, b.k);
  const Vec3 g = strain_gather(E, c, S, Y, G.c1, G.c2);
  const float un = ub[c] + g.u, vn = vb[c] + g.v, wn = wb[c] + g.w;
  ub[c] = un;
  vb[c] = vn;
  wb[c] = wn;
  if (b.k == G.top + 1) ghost1_column(b.i, b.j, un, vn, wn, V, lam_mu, col, glm, G);
}

// The free-surface band, second part (levels top-1 and top, band_cell's lane
// map: the two lanes of a column are neighbours in one wave).  The lane of
// level top runs ghost^T part 2 of the column (its ghost adjoints gathered
// from the neighbours' part 1) and hands the w ghost's adjoint wb[top+1],
// read before part 2 clears it, to the lane of level top-1, which adds it to
// its own cell; then each lane runs velocity^T on its cell, which reads only
// that cell's adjoints.
__global__ __launch_bounds__(256) void adj_band_ghost_kernel(
    float* __restrict__ ub, float* __restrict__ vb, float* __restrict__ wb,
    const float* __restrict__ col, Ptr6 Aw, CPtr6 Sg, Medium M, Grad8 Gr, const AdjGrid G) {
  const BandCell b = band_cell(G, G.top - 1, G.top);
  if (!b.ok) return;  // per column: both lanes of a column return or neither
  const bool at_top = b.k == G.top;
  const long long p = aidx(G, b.i, b.j, G.top);
  const float wg = at_top ? wb[p + 1] : 0.0f;
  const float wg_top = __shfl_down(wg, 1);  // the top lane's wb[top+1], on the top-1 lane
  if (at_top)
    ghost2_column(b.i, b.j, ub, vb, wb, col, G);
  else
    wb[p - 1] += wg_top;  // w[top-1] of the w ghost
  const CPtr6 none = {};
  vel_local_cell<false>(b.i, b.j, b.k, ub, vb, wb, Aw, Sg, M, Gr, G, none);
}

// velocity^T, gather part: stress adjoints from the weighted velocity
// adjoints of the velocity cells whose stencils read them (A is zero outside
// the velocity interior).  On the levels below the free-surface band
// (k <= top-3) this gather is left pending at the end of a step and applied
// by the next step's adj_stress_local_kernel, which reads those stress
// adjoints anyway (adj_settle_kernel applies it after the last step); the
// band, which fstr^T touches, is gathered at once.
__global__ __launch_bounds__(256) void adj_vel_gather_kernel(Adj15 Ab, CPtr6 Aw,
                                                             const AdjGrid G, int k0, int k1) {
  const BandCell b = band_cell(G, k0, k1);
  if (!b.ok) return;
  const long long q = aidx(G, b.i, b.j, b.k);
  float g[6];
  vel_gather6(Aw, q, G, g);
#pragma unroll
  for (int c = 0; c < 6; ++c) Ab.f[c][q] += g[c];
}

// The pending velocity^T gather after the last step (levels align+3 ..
// top-3).
__global__ __launch_bounds__(256) void adj_settle_kernel(Adj15 Ab, CPtr6 Aw, const AdjGrid G) {
  const Blk3 B = adj_block(G);
  const int k = AWP_ALIGN + 3 + (int)(B.x * 64 + threadIdx.x);
  const int j = 2 + AWP_NGSL + (int)(B.y * 4 + threadIdx.y);
  const int i = 2 + AWP_NGSL + B.z;
  if (k > G.top - 3 || j >= G.ny + 2 + AWP_NGSL) return;
  const long long q = aidx(G, i, j, k);
  float g[6];
  vel_gather6(Aw, q, G, g);
#pragma unroll
  for (int c = 0; c < 6; ++c) Ab.f[c][q] += g[c];
}

__global__ __launch_bounds__(256) void adj_inject_kernel(float* __restrict__ ub,
                                                         float* __restrict__ vb,
                                                         float* __restrict__ wb,
                                                         const int* __restrict__ idx, int npts,
                                                         const double* __restrict__ res,
                                                         const AdjGrid G) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= npts) return;
  const long long p = aidx(G, idx[3 * n], idx[3 * n + 1], idx[3 * n + 2]);
  ub[p] += (float)res[3 * n];
  vb[p] += (float)res[3 * n + 1];
  wb[p - 1] += (float)res[3 * n + 2];
}

__global__ __launch_bounds__(256) void adj_flush_kernel(float* __restrict__ g32,
                                                        double* __restrict__ g64, long long n) {
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    g64[i] += (double)g32[i];
    g32[i] = 0.0f;
  }
}

// residuals d - obs and the misfit 1/2 sum (d - obs)^2 in binary64: one
// partial sum per block, added by one lane per block into *J.
__global__ __launch_bounds__(256) void misfit_kernel(const float* __restrict__ d,
                                                     const float* __restrict__ obs, long long n,
                                                     double* __restrict__ res,
                                                     double* __restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double r = (double)d[i] - (double)obs[i];
    res[i] = r;
    acc += 0.5 * r * r;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void misfit_final_kernel(const double* __restrict__ part, int nb,
                                    double* __restrict__ J) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += part[b];
    J[0] = s;
  }
}

// Coefficient gradients -> gradients in lam (1/lambda), mu (1/mu), d1 (rho)
// per cell, binary64; every cell gathers from the coefficients whose means
// include it.
struct MapIn {
  const double *gd1, *gd2, *gd3, *gxl, *gxm, *gx1, *gx2, *gx3, *glm;
};

__global__ __launch_bounds__(256) void adj_map_kernel(MapIn Gi, Medium M, double* __restrict__ glam,
                                                      double* __restrict__ gmu,
                                                      double* __restrict__ grho,
                                                      const AdjGrid G) {
  const int k = (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = (int)(blockIdx.y * 4 + threadIdx.y);
  const int i = (int)blockIdx.z;
  const int NX = G.nx + 4 + AWP_NGSL2, NY = G.ny + 4 + AWP_NGSL2, NZ = G.yline;
  if (k >= NZ || j >= NY) return;
  const long long S = G.slice, Y = G.yline;
  const long long q = aidx(G, i, j, k);
  const int vi0 = 2 + AWP_NGSL, vi1 = G.nx + 1 + AWP_NGSL, vj1 = G.ny + 1 + AWP_NGSL;
  auto in_v = [&](int ii, int jj, int kk) {
    return ii >= vi0 && ii <= vi1 && jj >= vi0 && jj <= vj1 && kk >= AWP_ALIGN && kk <= G.top;
  };
  auto in_s = [&](int ii, int jj, int kk) {
    return ii >= vi0 && ii <= vi1 && jj >= vi0 && jj <= vj1 && kk >= AWP_ALIGN + 3 && kk <= G.top;
  };
  const float* d1 = M.d1;
  const double dth = (double)G.dth;
  double gr = 0.0;
  // f_d = dth / (0.25 sum4) : d f / d rho_q = -0.25 f^2 / dth
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      {  // d1: rho[c], rho[c-y], rho[c-z], rho[c-y-z] -> c = q + b y + a z
        const int ci = i, cj = j + b, ck = k + a;
        if (in_v(ci, cj, ck)) {
          const long long c = aidx(G, ci, cj, ck);
          const double f = dth / (0.25 * ((double)d1[c] + d1[c - Y] + d1[c - 1] + d1[c - Y - 1]));
          gr += Gi.gd1[c] * (-0.25 * f * f / dth);
        }
      }
      {  // d2: rho[c], rho[c+x], rho[c-z], rho[c+x-z] -> c = q - b x + a z
        const int ci = i - b, cj = j, ck = k + a;
        if (in_v(ci, cj, ck)) {
          const long long c = aidx(G, ci, cj, ck);
          const double f = dth / (0.25 * ((double)d1[c] + d1[c + S] + d1[c - 1] + d1[c + S - 1]));
          gr += Gi.gd2[c] * (-0.25 * f * f / dth);
        }
      }
      {  // d3: rho[c], rho[c+x], rho[c-y], rho[c+x-y] -> c = q - b x + a y
        const int ci = i - b, cj = j + a, ck = k;
        if (in_v(ci, cj, ck)) {
          const long long c = aidx(G, ci, cj, ck);
          const double f = dth / (0.25 * ((double)d1[c] + d1[c + S] + d1[c - Y] + d1[c + S - Y]));
          gr += Gi.gd3[c] * (-0.25 * f * f / dth);
        }
      }
    }
  double gl = 0.0, gm = 0.0;
  // 8-point means: cell p + a x - b y - c z includes q when p = q - a x + b y + c z
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const int pi = i - a, pj = j + b, pk = k + c;
        if (!in_s(pi, pj, pk)) continue;
        const long long p = aidx(G, pi, pj, pk);
        double sl8 = 0.0, sm8 = 0.0;
        for (int aa = 0; aa < 2; ++aa)
          for (int bb = 0; bb < 2; ++bb)
            for (int cc = 0; cc < 2; ++cc) {
              const long long r = p + aa * S - bb * Y - cc;
              sl8 += M.lam[r];
              sm8 += M.mu[r];
            }
        gl += Gi.gxl[p] * (-8.0 / (sl8 * sl8));
        gm += (Gi.gxl[p] + Gi.gxm[p]) * (-16.0 / (sm8 * sm8));
      }
  // 2-point means: xmu1 (p, p-z), xmu2 (p, p-y), xmu3 (p, p+x)
  {
    const int pk[2] = {k, k + 1};
    for (int t = 0; t < 2; ++t)
      if (in_s(i, j, pk[t])) {
        const long long p = aidx(G, i, j, pk[t]);
        const double s2 = (double)M.mu[p] + M.mu[p - 1];
        gm += Gi.gx1[p] * (-2.0 / (s2 * s2));
      }
    const int pj[2] = {j, j + 1};
    for (int t = 0; t < 2; ++t)
      if (in_s(i, pj[t], k)) {
        const long long p = aidx(G, i, pj[t], k);
        const double s2 = (double)M.mu[p] + M.mu[p - Y];
        gm += Gi.gx2[p] * (-2.0 / (s2 * s2));
      }
    const int pi[2] = {i, i - 1};
    for (int t = 0; t < 2; ++t)
      if (in_s(pi[t], j, k)) {
        const long long p = aidx(G, pi[t], j, k);
        const double s2 = (double)M.mu[p] + M.mu[p + S];
        gm += Gi.gx3[p] * (-2.0 / (s2 * s2));
      }
  }
  // lam_mu = (1/lam) / (2/mu + 1/lam) at the surface cells
  if (k == G.top && in_v(i, j, k)) {
    const double lc = M.lam[q], mc = M.mu[q];
    const double txl = 1.0 / lc, t2 = 2.0 / mc + txl;
    const double g = Gi.glm[(long long)i * G.lam_mu_ld + j];
    gl += g * (-(2.0 / mc) / (t2 * t2) / (lc * lc));
    gm += g * (txl * (2.0 / (mc * mc)) / (t2 * t2));
  }
  glam[q] = gl;
  gmu[q] = gm;
  grho[q] = gr;
}

AdjGrid adj_grid(int d_i) {
  const AwpGrid& g = awp_state().grid[d_i];
  AdjGrid G;
  G.nx = g.nxt; G.ny = g.nyt; G.nz = g.nzt;
  G.top = g.nzt + AWP_ALIGN - 1;
  G.yline = g.yline; G.slice = g.slice;
  G.lam_mu_ld = g.nyt + 4 + AWP_NGSL2;
  G.c1 = g.c1; G.c2 = g.c2; G.dth = g.dth; G.dh1 = g.dh1; G.DT = g.DT;
  G.jgroup = awp_tuning().adj_jgroup;
  return G;
}

Medium adj_medium(const float* const* med, const int* ww, const float* coeff,
                  const float* dcrjx, const float* dcrjy, const float* dcrjz) {
  Medium M;
  M.d1 = med[0]; M.lam = med[1]; M.mu = med[2]; M.qp = med[3]; M.qs = med[4];
  M.vx1 = med[5]; M.vx2 = med[6]; M.wwo = med[7]; M.lam_mu = med[8];
  M.coeff = coeff; M.dcrjx = dcrjx; M.dcrjy = dcrjy; M.dcrjz = dcrjz; M.ww = ww;
  return M;
}

}  // namespace

extern "C" {

void awp_adjoint_step(float* const* adj, const float* const* fwd_v, const float* const* fwd_s,
                      const float* const* med, const int* ww, const float* coeff,
                      const float* dcrjx, const float* dcrjy, const float* dcrjz,
                      float* const* scratch, float* const* grad, float* grad_lam_mu,
                      float* col, hipStream_t St) {
  const char* W = "awp_adjoint_step";
  if (!awp_grid_ok(0, W)) return;
  if (awp_state().ngrids != 1) {
    awp_set_error(AWP_ERR_STATE, "%s: the adjoint runs one grid (ngrids=%d)", W,
                  awp_state().ngrids);
    return;
  }
  const AdjGrid G = adj_grid(0);
  Adj15 A;
  for (int f = 0; f < 15; ++f) A.f[f] = adj[f];
  CPtr3 V = {{fwd_v[0], fwd_v[1], fwd_v[2]}};
  CPtr6 Sg = {{fwd_s[0], fwd_s[1], fwd_s[2], fwd_s[3], fwd_s[4], fwd_s[5]}};
  Ptr6 E = {{scratch[0], scratch[1], scratch[2], scratch[3], scratch[4], scratch[5]}};
  CPtr6 Ec = {{scratch[0], scratch[1], scratch[2], scratch[3], scratch[4], scratch[5]}};
  Ptr6 Aw = {{scratch[6], scratch[7], scratch[8], nullptr, nullptr, nullptr}};
  CPtr6 Awc = {{scratch[6], scratch[7], scratch[8], nullptr, nullptr, nullptr}};
  Grad8 Gr;
  for (int f = 0; f < 8; ++f) Gr.f[f] = grad[f];
  const Medium M = adj_medium(med, ww, coeff, dcrjx, dcrjy, dcrjz);
  const dim3 b2(64, 4);
  const dim3 g2(awp_cdiv(G.ny, 64), awp_cdiv(G.nx, 4));
  hipLaunchKernelGGL(adj_fstr_kernel, g2, b2, 0, St, A.f[2], A.f[4], A.f[5], G);
  const int nks = G.top - (AWP_ALIGN + 3) + 1;
  hipLaunchKernelGGL(adj_stress_local_kernel, dim3(awp_cdiv(nks, 64), awp_cdiv(G.ny, 4), G.nx),
                     dim3(64, 4), 0, St, A, E, V, M, Gr, G, Awc);
  // levels align .. top-2: strain gather and velocity^T in one pass (the
  // free-surface ghosts' transpose reads and writes only top-1 .. top+1)
  const int kf1 = G.top - 2;
  const int nkf = kf1 - AWP_ALIGN + 1;
  hipLaunchKernelGGL(adj_vel_local_kernel, dim3(awp_cdiv(nkf, 64), awp_cdiv(G.ny, 4), G.nx),
                     dim3(64, 4), 0, St, A.f[12], A.f[13], A.f[14], Aw, Sg, M, Gr, G, Ec,
                     AWP_ALIGN, kf1);
  // levels top-1 .. top+1: gather and ghost^T part 1; then per column ghost^T
  // part 2 and velocity^T on top-1 .. top (band launches: (cdiv(nk * ny,
  // 256), 1, nx) blocks for nk levels)
  auto band = [&](int nk) { return dim3(awp_cdiv(nk * G.ny, 256), 1, G.nx); };
  hipLaunchKernelGGL(adj_band_gather_kernel, band(3), dim3(64, 4), 0, St, A.f[12], A.f[13],
                     A.f[14], Ec, V, M.lam_mu, col, grad_lam_mu, G);
  hipLaunchKernelGGL(adj_band_ghost_kernel, band(2), dim3(64, 4), 0, St, A.f[12], A.f[13],
                     A.f[14], col, Aw, Sg, M, Gr, G);
  // the free-surface band now; the levels below stay pending for the next
  // step (or awp_adjoint_settle)
  hipLaunchKernelGGL(adj_vel_gather_kernel, band(5), dim3(64, 4), 0, St, A, Awc, G, G.top - 2,
                     G.top + 2);
  awp_check_launch(W);
}

void awp_adjoint_settle(float* const* adj, float* const* scratch, hipStream_t St) {
  const char* W = "awp_adjoint_settle";
  if (!awp_grid_ok(0, W)) return;
  const AdjGrid G = adj_grid(0);
  Adj15 A;
  for (int f = 0; f < 15; ++f) A.f[f] = adj[f];
  CPtr6 Awc = {{scratch[6], scratch[7], scratch[8], nullptr, nullptr, nullptr}};
  const int nk = G.top - 3 - (AWP_ALIGN + 3) + 1;
  if (nk <= 0) return;
  hipLaunchKernelGGL(adj_settle_kernel, dim3(awp_cdiv(nk, 64), awp_cdiv(G.ny, 4), G.nx),
                     dim3(64, 4), 0, St, A, Awc, G);
  awp_check_launch(W);
}

void awp_adjoint_inject(float* ub, float* vb, float* wb, const int* idx, int npts,
                        const double* res, hipStream_t St) {
  const char* W = "awp_adjoint_inject";
  if (!awp_grid_ok(0, W) || npts <= 0) return;
  const AdjGrid G = adj_grid(0);
  hipLaunchKernelGGL(adj_inject_kernel, dim3(awp_cdiv(npts, 256)), dim3(256), 0, St, ub, vb, wb,
                     idx, npts, res, G);
  awp_check_launch(W);
}

void awp_adjoint_flush(float* g32, double* g64, long long n, hipStream_t St) {
  const char* W = "awp_adjoint_flush";
  if (n <= 0) return;
  long long b = (n + 255) / 256;
  hipLaunchKernelGGL(adj_flush_kernel, dim3((unsigned)(b < 4096 ? b : 4096)), dim3(256), 0, St,
                     g32, g64, n);
  awp_check_launch(W);
}

void awp_misfit(const float* traces, const float* obs, long long n, double* res, double* J,
                double* work, hipStream_t St) {
  const char* W = "awp_misfit";
  if (n <= 0) return;
  const int nb = AWP_MISFIT_PARTS;
  hipLaunchKernelGGL(misfit_kernel, dim3(nb), dim3(256), 0, St, traces, obs, n, res, work);
  hipLaunchKernelGGL(misfit_final_kernel, dim3(1), dim3(64), 0, St, work, nb, J);
  awp_check_launch(W);
}

}  // extern "C"

// Scratch of awp_misfit_noalloc: one partial sum per block of misfit_kernel.
// Every GPU the library runs on holds its own copy until the process ends.
// Its size does not depend on n (the blocks stride over the samples), and
// every call overwrites all AWP_MISFIT_PARTS entries before the final sum
// reads them, so nothing needs clearing between calls.
static __device__ __align__(256) double scratch[AWP_MISFIT_PARTS];

namespace {

// Device address of `scratch`, looked up once per device with that device
// current.  Host code uses only these addresses: the array's own name is a
// host-side stand-in, not a device address.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

double* misfit_scratch(const char* W) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    awp_set_error(AWP_ERR_HIP, "%s: hipGetDevice: %s", W, hipGetErrorString(e));
    return nullptr;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    awp_set_error(AWP_ERR_ARG, "%s: device %d outside [0,%d)", W, dev, kMaxDevices);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      awp_set_error(AWP_ERR_HIP, "%s: hipGetSymbolAddress(scratch) on device %d: %s", W, dev,
                    hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

void awp_misfit_noalloc(const float* traces, const float* obs, long long n, double* res,
                        double* J, hipStream_t St) {
  const char* W = "awp_misfit_noalloc";
  if (n < 0) {
    awp_set_error(AWP_ERR_ARG, "%s: n=%lld is negative", W, n);
    return;
  }
  if (n == 0) return;
  double* work = misfit_scratch(W);
  if (!work) return;
  const int nb = AWP_MISFIT_PARTS;
  hipLaunchKernelGGL(misfit_kernel, dim3(nb), dim3(256), 0, St, traces, obs, n, res, work);
  hipLaunchKernelGGL(misfit_final_kernel, dim3(1), dim3(64), 0, St, work, nb, J);
  awp_check_launch(W);
}

void awp_adjoint_map(const double* const* gcoef, const double* grad_lam_mu,
                     const float* const* med, const int* ww, const float* coeff,
                     const float* dcrjx, const float* dcrjy, const float* dcrjz, double* glam,
                     double* gmu, double* grho, hipStream_t St) {
  const char* W = "awp_adjoint_map";
  if (!awp_grid_ok(0, W)) return;
  const AdjGrid G = adj_grid(0);
  MapIn Gi = {gcoef[0], gcoef[1], gcoef[2], gcoef[3], gcoef[4], gcoef[5], gcoef[6], gcoef[7],
              grad_lam_mu};
  const Medium M = adj_medium(med, ww, coeff, dcrjx, dcrjy, dcrjz);
  const int NX = G.nx + 4 + AWP_NGSL2, NY = G.ny + 4 + AWP_NGSL2;
  hipLaunchKernelGGL(adj_map_kernel, dim3(awp_cdiv(G.yline, 64), awp_cdiv(NY, 4), NX),
                     dim3(64, 4), 0, St, Gi, M, glam, gmu, grho, G);
  awp_check_launch(W);
}

}  // extern "C"
