// Discontinuous mesh: the kernels that couple a fine grid (index d) to the
// coarse grid below it (index d+1, 3x the spacing, 1/3 the lateral size).
//
//   dvelc2_H   2nd-order velocity at k = align+2 of a fine grid
//              (reference dvelc2, src/awp/kernel.cu:2289-2356)
//   dstrqc2_H  2nd-order stress + Q at k = align+2 of a fine grid
//              (dstrqc2, kernel.cu:2564-2784)
//   intp3d_H   coarse level nzt+align-3 -> fine level align+1, 4x4 bilinear
//              (intp3d, kernel.cu:2359-2561, 2786-2804)
//   swap_H     fine levels align+1..align+8 -> coarse levels nzt+align-2 and
//              nzt+align-1, triangular 5x5x5 weighted average
//              (swap3, kernel.cu:3032-3234)
//
// All four work on planes (O(nx*ny) per call); k is not a loop dimension, so
// the lane map is one lane per (i, j) column with j fastest.  swap3 reads
// 125 fine values per output; it stages the 8 fine levels of a tile into LDS
// once per field instead of re-reading them through the caches.
#include "awp_common.h"
#include "../../include/awp_hip.h"

namespace {

struct PlaneLaunch {
  int i0, i1, j0, j1;  // inclusive column range
  int yline;
  long long slice;
  float dth, dh1, DT;
};

// ---------------------------------------------------------------------------
// dvelc2: w from the xz/yz/zz differences, u and v from 2nd-order stress
// differences; the density means are float sums scaled by the double 0.25.
__global__ __launch_bounds__(256) void dvelc2_kernel(
    float* __restrict__ u1, float* __restrict__ v1, float* __restrict__ w1,
    const float* __restrict__ xx, const float* __restrict__ yy, const float* __restrict__ zz,
    const float* __restrict__ xy, const float* __restrict__ xz, const float* __restrict__ yz,
    const float* __restrict__ dcrjx, const float* __restrict__ dcrjy,
    const float* __restrict__ dcrjz, const float* __restrict__ d1, const PlaneLaunch L) {
  const int j = L.j0 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int i = L.i0 + (int)(blockIdx.y * 4 + threadIdx.y);
  if (j > L.j1 || i > L.i1) return;
  const int k = AWP_ALIGN + 2;
  const long long S = L.slice, Y = L.yline;
  const long long p = (long long)i * S + (long long)j * Y + k;
  float f_d3 = (float)(0.25 * (double)(d1[p] + d1[p + S] + d1[p - Y] + d1[p + S - Y]));
  f_d3 = L.dth / f_d3;
  const float f_dcrj = dcrjz[k] * dcrjx[i] * dcrjy[j];
  w1[p] = (w1[p] + f_d3 * ((xz[p + S] - xz[p]) + (yz[p] - yz[p - Y]) + (zz[p + 1] - zz[p]))) *
          f_dcrj;
  float f_d1 = (float)(0.25 * (double)(d1[p] + d1[p - Y] + d1[p - 1] + d1[p - Y - 1]));
  float f_d2 = (float)(0.25 * (double)(d1[p] + d1[p + S] + d1[p - 1] + d1[p + S - 1]));
  f_d1 = L.dth / f_d1;
  f_d2 = L.dth / f_d2;
  u1[p] = (u1[p] + f_d1 * ((xx[p] - xx[p - S]) + (xy[p] - xy[p - Y]) + (xz[p] - xz[p - 1]))) *
          f_dcrj;
  v1[p] = (v1[p] + f_d2 * ((xy[p + S] - xy[p]) + (yy[p + Y] - yy[p]) + (yz[p] - yz[p - 1]))) *
          f_dcrj;
}

// ---------------------------------------------------------------------------
// dstrqc2.  Q weights with this kernel's literal types: the qpa test is single
// precision and the else branches carry no factor 2 (kernel.cu:2638-2683).
// ca, cb: coeff[ww*2-2], coeff[ww*2-1], loaded once per cell before the
// branches.
__device__ __forceinline__ float dm_qweight(float ca, float cb, float wwo, float q) {
  float w;
  if (1. / ((double)q * 2.0) <= 200.0)
    w = (float)((double)ca * (2. * (double)q) * (2. * (double)q) + (double)cb * (2. * (double)q));
  else
    w = wwo * q;
  return w / wwo;
}

// Relaxation texture of dstrqc2 (as in dstrqc_H's table forms): the
// per-cell vx1, vx2, ww, wwo arrays (DM2_TEX_ARRAYS), or vx1, vx2, wwo from
// the 8-entry table indexed by ww - 1 with ww read (DM2_TEX_TABLE) or from
// init_texture's parity rule (DM2_TEX_RULE, mesh.c:1103-1137): ni = i - xls,
// col = ni*nj + j - yls, ww = 1 + 4((ni+1)&1) + 2((col+1)&1) + ((nzt +
// col*nzt + 2)&1) at this kernel's level k - align = 2.  On one fine column
// the level touches a line of each array it reads, so each per-cell array
// dropped saves a 128-byte line per cell.
enum { DM2_TEX_ARRAYS = 0, DM2_TEX_TABLE = 1, DM2_TEX_RULE = 2 };
struct Dm2Tex {
  float t[24];  // vx1[8], vx2[8], wwo[8]
  int xls, yls, nj, nzt;
};

template <int TEX>
__global__ __launch_bounds__(256) void dstrqc2_kernel(
    float* __restrict__ xx, float* __restrict__ yy, float* __restrict__ zz,
    float* __restrict__ xy, float* __restrict__ xz, float* __restrict__ yz,
    float* __restrict__ r1, float* __restrict__ r2, float* __restrict__ r3,
    float* __restrict__ r4, float* __restrict__ r5, float* __restrict__ r6,
    const float* __restrict__ u1, const float* __restrict__ v1, const float* __restrict__ w1,
    const float* __restrict__ lam, const float* __restrict__ mu, const float* __restrict__ qp,
    const float* __restrict__ qs, const float* __restrict__ dcrjx,
    const float* __restrict__ dcrjy, const float* __restrict__ dcrjz,
    const float* __restrict__ coeff, const float* __restrict__ vx1p,
    const float* __restrict__ vx2p, const int* __restrict__ wwp,
    const float* __restrict__ wwop, const PlaneLaunch L, const Dm2Tex TT) {
  __shared__ float s_tex[24];
  if (TEX != DM2_TEX_ARRAYS) {
    const int tid = (int)(threadIdx.y * 64 + threadIdx.x);
    if (tid < 24) s_tex[tid] = TT.t[tid];
    __syncthreads();
  }
  const int j = L.j0 + (int)(blockIdx.x * 64 + threadIdx.x);
  const int i = L.i0 + (int)(blockIdx.y * 4 + threadIdx.y);
  if (j > L.j1 || i > L.i1) return;
  const int k = AWP_ALIGN + 2;
  const long long S = L.slice, Y = L.yline;
  const long long p = (long long)i * S + (long long)j * Y + k;
  const long long ip1 = p + S, jm1 = p - Y, ijk = p + S - Y, km1 = p - 1, ik1 = p + S - 1,
                  jk1 = p - Y - 1, ijk1 = p + S - Y - 1;
  const float DT = L.DT, dth = L.dth, dh1 = L.dh1;
  const float f_dcrj = dcrjz[k] * dcrjx[i] * dcrjy[j];
  float f_vx1, f_vx2, f_wwo;
  int f_ww;
  if (TEX == DM2_TEX_ARRAYS) {
    f_vx1 = vx1p[p]; f_vx2 = vx2p[p]; f_wwo = wwop[p];
    f_ww = wwp[p];
  } else {
    if (TEX == DM2_TEX_TABLE) {
      f_ww = wwp[p];
    } else {
      const int ni = i - TT.xls, col = ni * TT.nj + (j - TT.yls);
      f_ww = 1 + 4 * ((ni + 1) & 1) + 2 * ((col + 1) & 1) + ((TT.nzt + col * TT.nzt + 2) & 1);
    }
    const int t = (f_ww - 1) & 7;
    f_vx1 = s_tex[t]; f_vx2 = s_tex[8 + t]; f_wwo = s_tex[16 + t];
  }
  float xl = awp_quot(8.0f, lam[p] + lam[ip1] + lam[jm1] + lam[ijk] + lam[km1] + lam[ik1] +
                               lam[jk1] + lam[ijk1]);
  float xm = awp_quot(16.0f, mu[p] + mu[ip1] + mu[jm1] + mu[ijk] + mu[km1] + mu[ik1] + mu[jk1] +
                                 mu[ijk1]);
  float xmu1 = awp_quot(2.0f, mu[p] + mu[km1]);
  float xmu2 = awp_quot(2.0f, mu[p] + mu[jm1]);
  float xmu3 = awp_quot(2.0f, mu[p] + mu[ip1]);
  xl = xl + xm;
  float qpa = (float)(0.0625 * (double)(qp[p] + qp[ip1] + qp[jm1] + qp[ijk] + qp[km1] +
                                        qp[ik1] + qp[jk1] + qp[ijk1]));
  const float ca = coeff[f_ww * 2 - 2], cb = coeff[f_ww * 2 - 1];
  float qpaw;
  if (1.0f / (qpa * 2.0f) <= 200.0f)
    qpaw = (float)((double)ca * (2. * (double)qpa) * (2. * (double)qpa) +
                   (double)cb * (2. * (double)qpa));
  else
    qpaw = f_wwo * qpa;
  qpaw = qpaw / f_wwo;
  float h = 0.0625f * (qs[p] + qs[ip1] + qs[jm1] + qs[ijk] + qs[km1] + qs[ik1] + qs[jk1] +
                       qs[ijk1]);
  const float hw = dm_qweight(ca, cb, f_wwo, h);
  float h1 = (float)(0.25 * (double)(qs[p] + qs[km1]));
  const float h1w = dm_qweight(ca, cb, f_wwo, h1);
  float h2 = (float)(0.250 * (double)(qs[p] + qs[jm1]));
  const float h2w = dm_qweight(ca, cb, f_wwo, h2);
  float h3 = (float)(0.250 * (double)(qs[p] + qs[ip1]));
  const float h3w = dm_qweight(ca, cb, f_wwo, h3);
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
  float f_r, f_rtmp, t, vs1, vs2, vs3;

  vs1 = u1[p + 1] - u1[p];
  vs2 = w1[p] - w1[p - S];
  f_r = r5[p];
  f_rtmp = h2 * (vs1 + vs2);
  t = xz[p] + xmu2 * (vs1 + vs2) + vx1 * f_r;
  r5[p] = f_vx2 * f_r + f_wwo * f_rtmp;
  f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
  xz[p] = (t + DT * f_rtmp) * f_dcrj;

  vs1 = v1[p + 1] - v1[p];
  vs2 = w1[p + Y] - w1[p];
  f_r = r6[p];
  f_rtmp = h3 * (vs1 + vs2);
  t = yz[p] + xmu3 * (vs1 + vs2) + vx1 * f_r;
  r6[p] = f_vx2 * f_r + f_wwo * f_rtmp;
  f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
  yz[p] = (t + DT * f_rtmp) * f_dcrj;

  vs1 = u1[p + S] - u1[p];
  vs2 = v1[p] - v1[p - Y];
  vs3 = w1[p] - w1[p - 1];
  float tmp = xl * (vs1 + vs2 + vs3);
  const float a1 = qpa * (vs1 + vs2 + vs3);
  tmp = tmp + DT * a1;

  f_r = r1[p];
  f_rtmp = -h * (vs2 + vs3) + a1;
  t = xx[p] + tmp - xm * (vs2 + vs3) + vx1 * f_r;
  r1[p] = f_vx2 * f_r + f_wwo * f_rtmp;
  f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
  xx[p] = (t + DT * f_rtmp) * f_dcrj;

  // yy and zz take f_dcrj twice, as the reference (kernel.cu:2727-2740)
  f_r = r2[p];
  f_rtmp = -h * (vs1 + vs3) + a1;
  t = (yy[p] + tmp - xm * (vs1 + vs3) + vx1 * f_r) * f_dcrj;
  r2[p] = f_vx2 * f_r + f_wwo * f_rtmp;
  f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
  yy[p] = (t + DT * f_rtmp) * f_dcrj;

  f_r = r3[p];
  f_rtmp = -h * (vs1 + vs2) + a1;
  t = (zz[p] + tmp - xm * (vs1 + vs2) + vx1 * f_r) * f_dcrj;
  r3[p] = f_vx2 * f_r + f_wwo * f_rtmp;
  f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
  zz[p] = (t + DT * f_rtmp) * f_dcrj;

  vs1 = u1[p + Y] - u1[p];
  vs2 = v1[p] - v1[p - S];
  f_r = r4[p];
  f_rtmp = h1 * (vs1 + vs2);
  t = xy[p] + xmu1 * (vs1 + vs2) + vx1 * f_r;
  r4[p] = f_vx2 * f_r + f_wwo * f_rtmp;
  f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
  xy[p] = (t + DT * f_rtmp) * f_dcrj;
}

// ---------------------------------------------------------------------------
// Field order of the nine-pointer packs below: u1 v1 w1 xx yy zz xy xz yz
// (the swap-buffer order, include/awp/pmcl3d_cons.h:80-88).
struct Fields9 {
  float* f[9];
};

struct Coupling {
  int ci0, ci1, cj0, cj1;  // coarse (i, j) ranges, inclusive
  int kc;                 // coarse level (intp3d) or top coarse level (swap)
  int ylc, ylf;
  long long slc, slf;
  int nxf_pad, nyf_pad;   // fine padded extents nxt+4+ngsl2, nyt+4+ngsl2
};

// intp3d: one lane per fine cell of one field (blockIdx.z).  Each coarse
// column (i, j) writes a 4x4 block of fine cells at (ii-1, jj-1) = 0..3 from
// its base; neighbouring blocks overlap in their last and first rows.  In
// the reference's loop order (i outer, j inner, kernel.cu:2379-2557 restated
// sequentially) an overlapped cell keeps the value of the column with the
// largest i and then the largest j, so the lane of fine offset t (0 .. 3*Ni+3
// from the first block's base) takes coarse column min(t/3, Ni) and weight
// index t - 3*that, and every fine cell is written once.  Fine offsets (ox,
// oy) per field, kernel.cu:2397-2557.
__constant__ int c_intp_ox[9] = {0, 1, 1, 1, 1, 1, 0, 0, 1};
__constant__ int c_intp_oy[9] = {0, 1, 0, 0, 0, 0, 1, 0, 1};

__device__ __forceinline__ float intp_w(int a) {
  return a == 0 ? 1.0f : a == 1 ? (float)(2. / 3.) : a == 2 ? (float)(1. / 3.) : 0.0f;
}

__global__ __launch_bounds__(256) void intp3d_kernel(const Fields9 L9, Fields9 H9,
                                                     const Coupling C) {
  const int f = (int)blockIdx.z;
  const int Ni = C.ci1 - C.ci0, Nj = C.cj1 - C.cj0;
  const int tj = (int)(blockIdx.x * 64 + threadIdx.x);
  const int ti = (int)(blockIdx.y * 4 + threadIdx.y);
  if (tj > 3 * Nj + 3 || ti > 3 * Ni + 3) return;
  const int ir = min(ti / 3, Ni), jr = min(tj / 3, Nj);
  const int a = ti - 3 * ir, b = tj - 3 * jr;  // ii-1, jj-1
  const long long pl =
      (long long)(C.ci0 + ir) * C.slc + (long long)(C.cj0 + jr) * C.ylc + C.kc;
  const float* A = L9.f[f];
  const float v0 = A[pl], v1 = A[pl + C.slc], v2 = A[pl + C.ylc], v3 = A[pl + C.slc + C.ylc];
  const float wa = intp_w(a), wb = intp_w(b), wa1 = intp_w(3 - a), wb1 = intp_w(3 - b);
  // fine base of the first coarse column, 3*(ci0-2-ngsl) + 2 + ngsl
  const int fi = (C.ci0 - 2 - AWP_NGSL) * 3 + 2 + AWP_NGSL + c_intp_ox[f] + ti;
  const int fj = (C.cj0 - 2 - AWP_NGSL) * 3 + 2 + AWP_NGSL + c_intp_oy[f] + tj;
  H9.f[f][(long long)fi * C.slf + (long long)fj * C.ylf + AWP_ALIGN + 1] =
      v0 * wa * wb + v1 * wa1 * wb + v2 * wa * wb1 + v3 * wa1 * wb1;
}

// swap3.  One wave per tile of SW_TI x SW_TJ coarse columns.  For each field
// the fine slab (3*TI+3) x (3*TJ+3) columns x 8 levels (align+1..align+8) is
// copied to LDS with 32-byte column segments, then every lane forms its
// outputs for that field from LDS.  Sums follow the reference exactly:
// loops jj, ii, kk; term = ((1/729 * wi) * wj) * wk * v in double; the
// running sum is rounded to float after every term.
constexpr int SW_TI = 4, SW_TJ = 16;
constexpr int SW_FI = 3 * SW_TI + 3, SW_FJ = 3 * SW_TJ + 3, SW_NL = 8;
constexpr int SW_CP = SW_NL + 1;  // column pitch in LDS (odd: spreads banks)
constexpr int SW_WWL = 2;         // include/awp/pmcl3d_cons.h:92
constexpr int SW_G = 4;           // slab columns per lane loaded together

struct SwapPass {
  int kc;      // 0: coarse level nzt+align-2, 1: nzt+align-1
  int kh;      // fine centre level minus (align+1)
  int oi, oj;  // fine base offsets: ih = 3*(i-10) + oi + ngsl
};

// Per field: number of passes and the passes (kernel.cu:3071-3213).
__constant__ int c_swap_np[9] = {2, 2, 1, 2, 2, 2, 2, 1, 1};
__constant__ SwapPass c_swap_pass[9][2] = {
    {{0, 2, 2, 2}, {1, 5, 2, 2}},  // u1: kh align+3, align+6
    {{0, 2, 3, 3}, {1, 5, 3, 3}},  // v1
    {{0, 3, 3, 2}, {0, 3, 3, 2}},  // w1: kh align+4
    {{0, 2, 3, 2}, {1, 5, 3, 2}},  // xx
    {{0, 2, 3, 2}, {1, 5, 3, 2}},  // yy
    {{0, 2, 3, 2}, {1, 5, 3, 2}},  // zz
    {{0, 2, 2, 3}, {1, 5, 2, 3}},  // xy
    {{0, 3, 2, 2}, {0, 3, 2, 2}},  // xz: kh align+4
    {{0, 3, 3, 3}, {0, 3, 3, 3}},  // yz: kh align+4
};

// One block per (field, tile): block f * ntiles + tile of the XCD-remapped
// order, so the blocks of one field run together and neighbouring tiles
// share their halo columns in one L2.  vec: each fine column's levels align .. align+8 arrive in
// two 16-byte loads and one 4-byte load (yline a multiple of 4 floats and
// 16-byte aligned field bases), else one 4-byte load per level.
__global__ __launch_bounds__(64) void swap3_kernel(Fields9 L9, const Fields9 H9,
                                                   const Coupling C, int ntj, int ntiles,
                                                   int vec) {
  __shared__ float slab[SW_FI * SW_FJ * SW_CP];
  __shared__ double wt[125];
  const int lane = (int)threadIdx.x;
  // XCD-aware order: each XCD takes a contiguous range of (field, tile)
  // blocks, so neighbouring tiles, which share halo columns, share an L2
  const int blk = awp_xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int f = blk / ntiles;
  const int tile = blk % ntiles;
  const int bi = tile / ntj, bj = tile % ntj;
  const int ti = lane / SW_TJ, tj = lane % SW_TJ;
  const int i0 = C.ci0 + bi * SW_TI, j0 = C.cj0 + bj * SW_TJ;
  const int i = i0 + ti, j = j0 + tj;
  const bool active = (i <= C.ci1 && j <= C.cj1);
  // fine row/column of slab index 0: 3*(i0-2-ngsl) + 2 + ngsl - WWL
  const int fi0 = (i0 - 2 - AWP_NGSL) * 3 + AWP_NGSL;
  const int fj0 = (j0 - 2 - AWP_NGSL) * 3 + AWP_NGSL;

  // weights ((1/ttl * (3-|ii|)) * (3-|jj|)) * (3-|kk|), n = (jj*5 + ii)*5 + kk;
  // ttl = sum of the unnormalised products (kernel.cu:3053-3062), exact 729
  for (int n = lane; n < 125; n += 64) {
    double ttl = 0.;
    for (int jj = -SW_WWL; jj <= SW_WWL; ++jj)
      for (int ii = -SW_WWL; ii <= SW_WWL; ++ii)
        for (int kk = -SW_WWL; kk <= SW_WWL; ++kk)
          ttl += (1. + SW_WWL - abs(jj)) * (1. + SW_WWL - abs(ii)) * (1. + SW_WWL - abs(kk));
    const int jj = n / 25 - SW_WWL, ii = (n / 5) % 5 - SW_WWL, kk = n % 5 - SW_WWL;
    wt[n] = 1 / ttl * (1. + SW_WWL - abs(ii)) * (1. + SW_WWL - abs(jj)) *
            (1. + SW_WWL - abs(kk));
  }

  // the fine slab of this field: one lane per column, levels align+1 .. +8.
  // The columns go in groups of SW_G per lane, each group's loads issued
  // before any is stored to LDS; a column past the slab or the padded grid
  // loads column 0 of the slab instead and stores nothing.
  const float* src = H9.f[f];
  constexpr int NCOL = SW_FI * SW_FJ;
  constexpr int NIT = (NCOL + 63) / 64;
  for (int g0 = 0; g0 < NIT; g0 += SW_G) {
    float v[SW_G][SW_NL];
    bool ok[SW_G];
#pragma unroll
    for (int u = 0; u < SW_G; ++u) {
      const int col = lane + (g0 + u) * 64;
      const int a = col / SW_FJ, b = col % SW_FJ;
      ok[u] = g0 + u < NIT && col < NCOL && fi0 + a < C.nxf_pad && fj0 + b < C.nyf_pad;
      const int fi = ok[u] ? fi0 + a : fi0, fj = ok[u] ? fj0 + b : fj0;
      const float* q = src + (long long)fi * C.slf + (long long)fj * C.ylf + AWP_ALIGN;
      if (vec) {
        const float4 lo = *reinterpret_cast<const float4*>(q);
        const float4 hi = *reinterpret_cast<const float4*>(q + 4);
        v[u][0] = lo.y; v[u][1] = lo.z; v[u][2] = lo.w;
        v[u][3] = hi.x; v[u][4] = hi.y; v[u][5] = hi.z; v[u][6] = hi.w;
        v[u][7] = q[8];
      } else {
#pragma unroll
        for (int l = 0; l < SW_NL; ++l) v[u][l] = q[1 + l];
      }
    }
#pragma unroll
    for (int u = 0; u < SW_G; ++u)
      if (ok[u]) {
        float* d = &slab[(lane + (g0 + u) * 64) * SW_CP];
#pragma unroll
        for (int l = 0; l < SW_NL; ++l) d[l] = v[u][l];
      }
  }
  __syncthreads();
  if (!active) return;
  const long long pc = (long long)i * C.slc + (long long)j * C.ylc + C.kc;
  const int np = c_swap_np[f];
  for (int q = 0; q < np; ++q) {
    const SwapPass P = c_swap_pass[f][q];
    const int a0 = 3 * ti + P.oi - SW_WWL, b0 = 3 * tj + P.oj - SW_WWL, l0 = P.kh - SW_WWL;
    float sum = 0.0f;
    int n = 0;
    for (int jj = 0; jj < 5; ++jj)
      for (int ii = 0; ii < 5; ++ii) {
        const float* col = &slab[((a0 + ii) * SW_FJ + (b0 + jj)) * SW_CP + l0];
#pragma unroll
        for (int kk = 0; kk < 5; ++kk, ++n)
          sum = (float)((double)sum + wt[n] * (double)col[kk]);
      }
    L9.f[f][pc + P.kc] = sum;
  }
}

bool dm_grid_pair(int fine, const char* W, const AwpGrid** gf, const AwpGrid** gc) {
  const AwpState& s = awp_state();
  if (!awp_grid_ok(fine, W) || !awp_grid_ok(fine + 1, W)) return false;
  *gf = &s.grid[fine];
  *gc = &s.grid[fine + 1];
  if ((*gf)->nxt != 3 * (*gc)->nxt || (*gf)->nyt != 3 * (*gc)->nyt) {
    awp_set_error(AWP_ERR_ARG, "%s: fine grid %d (%d x %d) is not 3x coarse grid %d (%d x %d)",
                  W, fine, (*gf)->nxt, (*gf)->nyt, fine + 1, (*gc)->nxt, (*gc)->nyt);
    return false;
  }
  if ((*gc)->nzt < 4) {
    awp_set_error(AWP_ERR_ARG, "%s: coarse grid %d has nzt=%d < 4", W, fine + 1, (*gc)->nzt);
    return false;
  }
  return true;
}

Coupling make_coupling(const AwpGrid& gf, const AwpGrid& gc) {
  Coupling C;
  C.ylc = gc.yline;
  C.slc = gc.slice;
  C.ylf = gf.yline;
  C.slf = gf.slice;
  C.nxf_pad = gf.nxt + 4 + AWP_NGSL2;
  C.nyf_pad = gf.nyt + 4 + AWP_NGSL2;
  return C;
}

}  // namespace

extern "C" {

void dvelc2_H(float* u1, float* v1, float* w1, float* xx, float* yy, float* zz, float* xy,
              float* xz, float* yz, float* dcrjx, float* dcrjy, float* dcrjz, float* d_1, int nxt,
              int nyt, hipStream_t St, int d_i) {
  const char* W = "dvelc2_H";
  if (!awp_grid_ok(d_i, W)) return;
  const AwpGrid& g = awp_state().grid[d_i];
  if (nxt != g.nxt || nyt != g.nyt) {
    awp_set_error(AWP_ERR_ARG, "%s: nxt=%d nyt=%d disagree with grid %d (%d, %d)", W, nxt, nyt,
                  d_i, g.nxt, g.nyt);
    return;
  }
  if (g.nzt < 3) {
    awp_set_error(AWP_ERR_ARG, "%s: grid %d has nzt=%d < 3", W, d_i, g.nzt);
    return;
  }
  PlaneLaunch L;
  L.i0 = 2 + AWP_NGSL; L.i1 = nxt + 1 + AWP_NGSL;   // kernel.cu:2301
  L.j0 = 2 + AWP_NGSL; L.j1 = nyt + 1 + AWP_NGSL;
  L.yline = g.yline; L.slice = g.slice;
  L.dth = g.dth; L.dh1 = g.dh1; L.DT = g.DT;
  dim3 grid(awp_cdiv(L.j1 - L.j0 + 1, 64), awp_cdiv(L.i1 - L.i0 + 1, 4));
  hipLaunchKernelGGL(dvelc2_kernel, grid, dim3(64, 4), 0, St, u1, v1, w1, xx, yy, zz, xy, xz, yz,
                     dcrjx, dcrjy, dcrjz, d_1, L);
  awp_check_launch(W);
}

}  // extern "C"

namespace {

void dstrqc2_common(const char* W, float* xx, float* yy, float* zz, float* xy, float* xz,
                    float* yz, float* r1, float* r2, float* r3, float* r4, float* r5, float* r6,
                    float* u1, float* v1, float* w1, float* lam, float* mu, float* qp, float* qs,
                    float* dcrjx, float* dcrjy, float* dcrjz, int nxt, int nyt, hipStream_t St,
                    float* coeff, float* d_vx1, float* d_vx2, int* d_ww, float* d_wwo, int s_i,
                    int e_i, int s_j, int e_j, int d_i, int tex_mode, const float* tex) {
  if (!awp_grid_ok(d_i, W)) return;
  const AwpGrid& g = awp_state().grid[d_i];
  if (nxt != g.nxt || nyt != g.nyt) {
    awp_set_error(AWP_ERR_ARG, "%s: nxt=%d nyt=%d disagree with grid %d (%d, %d)", W, nxt, nyt,
                  d_i, g.nxt, g.nyt);
    return;
  }
  if (g.nzt < 3) {
    awp_set_error(AWP_ERR_ARG, "%s: grid %d has nzt=%d < 3", W, d_i, g.nzt);
    return;
  }
  if (s_i > e_i || s_j > e_j) return;
  // i-1 .. i+1 and j-1 .. j+1 must stay in the padded grid
  if (s_i < 2 || e_i + 2 > g.nxt + 3 + AWP_NGSL2 || s_j < 2 || e_j + 2 > g.nyt + 3 + AWP_NGSL2) {
    awp_set_error(AWP_ERR_ARG, "%s: range i[%d,%d] j[%d,%d] leaves the padded grid (%d x %d)", W,
                  s_i, e_i, s_j, e_j, g.nxt, g.nyt);
    return;
  }
  PlaneLaunch L;
  L.i0 = s_i; L.i1 = e_i; L.j0 = s_j; L.j1 = e_j;
  L.yline = g.yline; L.slice = g.slice;
  L.dth = g.dth; L.dh1 = g.dh1; L.DT = g.DT;
  dim3 grid(awp_cdiv(e_j - s_j + 1, 64), awp_cdiv(e_i - s_i + 1, 4));
  Dm2Tex TT = {};
  if (tex_mode != DM2_TEX_ARRAYS) {
    for (int m = 0; m < 24; ++m) TT.t[m] = tex[m];
    // init_texture's region on one rank: the interior columns, all levels
    TT.xls = 2 + AWP_NGSL;
    TT.yls = 2 + AWP_NGSL;
    TT.nj = g.nyt;
    TT.nzt = g.nzt;
  }
  if (tex_mode == DM2_TEX_RULE)
    hipLaunchKernelGGL(dstrqc2_kernel<DM2_TEX_RULE>, grid, dim3(64, 4), 0, St, xx, yy, zz, xy,
                       xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1, lam, mu, qp, qs, dcrjx, dcrjy,
                       dcrjz, coeff, nullptr, nullptr, nullptr, nullptr, L, TT);
  else if (tex_mode == DM2_TEX_TABLE)
    hipLaunchKernelGGL(dstrqc2_kernel<DM2_TEX_TABLE>, grid, dim3(64, 4), 0, St, xx, yy, zz, xy,
                       xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1, lam, mu, qp, qs, dcrjx, dcrjy,
                       dcrjz, coeff, nullptr, nullptr, d_ww, nullptr, L, TT);
  else
    hipLaunchKernelGGL(dstrqc2_kernel<DM2_TEX_ARRAYS>, grid, dim3(64, 4), 0, St, xx, yy, zz, xy,
                       xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1, lam, mu, qp, qs, dcrjx, dcrjy,
                       dcrjz, coeff, d_vx1, d_vx2, d_ww, d_wwo, L, TT);
  awp_check_launch(W);
}

}  // namespace

extern "C" {

void dstrqc2_H(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz, float* r1,
               float* r2, float* r3, float* r4, float* r5, float* r6, float* u1, float* v1,
               float* w1, float* lam, float* mu, float* qp, float* qs, float* dcrjx, float* dcrjy,
               float* dcrjz, int nxt, int nyt, hipStream_t St, float* coeff, float* d_vx1,
               float* d_vx2, int* d_ww, float* d_wwo, int s_i, int e_i, int s_j, int e_j,
               int d_i) {
  const char* W = "dstrqc2_H";
  if (!d_vx1 || !d_vx2 || !d_ww || !d_wwo) {
    awp_set_error(AWP_ERR_ARG, "%s: a relaxation-texture pointer is NULL", W);
    return;
  }
  dstrqc2_common(W, xx, yy, zz, xy, xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1, lam, mu, qp, qs,
                 dcrjx, dcrjy, dcrjz, nxt, nyt, St, coeff, d_vx1, d_vx2, d_ww, d_wwo, s_i, e_i,
                 s_j, e_j, d_i, DM2_TEX_ARRAYS, nullptr);
}

void awp_dstrqc2_tex(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz,
                     float* r1, float* r2, float* r3, float* r4, float* r5, float* r6, float* u1,
                     float* v1, float* w1, float* lam, float* mu, float* qp, float* qs,
                     float* dcrjx, float* dcrjy, float* dcrjz, int nxt, int nyt, hipStream_t St,
                     float* coeff, const float* tex, int* ww, int s_i, int e_i, int s_j, int e_j,
                     int d_i) {
  const char* W = "awp_dstrqc2_tex";
  if (!tex) {
    awp_set_error(AWP_ERR_ARG, "%s: the texture table is NULL", W);
    return;
  }
  dstrqc2_common(W, xx, yy, zz, xy, xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1, lam, mu, qp, qs,
                 dcrjx, dcrjy, dcrjz, nxt, nyt, St, coeff, nullptr, nullptr, ww, nullptr, s_i,
                 e_i, s_j, e_j, d_i, ww ? DM2_TEX_TABLE : DM2_TEX_RULE, tex);
}

void intp3d_H(float* u1l, float* v1l, float* w1l, float* xxl, float* yyl, float* zzl, float* xyl,
              float* xzl, float* yzl, float* u1h, float* v1h, float* w1h, float* xxh, float* yyh,
              float* zzh, float* xyh, float* xzh, float* yzh, int nxtl, int nytl, int rank,
              hipStream_t St, int d_i) {
  (void)rank;
  const char* W = "intp3d_H";
  const AwpGrid *gf, *gc;
  if (d_i < 1) {
    awp_set_error(AWP_ERR_ARG, "%s: d_i=%d names the coarse grid and must be >= 1", W, d_i);
    return;
  }
  if (!dm_grid_pair(d_i - 1, W, &gf, &gc)) return;
  if (nxtl != gc->nxt || nytl != gc->nyt) {
    awp_set_error(AWP_ERR_ARG, "%s: nxtl=%d nytl=%d disagree with grid %d (%d, %d)", W, nxtl,
                  nytl, d_i, gc->nxt, gc->nyt);
    return;
  }
  Coupling C = make_coupling(*gf, *gc);
  C.ci0 = AWP_NGSL; C.ci1 = nxtl + 3 + AWP_NGSL;  // kernel.cu:2379
  C.cj0 = AWP_NGSL; C.cj1 = nytl + 3 + AWP_NGSL;
  C.kc = gc->nzt + AWP_ALIGN - 3;
  Fields9 L9 = {{u1l, v1l, w1l, xxl, yyl, zzl, xyl, xzl, yzl}};
  Fields9 H9 = {{u1h, v1h, w1h, xxh, yyh, zzh, xyh, xzh, yzh}};
  // one lane per fine cell: 3*(coarse columns - 1) + 4 per direction, 9 fields
  dim3 grid(awp_cdiv(3 * (C.cj1 - C.cj0) + 4, 64), awp_cdiv(3 * (C.ci1 - C.ci0) + 4, 4), 9);
  hipLaunchKernelGGL(intp3d_kernel, grid, dim3(64, 4), 0, St, L9, H9, C);
  awp_check_launch(W);
}

void swap_H(float* xxl, float* yyl, float* zzl, float* xyl, float* xzl, float* yzl, float* u1l,
            float* v1l, float* w1l, float* xxh, float* yyh, float* zzh, float* xyh, float* xzh,
            float* yzh, float* u1h, float* v1h, float* w1h, int nxtl, int nytl, float* buf_L,
            float* buf_R, float* buf_F, float* buf_B, int rank, hipStream_t St, int d_i) {
  // The halo buffers serve bgcaccess (kernel.cu:2809-2852) for fine indices
  // outside the padded array; with WWL = 2 the indices read here span
  // [2, nxt_fine+19] x [2, nyt_fine+19], inside it, so they are never read.
  (void)buf_L; (void)buf_R; (void)buf_F; (void)buf_B; (void)rank;
  const char* W = "swap_H";
  const AwpGrid *gf, *gc;
  if (!dm_grid_pair(d_i, W, &gf, &gc)) return;
  if (nxtl != gc->nxt || nytl != gc->nyt) {
    awp_set_error(AWP_ERR_ARG, "%s: nxtl=%d nytl=%d disagree with grid %d (%d, %d)", W, nxtl,
                  nytl, d_i + 1, gc->nxt, gc->nyt);
    return;
  }
  if (gf->nzt < 9) {
    awp_set_error(AWP_ERR_ARG, "%s: fine grid %d has nzt=%d < 9", W, d_i, gf->nzt);
    return;
  }
  Coupling C = make_coupling(*gf, *gc);
  C.ci0 = AWP_NGSL; C.ci1 = nxtl + AWP_NGSL + 4;  // kernel.cu:3066
  C.cj0 = AWP_NGSL; C.cj1 = nytl + AWP_NGSL + 4;
  C.kc = gc->nzt + AWP_ALIGN - 2;
  Fields9 L9 = {{u1l, v1l, w1l, xxl, yyl, zzl, xyl, xzl, yzl}};
  Fields9 H9 = {{u1h, v1h, w1h, xxh, yyh, zzh, xyh, xzh, yzh}};
  const int nti = awp_cdiv(C.ci1 - C.ci0 + 1, SW_TI);
  const int ntj = awp_cdiv(C.cj1 - C.cj0 + 1, SW_TJ);
  int vec = (C.ylf % 4 == 0) ? 1 : 0;
  for (int f = 0; f < 9; ++f)
    if (reinterpret_cast<uintptr_t>(H9.f[f]) % 16 != 0) vec = 0;
  hipLaunchKernelGGL(swap3_kernel, dim3(9 * nti * ntj), dim3(64), 0, St, L9, H9, C, ntj,
                     nti * ntj, vec);
  awp_check_launch(W);
}

}  // extern "C"
