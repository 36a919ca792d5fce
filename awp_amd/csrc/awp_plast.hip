// Drucker-Prager plasticity, gfx950: yield factor + plastic strain
// (reference drprecpc_calc_opt, src/awp/kernel.cu:1253-1422) and its
// application to the stresses (drprecpc_app, :1650-1778), with the initial
// stress rotated from the principal frame on the fly (rotate_principal,
// :43-64).
//
// drprecpc_calc: wave64 along k, BJ waves along j, i-march with the plane
// (i+1) values of sigma2/xy/xz/mu and the rotated initial stress carried in
// registers, so every per-cell array is fetched once and each plane's
// initial stress is rotated once.  drprecpc_app only works where the
// averaged yield factor is below one; it is a pointwise update read from
// L1/L2 neighbours.
//
// Arithmetic follows the reference expression trees (double-literal
// sub-expressions in double, -ffp-contract=off).  cos/sin/exp are evaluated
// in double and rounded to float; the CPU oracle does the same.
#include <math.h>

#include "awp_common.h"
#include "awp_tuning.h"
#include "../../include/awp_hip.h"

namespace {

struct PlastLaunch {
  int k0, k1, j0, j1, i0, ni, ichunk, nkt, njt, nblocks, remap, srfpos, yline;
  int kw;  // k-run width (awp_pick_kw)
  long long slice;
  float DH, DT, Tv_fixed;
  float fmajor, fminor;
  float Rz[9], RzT[9];
};

// rotate_principal: ssp = RzT * diag((s2+pf)*fmajor, s2+pf, (s2+pf)*fminor) * Rz
// - pf on the diagonal, with matmul3's accumulation order c = ((0+a0b0)+a1b1)+a2b2.
__device__ __forceinline__ void rotate6(float sigma2, float pfluid, const PlastLaunch& L,
                                        float* o /* xx, yy, zz, xy, xz, yz */) {
  float ss[9], tmp[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) ss[q] = 0.0f;
  ss[0] = (sigma2 + pfluid) * L.fmajor;
  ss[4] = sigma2 + pfluid;
  ss[8] = (sigma2 + pfluid) * L.fminor;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      float acc = 0.0f;
#pragma unroll
      for (int c = 0; c < 3; ++c) acc += L.RzT[a * 3 + c] * ss[c * 3 + b];
      tmp[a * 3 + b] = acc;
    }
  auto ent = [&](int a, int b) {
    float acc = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c) acc += tmp[a * 3 + c] * L.Rz[c * 3 + b];
    return acc;
  };
  o[0] = ent(0, 0) - pfluid;
  o[1] = ent(1, 1) - pfluid;
  o[2] = ent(2, 2) - pfluid;
  o[3] = ent(0, 1);
  o[4] = ent(0, 2);
  o[5] = ent(1, 2);
}

__device__ __forceinline__ float pfluid_at(int srfpos, int k, float DH) {
  const float depth = (float)(srfpos - k) * DH;
  if (depth > 0) return (float)(((double)depth + (double)DH * 0.5) * 9.81e3);
  return (float)((double)DH / 2. * 9.81e3);
}

template <int BJ, int KW = 64>
__global__ __launch_bounds__(64 * BJ) void drprecpc_calc_kernel(
    const float* __restrict__ xx, const float* __restrict__ yy, const float* __restrict__ zz,
    const float* __restrict__ xy, const float* __restrict__ xz, const float* __restrict__ yz,
    const float* __restrict__ mu, const float* __restrict__ d1,
    const float* __restrict__ sigma2, float* __restrict__ yldfac,
    const float* __restrict__ cohes, const float* __restrict__ phi, float* __restrict__ neta,
    const PlastLaunch L) {
  const int b = L.remap ? awp_xcd_remap(blockIdx.x, L.nblocks) : (int)blockIdx.x;
  const int kt = b % L.nkt;
  const int jt = (b / L.nkt) % L.njt;
  const int ic = b / (L.nkt * L.njt);
  // KW = 32: a wave covers 32 levels of each of two j rows
  const int lane = (int)threadIdx.x;
  const int k = L.k0 + kt * KW + (KW == 64 ? lane : lane % KW);
  const int j = L.j0 + (jt * BJ + (int)threadIdx.y) * (64 / KW) + (KW == 64 ? 0 : lane / KW);
  if (k > L.k1 || j > L.j1) return;
  const int ia = L.i0 + ic * L.ichunk;
  int nsteps = L.ni - ic * L.ichunk;
  if (nsteps > L.ichunk) nsteps = L.ichunk;
  const int Y = L.yline;
  const long long S = L.slice;
  const int o = j * Y + k;
  const float pfluid = pfluid_at(L.srfpos, k, L.DH);
  const float DT = L.DT, DH = L.DH;

  // plane ia: carried values
  long long p = (long long)ia * S + o;
  float s2c = sigma2[p];
  float xyc = xy[p], xyjm = xy[p - Y];
  float xzc = xz[p], xzkm = xz[p - 1];
  float muc = mu[p];
  float ini[6];
  rotate6(s2c, pfluid, L, ini);

  for (int s = 0; s < nsteps; ++s) {
    p = (long long)(ia + s) * S + o;
    const long long pn = p + S;
    const float s2n = sigma2[pn];
    const float xyn = xy[pn], xynjm = xy[pn - Y];
    const float xzn = xz[pn], xznkm = xz[pn - 1];
    const float mun = mu[pn];
    float inn[6];
    rotate6(s2n, pfluid, L, inn);

    const float iixx = 0.5f * (ini[0] + inn[0]);
    const float iiyy = 0.5f * (ini[1] + inn[1]);
    const float iizz = 0.5f * (ini[2] + inn[2]);
    float Sxx = xx[p] + iixx;
    float Syy = yy[p] + iiyy;
    float Szz = zz[p] + iizz;
    float Sxz = 0.25f * (xzc + xzn + xzkm + xznkm) + 0.5f * (ini[4] + inn[4]);
    float Syz = 0.25f * (yz[p] + yz[p - Y] + yz[p - 1] + yz[p - Y - 1]) + 0.5f * (ini[5] + inn[5]);
    float Sxy = 0.25f * (xyc + xyn + xyjm + xynjm) + 0.5f * (ini[3] + inn[3]);
    float Tv = DH / sqrtf(1.0f / (muc * d1[p]));
    if (L.Tv_fixed > 0.0f) Tv = L.Tv_fixed;
    const float Sxxp = Sxx, Syyp = Syy, Szzp = Szz, Sxyp = Sxy, Sxzp = Sxz, Syzp = Syz;
    const float rphi = phi[p] * 0.017453292519943295f;
    const float sigma_m = (Sxx + Syy + Szz) / 3.0f;
    const float SDxx = Sxx - sigma_m, SDyy = Syy - sigma_m, SDzz = Szz - sigma_m;
    const float secinv =
        0.5f * (SDxx * SDxx + SDyy * SDyy + SDzz * SDzz) + Sxz * Sxz + Sxy * Sxy + Syz * Syz;
    const float sqrtSecinv = sqrtf(secinv);
    const float taulim2 = cohes[p] * (float)cos((double)rphi) -
                          (sigma_m + pfluid) * (float)sin((double)rphi);
    const float taulim = taulim2 > 0.0f ? taulim2 : 0.0f;
    float iyldfac;
    if (sqrtSecinv > taulim) {
      iyldfac = taulim / sqrtSecinv +
                (1.0f - taulim / sqrtSecinv) * (float)exp((double)(-DT / Tv));
      Sxx = SDxx * iyldfac + sigma_m;
      Syy = SDyy * iyldfac + sigma_m;
      Szz = SDzz * iyldfac + sigma_m;
      Sxz = Sxz * iyldfac;
      Sxy = Sxy * iyldfac;
      Syz = Syz * iyldfac;
    } else {
      iyldfac = 1.0f;
    }
    if (iyldfac < 1.0f) {
      yldfac[p] = iyldfac;
      const float xm = 2.0f / (muc + mun);
      const float dxx = (Sxx - Sxxp) / xm, dyy = (Syy - Syyp) / xm, dzz = (Szz - Szzp) / xm;
      const float dxy = (Sxy - Sxyp) / xm, dxz = (Sxz - Sxzp) / xm, dyz = (Syz - Syzp) / xm;
      neta[p] = neta[p] + sqrtf(0.5f * (dxx * dxx + dyy * dyy + dzz * dzz) +
                                2.0f * (dxy * dxy + dxz * dxz + dyz * dyz));
    } else {
      yldfac[p] = 1.0f;
    }
    // advance the x window
    s2c = s2n;
    xyc = xyn; xyjm = xynjm;
    xzc = xzn; xzkm = xznkm;
    muc = mun;
#pragma unroll
    for (int q = 0; q < 6; ++q) ini[q] = inn[q];
  }
}

template <int BJ, int KW = 64>
__global__ __launch_bounds__(64 * BJ) void drprecpc_app_kernel(
    float* __restrict__ xx, float* __restrict__ yy, float* __restrict__ zz,
    float* __restrict__ xy, float* __restrict__ xz, float* __restrict__ yz,
    const float* __restrict__ sigma2, const float* __restrict__ yldfac, const PlastLaunch L) {
  const int b = L.remap ? awp_xcd_remap(blockIdx.x, L.nblocks) : (int)blockIdx.x;
  const int kt = b % L.nkt;
  const int jt = (b / L.nkt) % L.njt;
  const int ic = b / (L.nkt * L.njt);
  // KW = 32: a wave covers 32 levels of each of two j rows
  const int lane = (int)threadIdx.x;
  const int k = L.k0 + kt * KW + (KW == 64 ? lane : lane % KW);
  const int j = L.j0 + (jt * BJ + (int)threadIdx.y) * (64 / KW) + (KW == 64 ? 0 : lane / KW);
  if (k > L.k1 || j > L.j1) return;
  const int ia = L.i0 + ic * L.ichunk;
  int nsteps = L.ni - ic * L.ichunk;
  if (nsteps > L.ichunk) nsteps = L.ichunk;
  const int Y = L.yline;
  const long long S = L.slice;
  const int o = j * Y + k;
  const float pf = pfluid_at(L.srfpos, k, L.DH);
  const float pf1 = pfluid_at(L.srfpos, k - 1, L.DH);  // depth of the k+1 level

  for (int s = 0; s < nsteps; ++s) {
    const long long p = (long long)(ia + s) * S + o;
    const long long im1 = p - S, ip1 = p + S, jp1 = p + Y, kp1 = p + 1;
    const float y0 = yldfac[p];
    const float iy_xz = (float)(0.25 * (double)(yldfac[im1] + y0 + yldfac[im1 + 1] + yldfac[kp1]));
    const float iy_yz = (float)(0.25 * (double)(y0 + yldfac[jp1] + yldfac[jp1 + 1] + yldfac[kp1]));
    const float iy_xy = (float)(0.25 * (double)(y0 + yldfac[jp1] + yldfac[im1] + yldfac[im1 + Y]));
    if (!(y0 < 1.) && !(iy_xz < 1.) && !(iy_yz < 1.) && !(iy_xy < 1.)) continue;
    float r0[6];  // rotation at p with pfluid: shared by every branch
    rotate6(sigma2[p], pf, L, r0);
    if (y0 < 1.) {
      float rb[6];
      rotate6(sigma2[ip1], pf1, L, rb);
      const float iixx = (float)(0.5 * (double)(r0[0] + rb[0]));
      const float iiyy = (float)(0.5 * (double)(r0[1] + rb[1]));
      const float iizz = (float)(0.5 * (double)(r0[2] + rb[2]));
      const float x0 = xx[p], yv = yy[p], z0 = zz[p];
      const float Sxx = x0 + iixx, Syy = yv + iiyy, Szz = z0 + iizz;
      const float sigma_m = (float)((double)(Sxx + Syy + Szz) / 3.);
      const float SDxx = x0 + iixx - sigma_m;
      const float SDyy = yv + iiyy - sigma_m;
      const float SDzz = z0 + iizz - sigma_m;
      xx[p] = SDxx * y0 + sigma_m - iixx;
      yy[p] = SDyy * y0 + sigma_m - iiyy;
      zz[p] = SDzz * y0 + sigma_m - iizz;
    }
    if (iy_xz < 1.) {
      float rk[6];
      rotate6(sigma2[kp1], pf1, L, rk);
      const float iist = (float)(0.5 * (double)(r0[4] + rk[4]));
      xz[p] = (xz[p] + iist) * iy_xz - iist;
    }
    if (iy_yz < 1.) {
      float t2[6], t3[6], t4[6], t5[6], t6[6], t7[6], t8[6];
      rotate6(sigma2[ip1], pf, L, t2);
      rotate6(sigma2[kp1], pf1, L, t3);
      rotate6(sigma2[ip1 + 1], pf1, L, t4);
      rotate6(sigma2[jp1], pf, L, t5);
      rotate6(sigma2[ip1 + Y], pf, L, t6);
      rotate6(sigma2[jp1 + 1], pf1, L, t7);
      rotate6(sigma2[ip1 + Y + 1], pf1, L, t8);
      const float iist = (float)(0.125 * (double)(r0[5] + t2[5] + t3[5] + t4[5] + t5[5] + t6[5] +
                                                  t7[5] + t8[5]));
      yz[p] = (yz[p] + iist) * iy_yz - iist;
    }
    if (iy_xy < 1.) {
      float rj[6];
      rotate6(sigma2[jp1], pf, L, rj);
      const float iist = (float)(0.5 * (double)(r0[3] + rj[3]));
      xy[p] = (xy[p] + iist) * iy_xy - iist;
    }
  }
}

bool plast_setup(const char* W, int nzt, int xls, int xre, int yls, int yre, int d_i,
                 float Tv_fixed, bool app, PlastLaunch& L) {
  if (!awp_grid_ok(d_i, W)) return false;
  const AwpGrid& g = awp_state().grid[d_i];
  if (nzt != g.nzt) {
    awp_set_error(AWP_ERR_ARG, "%s: nzt=%d disagrees with grid %d (%d)", W, nzt, d_i, g.nzt);
    return false;
  }
  if (xls > xre || yls > yre) return false;
  // calc reads i+1, j-1, k-1; app reads i-1, i+1, j+1, k+1
  if (xls < 1 || xre + 1 >= g.nxt + 4 + AWP_NGSL2 || yls < 1 ||
      yre + 1 >= g.nyt + 4 + AWP_NGSL2) {
    awp_set_error(AWP_ERR_ARG, "%s: range i[%d,%d] j[%d,%d] leaves the padded grid", W, xls, xre,
                  yls, yre);
    return false;
  }
  const AwpPlastConst& pc = awp_state().plast;
  L.k0 = AWP_ALIGN;
  L.k1 = nzt + AWP_ALIGN - 1;
  L.j0 = yls;
  L.j1 = yre;
  L.i0 = xls;
  L.ni = xre - xls + 1;
  L.srfpos = g.nzt + AWP_ALIGN - 1;
  L.yline = g.yline;
  L.slice = g.slice;
  L.DH = g.DH;
  L.DT = g.DT;
  L.Tv_fixed = Tv_fixed;
  L.fmajor = pc.fmajor;
  L.fminor = pc.fminor;
  for (int q = 0; q < 9; ++q) {
    L.Rz[q] = pc.Rz[q];
    L.RzT[q] = pc.RzT[q];
  }
  L.remap = awp_tuning().xcd_remap;
  L.kw = awp_pick_kw(L.k1 - L.k0 + 1, awp_tuning().kw);
  L.nkt = awp_cdiv(L.k1 - L.k0 + 1, L.kw);
  L.njt = awp_cdiv(L.j1 - L.j0 + 1, 4 * (64 / L.kw));
  int nic, ichunk;
  awp_split_i(L.ni, L.nkt * L.njt, 0, app ? 4096 : 2048, &nic, &ichunk);
  L.ichunk = ichunk;
  L.nblocks = L.nkt * L.njt * nic;
  return true;
}

void calc_common(const char* W, float* xx, float* yy, float* zz, float* xy, float* xz,
                 float* yz, float* mu, float* d1, float* sigma2, float* yldfac, float* cohes,
                 float* phi, float* neta, int nzt, int xls, int xre, int yls, int yre,
                 hipStream_t St, int d_i, float Tv) {
  PlastLaunch L;
  if (!plast_setup(W, nzt, xls, xre, yls, yre, d_i, Tv, false, L)) return;
  if (L.kw == 32)
    hipLaunchKernelGGL((drprecpc_calc_kernel<4, 32>), dim3(L.nblocks), dim3(64, 4), 0, St, xx, yy,
                       zz, xy, xz, yz, mu, d1, sigma2, yldfac, cohes, phi, neta, L);
  else
    hipLaunchKernelGGL((drprecpc_calc_kernel<4, 64>), dim3(L.nblocks), dim3(64, 4), 0, St, xx, yy,
                       zz, xy, xz, yz, mu, d1, sigma2, yldfac, cohes, phi, neta, L);
  awp_check_launch(W);
}

}  // namespace

extern "C" {

void drprecpc_calc_H_opt(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz,
                         float* mu, float* d1, float* sigma2, float* yldfac, float* cohes,
                         float* phi, float* neta, int nzt, int xls, int xre, int yls, int yre,
                         hipStream_t St, int d_i) {
  calc_common("drprecpc_calc_H_opt", xx, yy, zz, xy, xz, yz, mu, d1, sigma2, yldfac, cohes, phi,
              neta, nzt, xls, xre, yls, yre, St, d_i, 0.0f);
}

void awp_drprecpc_calc_tv(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz,
                          float* mu, float* d1, float* sigma2, float* yldfac, float* cohes,
                          float* phi, float* neta, int nzt, int xls, int xre, int yls, int yre,
                          hipStream_t St, int d_i, float Tv) {
  if (!(Tv > 0.0f)) {
    awp_set_error(AWP_ERR_ARG, "awp_drprecpc_calc_tv: Tv=%g must be positive", (double)Tv);
    return;
  }
  calc_common("awp_drprecpc_calc_tv", xx, yy, zz, xy, xz, yz, mu, d1, sigma2, yldfac, cohes,
              phi, neta, nzt, xls, xre, yls, yre, St, d_i, Tv);
}

void drprecpc_app_H(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz,
                    float* mu, float* sigma2, float* yldfac, int nzt, int xls, int xre, int yls,
                    int yre, hipStream_t St, int d_i) {
  (void)mu;  // read by the reference only for a value it never uses (kernel.cu:1744-1772)
  const char* W = "drprecpc_app_H";
  PlastLaunch L;
  if (!plast_setup(W, nzt, xls, xre, yls, yre, d_i, 0.0f, true, L)) return;
  if (L.kw == 32)
    hipLaunchKernelGGL((drprecpc_app_kernel<4, 32>), dim3(L.nblocks), dim3(64, 4), 0, St, xx, yy,
                       zz, xy, xz, yz, sigma2, yldfac, L);
  else
    hipLaunchKernelGGL((drprecpc_app_kernel<4, 64>), dim3(L.nblocks), dim3(64, 4), 0, St, xx, yy,
                       zz, xy, xz, yz, sigma2, yldfac, L);
  awp_check_launch(W);
}

}  // extern "C"
