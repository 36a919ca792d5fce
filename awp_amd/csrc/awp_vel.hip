// Velocity update (reference dvelcx_opt, src/awp/kernel.cu:125-268), gfx950.
//
// Mapping: one wave64 covers 64 consecutive k (the contiguous axis, 256 B per
// field load), a workgroup stacks BJ waves along j, and each lane marches
// along i over one i-chunk with register pipelines for every x-stencil and
// for the 4-point density means, so each field plane is fetched from HBM
// once per march.  j/k neighbours come from L1/L2; the XCD-aware tile order
// keeps neighbouring tiles (which share those lines) on one L2.
//
// Arithmetic follows the reference expression trees (built with
// -ffp-contract=off), so results are reproducible bit for bit on any
// IEEE-754 binary32 implementation of the same operations.
#include "awp_common.h"
#include "awp_tuning.h"
#include "../../include/awp_hip.h"

namespace {

struct VelLaunch {
  int k0, k1, j0, j1, i0;
  int nkt, njt, ichunk, ni, nblocks, remap;
  int yline;
  long long slice;
  float c1, c2, dth;
};

template <int BJ>
__global__ __launch_bounds__(64 * BJ) void dvelcx_kernel(
    float* __restrict__ u1, float* __restrict__ v1, float* __restrict__ w1,
    const float* __restrict__ xx, const float* __restrict__ yy, const float* __restrict__ zz,
    const float* __restrict__ xy, const float* __restrict__ xz, const float* __restrict__ yz,
    const float* __restrict__ dcrjx, const float* __restrict__ dcrjy,
    const float* __restrict__ dcrjz, const float* __restrict__ d1, const VelLaunch L) {
  const int b = L.remap ? awp_xcd_remap(blockIdx.x, L.nblocks) : (int)blockIdx.x;
  const int kt = b % L.nkt;
  const int jt = (b / L.nkt) % L.njt;
  const int ic = b / (L.nkt * L.njt);
  const int k = L.k0 + kt * 64 + (int)threadIdx.x;
  const int j = L.j0 + jt * BJ + (int)threadIdx.y;
  if (k > L.k1 || j > L.j1) return;
  const int ia = L.i0 + ic * L.ichunk;
  int nsteps = L.ni - ic * L.ichunk;
  if (nsteps > L.ichunk) nsteps = L.ichunk;

  const int Y = L.yline;
  const long long S = L.slice;
  const int o = j * Y + k;  // in-plane offset
  const float c1 = L.c1, c2 = L.c2, dth = L.dth;
  const float dyz = dcrjy[j];
  const float dz = dcrjz[k];

  // Prime the x pipelines at i = ia.
  const long long p0 = (long long)ia * S + o;
  float xx_m2 = xx[p0 - 2 * S], xx_m1 = xx[p0 - S], xx_0 = xx[p0];
  float xy_m1 = xy[p0 - S], xy_0 = xy[p0], xy_p1 = xy[p0 + S];
  float xz_m1 = xz[p0 - S], xz_0 = xz[p0], xz_p1 = xz[p0 + S];
  // density at plane i: (j,k), (j-1,k), (j,k-1), (j-1,k-1)
  float dc = d1[p0], djm = d1[p0 - Y], dkm = d1[p0 - 1], djk = d1[p0 - Y - 1];

  for (int s = 0; s < nsteps; ++s) {
    const int i = ia + s;
    const float* xxp = xx + (long long)i * S;
    const float* xyp = xy + (long long)i * S;
    const float* xzp = xz + (long long)i * S;
    const float* yyp = yy + (long long)i * S;
    const float* yzp = yz + (long long)i * S;
    const float* zzp = zz + (long long)i * S;
    const float* d1n = d1 + (long long)(i + 1) * S;
    float* up = u1 + (long long)i * S;
    float* vp = v1 + (long long)i * S;
    float* wp = w1 + (long long)i * S;

    // pipeline leads
    const float xx_p1 = xxp[S + o];
    const float xy_p2 = xyp[2 * S + o];
    const float xz_p2 = xzp[2 * S + o];
    const float ec = d1n[o], ejm = d1n[o - Y], ekm = d1n[o - 1], ejk = d1n[o - Y - 1];

    // in-plane stencils
    const float xy_jm1 = xyp[o - Y], xy_jp1 = xyp[o + Y], xy_jm2 = xyp[o - 2 * Y];
    const float xz_km1 = xzp[o - 1], xz_kp1 = xzp[o + 1], xz_km2 = xzp[o - 2];
    const float yy_0 = yyp[o], yy_jp1 = yyp[o + Y], yy_jp2 = yyp[o + 2 * Y], yy_jm1 = yyp[o - Y];
    const float yz_0 = yzp[o], yz_km1 = yzp[o - 1], yz_kp1 = yzp[o + 1], yz_km2 = yzp[o - 2];
    const float yz_jm1 = yzp[o - Y], yz_jp1 = yzp[o + Y], yz_jm2 = yzp[o - 2 * Y];
    const float zz_0 = zzp[o], zz_kp1 = zzp[o + 1], zz_kp2 = zzp[o + 2], zz_km1 = zzp[o - 1];
    const float u = up[o], v = vp[o], w = wp[o];

    const float dcrj = dcrjx[i] * dyz * dz;
    float f_d1 = 0.25f * (dc + djm + dkm + djk);
    float f_d2 = 0.25f * (dc + ec + dkm + ekm);
    float f_d3 = 0.25f * (dc + ec + djm + ejm);
    f_d1 = dth / f_d1;
    f_d2 = dth / f_d2;
    f_d3 = dth / f_d3;

    up[o] = (u + f_d1 * (c1 * (xx_0 - xx_m1) + c2 * (xx_p1 - xx_m2) + c1 * (xy_0 - xy_jm1) +
                         c2 * (xy_jp1 - xy_jm2) + c1 * (xz_0 - xz_km1) + c2 * (xz_kp1 - xz_km2))) *
            dcrj;
    vp[o] = (v + f_d2 * (c1 * (xy_p1 - xy_0) + c2 * (xy_p2 - xy_m1) + c1 * (yy_jp1 - yy_0) +
                         c2 * (yy_jp2 - yy_jm1) + c1 * (yz_0 - yz_km1) + c2 * (yz_kp1 - yz_km2))) *
            dcrj;
    wp[o] = (w + f_d3 * (c1 * (xz_p1 - xz_0) + c2 * (xz_p2 - xz_m1) + c1 * (yz_0 - yz_jm1) +
                         c2 * (yz_jp1 - yz_jm2) + c1 * (zz_kp1 - zz_0) + c2 * (zz_kp2 - zz_km1))) *
            dcrj;

    // advance pipelines
    xx_m2 = xx_m1; xx_m1 = xx_0; xx_0 = xx_p1;
    xy_m1 = xy_0; xy_0 = xy_p1; xy_p1 = xy_p2;
    xz_m1 = xz_0; xz_0 = xz_p1; xz_p1 = xz_p2;
    dc = ec; djm = ejm; dkm = ekm; djk = ejk;
  }
}

template <int BJ>
void launch_vel(float* u1, float* v1, float* w1, float* xx, float* yy, float* zz, float* xy,
                float* xz, float* yz, float* dcrjx, float* dcrjy, float* dcrjz, float* d_1,
                hipStream_t St, VelLaunch L) {
  const int nkj = L.nkt * L.njt;
  int nic, ichunk;
  awp_split_i(L.ni, nkj, awp_tuning().vel_ichunk, awp_tuning().target_wgs, &nic, &ichunk);
  L.ichunk = ichunk;
  L.nblocks = nkj * nic;
  hipLaunchKernelGGL(dvelcx_kernel<BJ>, dim3(L.nblocks), dim3(64, BJ), 0, St, u1, v1, w1, xx,
                     yy, zz, xy, xz, yz, dcrjx, dcrjy, dcrjz, d_1, L);
}

}  // namespace

extern "C" void dvelcx_H_opt(float* u1, float* v1, float* w1, float* xx, float* yy, float* zz,
                             float* xy, float* xz, float* yz, float* dcrjx, float* dcrjy,
                             float* dcrjz, float* d_1, int nyt, int nzt, hipStream_t St,
                             int s_i, int e_i, int d_i, int ngrids) {
  const char* W = "dvelcx_H_opt";
  if (!awp_grid_ok(d_i, W)) return;
  const AwpGrid& g = awp_state().grid[d_i];
  if (nyt != g.nyt || nzt != g.nzt) {
    awp_set_error(AWP_ERR_ARG, "%s: nyt=%d nzt=%d disagree with grid %d (%d, %d)", W, nyt, nzt,
                  d_i, g.nyt, g.nzt);
    return;
  }
  // The x stencil reaches i-2 .. i+2; the padded array holds i in [0, nxt+20).
  if (s_i > e_i) return;
  if (s_i < 2 || e_i + 2 > g.nxt + 3 + AWP_NGSL2) {
    awp_set_error(AWP_ERR_ARG, "%s: i range [%d, %d] leaves the padded grid (nxt=%d)", W, s_i,
                  e_i, g.nxt);
    return;
  }
  VelLaunch L;
  L.k0 = AWP_ALIGN;
  L.k1 = nzt + AWP_ALIGN - 1;
  if (d_i > 0) L.k1 = nzt + AWP_ALIGN - 3;           // kernel.cu:146
  if (d_i < ngrids - 1) L.k0 = AWP_ALIGN + 3;        // kernel.cu:147
  L.j0 = 2 + AWP_NGSL;
  L.j1 = nyt + AWP_NGSL + 1;
  L.i0 = s_i;
  L.ni = e_i - s_i + 1;
  L.yline = g.yline;
  L.slice = g.slice;
  L.c1 = g.c1;
  L.c2 = g.c2;
  L.dth = g.dth;
  L.remap = awp_tuning().xcd_remap;
  const int bj = awp_tuning().vel_bj;
  L.nkt = awp_cdiv(L.k1 - L.k0 + 1, 64);
  L.njt = awp_cdiv(L.j1 - L.j0 + 1, bj);
  if (L.k1 < L.k0) return;
  switch (bj) {
    case 1: launch_vel<1>(u1, v1, w1, xx, yy, zz, xy, xz, yz, dcrjx, dcrjy, dcrjz, d_1, St, L); break;
    case 2: launch_vel<2>(u1, v1, w1, xx, yy, zz, xy, xz, yz, dcrjx, dcrjy, dcrjz, d_1, St, L); break;
    case 8: launch_vel<8>(u1, v1, w1, xx, yy, zz, xy, xz, yz, dcrjx, dcrjy, dcrjz, d_1, St, L); break;
    default: launch_vel<4>(u1, v1, w1, xx, yy, zz, xy, xz, yz, dcrjx, dcrjy, dcrjz, d_1, St, L); break;
  }
  awp_check_launch(W);
}
