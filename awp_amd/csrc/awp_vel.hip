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
  int kb, k0, k1, j0, j1, i0;  // kb: k of lane 0 in tile 0; [k0, k1] active
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
  // one wave per j row: j is wave-uniform, k is the lane
  const int j = L.j0 + jt * BJ + __builtin_amdgcn_readfirstlane((int)threadIdx.y);
  const int k = L.kb + kt * 64 + (int)threadIdx.x;
  if (j > L.j1) return;
  if (k < L.k0 || k > L.k1) return;
  const int ia = L.i0 + ic * L.ichunk;
  int nsteps = L.ni - ic * L.ichunk;
  if (nsteps > L.ichunk) nsteps = L.ichunk;

  const long long Y = L.yline;
  const long long S = L.slice;
  const long long J = (long long)j * Y;
  const unsigned kb = (unsigned)k * 4u;
  const float c1 = L.c1, c2 = L.c2, dth = L.dth;
  const float dy = dcrjy[j];
  const float dz = dcrjz[k];
// row pointer of field A at plane ii, row j+dj, shifted by dk in k (uniform)
#define RP(A, ii, dj, dk) ((A) + (long long)(ii) * S + J + (long long)(dj) * Y + (dk))

  // Prime the x pipelines at i = ia.
  float xx_m2 = awp_ld(RP(xx, ia - 2, 0, 0), kb), xx_m1 = awp_ld(RP(xx, ia - 1, 0, 0), kb),
        xx_0 = awp_ld(RP(xx, ia, 0, 0), kb);
  float xy_m1 = awp_ld(RP(xy, ia - 1, 0, 0), kb), xy_0 = awp_ld(RP(xy, ia, 0, 0), kb),
        xy_p1 = awp_ld(RP(xy, ia + 1, 0, 0), kb);
  float xz_m1 = awp_ld(RP(xz, ia - 1, 0, 0), kb), xz_0 = awp_ld(RP(xz, ia, 0, 0), kb),
        xz_p1 = awp_ld(RP(xz, ia + 1, 0, 0), kb);
  // density at plane i: (j,k), (j-1,k), (j,k-1), (j-1,k-1)
  float dc = awp_ld(RP(d1, ia, 0, 0), kb), djm = awp_ld(RP(d1, ia, -1, 0), kb),
        dkm = awp_ld(RP(d1, ia, 0, -1), kb), djk = awp_ld(RP(d1, ia, -1, -1), kb);

  for (int s = 0; s < nsteps; ++s) {
    // readfirstlane hides the plane recurrence from loop strength reduction,
    // which would otherwise turn every field address into a 64-bit VGPR
    // induction pointer; the plane bases stay in scalar registers.
    const int i = __builtin_amdgcn_readfirstlane(ia + s);
    // pipeline leads
    const float xx_p1 = awp_ld(RP(xx, i + 1, 0, 0), kb);
    const float xy_p2 = awp_ld(RP(xy, i + 2, 0, 0), kb);
    const float xz_p2 = awp_ld(RP(xz, i + 2, 0, 0), kb);
    const float ec = awp_ld(RP(d1, i + 1, 0, 0), kb), ejm = awp_ld(RP(d1, i + 1, -1, 0), kb),
                ekm = awp_ld(RP(d1, i + 1, 0, -1), kb), ejk = awp_ld(RP(d1, i + 1, -1, -1), kb);

    // in-plane stencils
    const float* XY = RP(xy, i, 0, 0);
    const float* XZ = RP(xz, i, 0, 0);
    const float* YY = RP(yy, i, 0, 0);
    const float* YZ = RP(yz, i, 0, 0);
    const float* ZZ = RP(zz, i, 0, 0);
    const float xy_jm1 = awp_ld(XY - Y, kb), xy_jp1 = awp_ld(XY + Y, kb),
                xy_jm2 = awp_ld(XY - 2 * Y, kb);
    const float xz_km1 = awp_ld(XZ - 1, kb), xz_kp1 = awp_ld(XZ + 1, kb), xz_km2 = awp_ld(XZ - 2, kb);
    const float yy_0 = awp_ld(YY, kb), yy_jp1 = awp_ld(YY + Y, kb), yy_jp2 = awp_ld(YY + 2 * Y, kb),
                yy_jm1 = awp_ld(YY - Y, kb);
    const float yz_0 = awp_ld(YZ, kb), yz_km1 = awp_ld(YZ - 1, kb), yz_kp1 = awp_ld(YZ + 1, kb),
                yz_km2 = awp_ld(YZ - 2, kb);
    const float yz_jm1 = awp_ld(YZ - Y, kb), yz_jp1 = awp_ld(YZ + Y, kb),
                yz_jm2 = awp_ld(YZ - 2 * Y, kb);
    const float zz_0 = awp_ld(ZZ, kb), zz_kp1 = awp_ld(ZZ + 1, kb), zz_kp2 = awp_ld(ZZ + 2, kb),
                zz_km1 = awp_ld(ZZ - 1, kb);
    float* U = RP(u1, i, 0, 0);
    float* V = RP(v1, i, 0, 0);
    float* W = RP(w1, i, 0, 0);
    const float u = awp_ld(U, kb), v = awp_ld(V, kb), w = awp_ld(W, kb);

    const float dcrj = dcrjx[i] * dy * dz;
    float f_d1 = 0.25f * (dc + djm + dkm + djk);
    float f_d2 = 0.25f * (dc + ec + dkm + ekm);
    float f_d3 = 0.25f * (dc + ec + djm + ejm);
    f_d1 = dth / f_d1;
    f_d2 = dth / f_d2;
    f_d3 = dth / f_d3;

    awp_st(U, kb, (u + f_d1 * (c1 * (xx_0 - xx_m1) + c2 * (xx_p1 - xx_m2) + c1 * (xy_0 - xy_jm1) +
                               c2 * (xy_jp1 - xy_jm2) + c1 * (xz_0 - xz_km1) +
                               c2 * (xz_kp1 - xz_km2))) * dcrj);
    awp_st(V, kb, (v + f_d2 * (c1 * (xy_p1 - xy_0) + c2 * (xy_p2 - xy_m1) + c1 * (yy_jp1 - yy_0) +
                               c2 * (yy_jp2 - yy_jm1) + c1 * (yz_0 - yz_km1) +
                               c2 * (yz_kp1 - yz_km2))) * dcrj);
    awp_st(W, kb, (w + f_d3 * (c1 * (xz_p1 - xz_0) + c2 * (xz_p2 - xz_m1) + c1 * (yz_0 - yz_jm1) +
                               c2 * (yz_jp1 - yz_jm2) + c1 * (zz_kp1 - zz_0) +
                               c2 * (zz_kp2 - zz_km1))) * dcrj);

    // advance pipelines
    xx_m2 = xx_m1; xx_m1 = xx_0; xx_0 = xx_p1;
    xy_m1 = xy_0; xy_0 = xy_p1; xy_p1 = xy_p2;
    xz_m1 = xz_0; xz_0 = xz_p1; xz_p1 = xz_p2;
    dc = ec; djm = ejm; dkm = ekm; djk = ejk;
  }
#undef RP
}

template <int BJ>
void launch_vel(float* u1, float* v1, float* w1, float* xx, float* yy, float* zz, float* xy,
                float* xz, float* yz, float* dcrjx, float* dcrjy, float* dcrjz, float* d_1,
                hipStream_t St, VelLaunch L) {
  const int nkj = L.nkt * L.njt;
  int nic, ichunk;
  awp_split_i(L.ni, nkj, awp_tuning().vel_ichunk, awp_tuning().vel_target_wgs, &nic, &ichunk);
  L.ichunk = ichunk;
  L.nblocks = nkj * nic;
  hipLaunchKernelGGL((dvelcx_kernel<BJ>), dim3(L.nblocks), dim3(64, BJ), 0, St, u1, v1, w1, xx,
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
  L.kb = awp_tuning().kalign ? (L.k0 & ~31) : L.k0;
  const int bj = awp_tuning().vel_bj;
  L.nkt = awp_cdiv(L.k1 - L.kb + 1, 64);
  L.njt = awp_cdiv(L.j1 - L.j0 + 1, bj);
  if (L.k1 < L.k0) return;
#define AWP_VEL_LAUNCH(B) \
  launch_vel<B>(u1, v1, w1, xx, yy, zz, xy, xz, yz, dcrjx, dcrjy, dcrjz, d_1, St, L);
  switch (bj) {
    case 1: AWP_VEL_LAUNCH(1) break;
    case 2: AWP_VEL_LAUNCH(2) break;
    case 8: AWP_VEL_LAUNCH(8) break;
    case 16: AWP_VEL_LAUNCH(16) break;
    default: AWP_VEL_LAUNCH(4) break;
  }
#undef AWP_VEL_LAUNCH
  awp_check_launch(W);
}
