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
#include "awp_store.h"
#include "awp_tuning.h"
#include "../../include/awp_hip.h"

namespace {

struct VelLaunch {
  int kb, k0, k1, j0, j1, i0;  // kb: k of lane 0 in tile 0; [k0, k1] active
  int nkt, njt, ichunk, ni, nblocks, remap, jfast;
  int yline;
  long long slice;
  float c1, c2, dth;
};

// OOP: the velocities before the step are read from u0/v0/w0 and the
// updated ones written to u1/v1/w1 (distinct arrays); otherwise u1/v1/w1 are
// updated in place and u0/v0/w0 are unused.
// DM2: the grid sits above a coarser one, and the lane at k = align+2 (which
// the 4th-order update skips there) applies dvelc2's 2nd-order update
// (kernel.cu:2289-2356) from the operands it has loaded anyway, so one
// launch does dvelcx_H_opt + dvelc2_H.
// KW: k-run width (awp_pick_kw); with KW = 32 a wave covers two j rows, and
// the second row's offset rides in the per-lane byte offsets.
template <int BJ, bool OOP, class ST, bool DM2 = false, int KW = 64>
__global__ __launch_bounds__(64 * BJ) void dvelcx_kernel(
    typename ST::T* __restrict__ u1, typename ST::T* __restrict__ v1,
    typename ST::T* __restrict__ w1, const typename ST::T* __restrict__ u0,
    const typename ST::T* __restrict__ v0, const typename ST::T* __restrict__ w0,
    const typename ST::T* __restrict__ xx,
    const typename ST::T* __restrict__ yy, const typename ST::T* __restrict__ zz,
    const typename ST::T* __restrict__ xy, const typename ST::T* __restrict__ xz,
    const typename ST::T* __restrict__ yz, const float* __restrict__ dcrjx,
    const float* __restrict__ dcrjy, const float* __restrict__ dcrjz,
    const float* __restrict__ d1, const VelLaunch L, const WaveScales SC) {
  const int b = L.remap ? awp_xcd_remap(blockIdx.x, L.nblocks) : (int)blockIdx.x;
  int kt, jt, ic;
  awp_tile3(b, L.nkt, L.njt, L.jfast, kt, jt, ic);
  // one wave per j row (KW = 64) or per two rows (KW = 32): the wave's first
  // row jw is wave-uniform, k is the lane (mod KW)
  constexpr int RW = 64 / KW;
  const int lane = (int)threadIdx.x;
  const int lr = KW == 64 ? 0 : lane / KW;  // row of this lane in the wave
  const int jw = L.j0 + (jt * BJ + __builtin_amdgcn_readfirstlane((int)threadIdx.y)) * RW;
  const int j = jw + lr;
  const int k = L.kb + kt * KW + (KW == 64 ? lane : lane % KW);
  if (j > L.j1) return;
  const bool dm2 = DM2 && k == AWP_ALIGN + 2;
  if ((k < L.k0 && !dm2) || k > L.k1) return;
  const int ia = L.i0 + ic * L.ichunk;
  int nsteps = L.ni - ic * L.ichunk;
  if (nsteps > L.ichunk) nsteps = L.ichunk;

  const long long Y = L.yline;
  const long long S = L.slice;
  const long long J = (long long)jw * Y;
  const unsigned ko = (unsigned)(lr * Y + k);                // lane offset from row jw
  const unsigned kb = ko * 4u;                               // medium (binary32)
  const unsigned kw = ko * (unsigned)sizeof(typename ST::T);  // wavefield
  const float sxx = SC.s[AWP_SXX], syy = SC.s[AWP_SYY], szz = SC.s[AWP_SZZ];
  const float sxy = SC.s[AWP_SXY], sxz = SC.s[AWP_SXZ], syz = SC.s[AWP_SYZ];
  const float c1 = L.c1, c2 = L.c2, dth = L.dth;
  const float dy = dcrjy[j];
  const float dz = dcrjz[k];
// row pointer of field A at plane ii, row jw+dj, shifted by dk in k (uniform)
#define RP(A, ii, dj, dk) ((A) + (long long)(ii) * S + J + (long long)(dj) * Y + (dk))

  // Prime the x pipelines at i = ia.
  float xx_m2 = ST::ldb(RP(xx, ia - 2, 0, 0), kw, sxx), xx_m1 = ST::ldb(RP(xx, ia - 1, 0, 0), kw, sxx),
        xx_0 = ST::ldb(RP(xx, ia, 0, 0), kw, sxx);
  float xy_m1 = ST::ldb(RP(xy, ia - 1, 0, 0), kw, sxy), xy_0 = ST::ldb(RP(xy, ia, 0, 0), kw, sxy),
        xy_p1 = ST::ldb(RP(xy, ia + 1, 0, 0), kw, sxy);
  float xz_m1 = ST::ldb(RP(xz, ia - 1, 0, 0), kw, sxz), xz_0 = ST::ldb(RP(xz, ia, 0, 0), kw, sxz),
        xz_p1 = ST::ldb(RP(xz, ia + 1, 0, 0), kw, sxz);
  // density at plane i: (j,k), (j-1,k), (j,k-1), (j-1,k-1)
  float dc = awp_ld(RP(d1, ia, 0, 0), kb), djm = awp_ld(RP(d1, ia, -1, 0), kb),
        dkm = awp_ld(RP(d1, ia, 0, -1), kb), djk = awp_ld(RP(d1, ia, -1, -1), kb);

  for (int s = 0; s < nsteps; ++s) {
    // readfirstlane hides the plane recurrence from loop strength reduction,
    // which would otherwise turn every field address into a 64-bit VGPR
    // induction pointer; the plane bases stay in scalar registers.
    const int i = __builtin_amdgcn_readfirstlane(ia + s);
    // pipeline leads
    const float xx_p1 = ST::ldb(RP(xx, i + 1, 0, 0), kw, sxx);
    const float xy_p2 = ST::ldb(RP(xy, i + 2, 0, 0), kw, sxy);
    const float xz_p2 = ST::ldb(RP(xz, i + 2, 0, 0), kw, sxz);
    const float ec = awp_ld(RP(d1, i + 1, 0, 0), kb), ejm = awp_ld(RP(d1, i + 1, -1, 0), kb),
                ekm = awp_ld(RP(d1, i + 1, 0, -1), kb), ejk = awp_ld(RP(d1, i + 1, -1, -1), kb);

    // in-plane stencils
    const typename ST::T* XY = RP(xy, i, 0, 0);
    const typename ST::T* XZ = RP(xz, i, 0, 0);
    const typename ST::T* YY = RP(yy, i, 0, 0);
    const typename ST::T* YZ = RP(yz, i, 0, 0);
    const typename ST::T* ZZ = RP(zz, i, 0, 0);
    const float xy_jm1 = ST::ldb(XY - Y, kw, sxy), xy_jp1 = ST::ldb(XY + Y, kw, sxy),
                xy_jm2 = ST::ldb(XY - 2 * Y, kw, sxy);
    const float xz_km1 = ST::ldb(XZ - 1, kw, sxz), xz_kp1 = ST::ldb(XZ + 1, kw, sxz), xz_km2 = ST::ldb(XZ - 2, kw, sxz);
    const float yy_0 = ST::ldb(YY, kw, syy), yy_jp1 = ST::ldb(YY + Y, kw, syy), yy_jp2 = ST::ldb(YY + 2 * Y, kw, syy),
                yy_jm1 = ST::ldb(YY - Y, kw, syy);
    const float yz_0 = ST::ldb(YZ, kw, syz), yz_km1 = ST::ldb(YZ - 1, kw, syz), yz_kp1 = ST::ldb(YZ + 1, kw, syz),
                yz_km2 = ST::ldb(YZ - 2, kw, syz);
    const float yz_jm1 = ST::ldb(YZ - Y, kw, syz), yz_jp1 = ST::ldb(YZ + Y, kw, syz),
                yz_jm2 = ST::ldb(YZ - 2 * Y, kw, syz);
    const float zz_0 = ST::ldb(ZZ, kw, szz), zz_kp1 = ST::ldb(ZZ + 1, kw, szz), zz_kp2 = ST::ldb(ZZ + 2, kw, szz),
                zz_km1 = ST::ldb(ZZ - 1, kw, szz);
    typename ST::T* U = RP(u1, i, 0, 0);
    typename ST::T* V = RP(v1, i, 0, 0);
    typename ST::T* W = RP(w1, i, 0, 0);
    const float u = ST::ldb(OOP ? RP(u0, i, 0, 0) : U, kw, SC.s[AWP_SU]),
                v = ST::ldb(OOP ? RP(v0, i, 0, 0) : V, kw, SC.s[AWP_SV]),
                w = ST::ldb(OOP ? RP(w0, i, 0, 0) : W, kw, SC.s[AWP_SW]);

    const float dcrj = dcrjx[i] * dy * dz;
    float f_d1 = 0.25f * (dc + djm + dkm + djk);
    float f_d2 = 0.25f * (dc + ec + dkm + ekm);
    float f_d3 = 0.25f * (dc + ec + djm + ejm);
    f_d1 = dth / f_d1;
    f_d2 = dth / f_d2;
    f_d3 = dth / f_d3;

    float un = (u + f_d1 * (c1 * (xx_0 - xx_m1) + c2 * (xx_p1 - xx_m2) + c1 * (xy_0 - xy_jm1) +
                            c2 * (xy_jp1 - xy_jm2) + c1 * (xz_0 - xz_km1) +
                            c2 * (xz_kp1 - xz_km2))) * dcrj;
    float vn = (v + f_d2 * (c1 * (xy_p1 - xy_0) + c2 * (xy_p2 - xy_m1) + c1 * (yy_jp1 - yy_0) +
                            c2 * (yy_jp2 - yy_jm1) + c1 * (yz_0 - yz_km1) +
                            c2 * (yz_kp1 - yz_km2))) * dcrj;
    float wn = (w + f_d3 * (c1 * (xz_p1 - xz_0) + c2 * (xz_p2 - xz_m1) + c1 * (yz_0 - yz_jm1) +
                            c2 * (yz_jp1 - yz_jm2) + c1 * (zz_kp1 - zz_0) +
                            c2 * (zz_kp2 - zz_km1))) * dcrj;
    if (DM2) {
      // dvelc2 on its level, computed on every lane and selected, so the
      // loop keeps one path.  Its density means are the same float sums
      // scaled by the double 0.25; a power-of-two scale is exact in both
      // precisions, so its dth/mean equal f_d1..3 bit for bit.  Its sponge
      // factor multiplies in the order dcrjz * dcrjx * dcrjy.
      const float f_dcrj = dz * dcrjx[i] * dy;
      const float u2 = (u + f_d1 * ((xx_0 - xx_m1) + (xy_0 - xy_jm1) + (xz_0 - xz_km1))) * f_dcrj;
      const float v2 = (v + f_d2 * ((xy_p1 - xy_0) + (yy_jp1 - yy_0) + (yz_0 - yz_km1))) * f_dcrj;
      const float w2 = (w + f_d3 * ((xz_p1 - xz_0) + (yz_0 - yz_jm1) + (zz_kp1 - zz_0))) * f_dcrj;
      un = dm2 ? u2 : un;
      vn = dm2 ? v2 : vn;
      wn = dm2 ? w2 : wn;
    }
    ST::stb(U, kw, un, SC.inv[AWP_SU]);
    ST::stb(V, kw, vn, SC.inv[AWP_SV]);
    ST::stb(W, kw, wn, SC.inv[AWP_SW]);

    // advance pipelines
    xx_m2 = xx_m1; xx_m1 = xx_0; xx_0 = xx_p1;
    xy_m1 = xy_0; xy_0 = xy_p1; xy_p1 = xy_p2;
    xz_m1 = xz_0; xz_0 = xz_p1; xz_p1 = xz_p2;
    dc = ec; djm = ejm; dkm = ekm; djk = ejk;
  }
#undef RP
}

template <int BJ, class ST, bool DM2 = false, int KW = 64>
void launch_vel(void* const* f9, void* const* v0, float* dcrjx, float* dcrjy, float* dcrjz,
                float* d_1, hipStream_t St, VelLaunch L, const WaveScales& SC) {
  typedef typename ST::T T;
  const int nkj = L.nkt * L.njt;
  int nic, ichunk;
  awp_split_i(L.ni, nkj, awp_tuning().vel_ichunk, awp_tuning().vel_target_wgs, &nic, &ichunk);
  L.ichunk = ichunk;
  L.nblocks = nkj * nic;
  if (DM2)
    hipLaunchKernelGGL((dvelcx_kernel<BJ, false, ST, true, KW>), dim3(L.nblocks), dim3(64, BJ), 0,
                       St, (T*)f9[0], (T*)f9[1], (T*)f9[2], nullptr, nullptr, nullptr,
                       (const T*)f9[3], (const T*)f9[4], (const T*)f9[5], (const T*)f9[6],
                       (const T*)f9[7], (const T*)f9[8], dcrjx, dcrjy, dcrjz, d_1, L, SC);
  else if (v0)
    hipLaunchKernelGGL((dvelcx_kernel<BJ, true, ST, false, KW>), dim3(L.nblocks), dim3(64, BJ), 0, St,
                       (T*)f9[0], (T*)f9[1], (T*)f9[2], (const T*)v0[0], (const T*)v0[1],
                       (const T*)v0[2], (const T*)f9[3], (const T*)f9[4], (const T*)f9[5],
                       (const T*)f9[6], (const T*)f9[7], (const T*)f9[8], dcrjx, dcrjy, dcrjz,
                       d_1, L, SC);
  else
    hipLaunchKernelGGL((dvelcx_kernel<BJ, false, ST, false, KW>), dim3(L.nblocks), dim3(64, BJ), 0, St,
                       (T*)f9[0], (T*)f9[1], (T*)f9[2], nullptr, nullptr, nullptr,
                       (const T*)f9[3], (const T*)f9[4], (const T*)f9[5], (const T*)f9[6],
                       (const T*)f9[7], (const T*)f9[8], dcrjx, dcrjy, dcrjz, d_1, L, SC);
}

// f9: u1 v1 w1 xx yy zz xy xz yz in the storage type of ST; v0: null (u1 v1
// w1 updated in place) or the velocities before the step (u0 v0 w0).
template <class ST, bool DM2 = false>
void dvelcx_common(const char* W, void* const* f9, void* const* v0, float* dcrjx, float* dcrjy,
                   float* dcrjz, float* d_1, int nyt, int nzt, hipStream_t St, int s_i, int e_i,
                   int d_i, int ngrids, const WaveScales& SC) {
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
  L.jfast = awp_tuning().vel_jfast;
  L.kb = awp_tuning().kalign ? (L.k0 & ~31) : L.k0;
  const int bj = awp_tuning().vel_bj;
  const int kw = awp_pick_kw(L.k1 - L.kb + 1, awp_tuning().kw);
  L.nkt = awp_cdiv(L.k1 - L.kb + 1, kw);
  L.njt = awp_cdiv(L.j1 - L.j0 + 1, bj * (64 / kw));
  if (L.k1 < L.k0) return;
#define AWP_VEL_LAUNCH(B, KW) \
  launch_vel<B, ST, DM2, KW>(f9, v0, dcrjx, dcrjy, dcrjz, d_1, St, L, SC)
  if (kw == 32) {
    switch (bj) {
      case 1: AWP_VEL_LAUNCH(1, 32); break;
      case 2: AWP_VEL_LAUNCH(2, 32); break;
      case 8: AWP_VEL_LAUNCH(8, 32); break;
      case 16: AWP_VEL_LAUNCH(16, 32); break;
      default: AWP_VEL_LAUNCH(4, 32); break;
    }
  } else {
    switch (bj) {
      case 1: AWP_VEL_LAUNCH(1, 64); break;
      case 2: AWP_VEL_LAUNCH(2, 64); break;
      case 8: AWP_VEL_LAUNCH(8, 64); break;
      case 16: AWP_VEL_LAUNCH(16, 64); break;
      default: AWP_VEL_LAUNCH(4, 64); break;
    }
  }
#undef AWP_VEL_LAUNCH
  awp_check_launch(W);
}

}  // namespace

extern "C" void dvelcx_H_opt(float* u1, float* v1, float* w1, float* xx, float* yy, float* zz,
                             float* xy, float* xz, float* yz, float* dcrjx, float* dcrjy,
                             float* dcrjz, float* d_1, int nyt, int nzt, hipStream_t St,
                             int s_i, int e_i, int d_i, int ngrids) {
  void* f9[9] = {u1, v1, w1, xx, yy, zz, xy, xz, yz};
  WaveScales SC = {};
  dvelcx_common<StF32>("dvelcx_H_opt", f9, nullptr, dcrjx, dcrjy, dcrjz, d_1, nyt, nzt, St,
                       s_i, e_i, d_i, ngrids, SC);
}

extern "C" void awp_dvelcx_dm(float* u1, float* v1, float* w1, float* xx, float* yy,
                              float* zz, float* xy, float* xz, float* yz, float* dcrjx,
                              float* dcrjy, float* dcrjz, float* d_1, int nyt, int nzt,
                              hipStream_t St, int s_i, int e_i, int d_i, int ngrids) {
  const char* W = "awp_dvelcx_dm";
  void* f9[9] = {u1, v1, w1, xx, yy, zz, xy, xz, yz};
  WaveScales SC = {};
  if (d_i >= ngrids - 1) {
    // the lowest grid has no dvelc2 level
    dvelcx_common<StF32>(W, f9, nullptr, dcrjx, dcrjy, dcrjz, d_1, nyt, nzt, St, s_i, e_i, d_i,
                         ngrids, SC);
    return;
  }
  if (!awp_grid_ok(d_i, W)) return;
  const AwpGrid& g = awp_state().grid[d_i];
  // dvelc2_H covers the whole interior (kernel.cu:2301): so must the sweep
  if (s_i != 2 + AWP_NGSL || e_i != g.nxt + AWP_NGSL + 1) {
    awp_set_error(AWP_ERR_ARG, "%s: i range [%d, %d] is not grid %d's interior [%d, %d]", W, s_i,
                  e_i, d_i, 2 + AWP_NGSL, g.nxt + AWP_NGSL + 1);
    return;
  }
  if (g.nzt < 3) {
    awp_set_error(AWP_ERR_ARG, "%s: grid %d has nzt=%d < 3", W, d_i, g.nzt);
    return;
  }
  if (!awp_tuning().kalign) {
    awp_set_error(AWP_ERR_STATE, "%s: needs k tiles that start at k = align (kalign=1)", W);
    return;
  }
  dvelcx_common<StF32, true>(W, f9, nullptr, dcrjx, dcrjy, dcrjz, d_1, nyt, nzt, St, s_i, e_i,
                             d_i, ngrids, SC);
}

extern "C" void awp_dvelcx_s16(uint16_t* u1, uint16_t* v1, uint16_t* w1, uint16_t* xx,
                               uint16_t* yy, uint16_t* zz, uint16_t* xy, uint16_t* xz,
                               uint16_t* yz, float* dcrjx, float* dcrjy, float* dcrjz,
                               float* d_1, int nyt, int nzt, hipStream_t St, int s_i, int e_i,
                               int d_i, int ngrids, int fmt, const float* scale) {
  const char* W = "awp_dvelcx_s16";
  WaveScales SC;
  if (!awp_scales_from_host(W, fmt, scale, &SC)) return;
  void* f9[9] = {u1, v1, w1, xx, yy, zz, xy, xz, yz};
  if (fmt == AWP_FMT_BF16)
    dvelcx_common<StBF16>(W, f9, nullptr, dcrjx, dcrjy, dcrjz, d_1, nyt, nzt, St, s_i, e_i, d_i,
                          ngrids, SC);
  else
    dvelcx_common<StF16>(W, f9, nullptr, dcrjx, dcrjy, dcrjz, d_1, nyt, nzt, St, s_i, e_i, d_i,
                         ngrids, SC);
}

extern "C" void awp_dvelcx_oop(float* u1, float* v1, float* w1, const float* u0,
                               const float* v0, const float* w0, float* xx, float* yy,
                               float* zz, float* xy, float* xz, float* yz, float* dcrjx,
                               float* dcrjy, float* dcrjz, float* d_1, int nyt, int nzt,
                               hipStream_t St, int s_i, int e_i, int d_i, int ngrids) {
  const char* W = "awp_dvelcx_oop";
  if (u1 == u0 || v1 == v0 || w1 == w0) {
    awp_set_error(AWP_ERR_ARG, "%s: the velocities before and after the step must be distinct "
                  "arrays (dvelcx_H_opt updates in place)", W);
    return;
  }
  void* f9[9] = {u1, v1, w1, xx, yy, zz, xy, xz, yz};
  void* vin[3] = {(void*)u0, (void*)v0, (void*)w0};
  WaveScales SC = {};
  dvelcx_common<StF32>(W, f9, vin, dcrjx, dcrjy, dcrjz, d_1, nyt, nzt, St, s_i, e_i, d_i, ngrids,
                       SC);
}
