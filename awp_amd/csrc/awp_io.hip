// Point-source injection (reference addsrc_cu, src/awp/kernel.cu:1780-1806),
// sub-volume output sampling (velbuffer, :3927-4011), receiver gathers and
// the launch-shape knobs.
#include <string.h>

#include "awp_common.h"
#include "awp_store.h"
#include "awp_tuning.h"
#include "../../include/awp_hip.h"

// Defaults from sweeps on MI355X at 512x512x256: velocity 4 rows x 2 i-chunks
// (all workgroups resident), stress 16 rows x ~32-plane i-chunks; the
// curvilinear velocity update as three launches (0.985 against 1.043 ms for
// one launch at 384x384x192), its strain-rate pass as three (normal rates,
// xy, xz and yz: 2.056 -> 2.006 -> 1.948 ms per stress update for one, two,
// three launches), the velocity waves ordering their LDS at wave scope
// (1.002 -> 0.971 ms); the adjoint kernels' blocks in groups of 8 j tiles over
// all planes (adj_stress_local 551 -> 533 us at 256x256x128).
static AwpTuning g_tuning = {4, 16, 1024, 2048, 0, 0, 1, 1, 8, 16, 1, 32, 2048, 0, 0, 0, 1, 8, 2, 0, 8, 1, 0};
AwpTuning& awp_tuning() { return g_tuning; }

namespace {

struct SrcLaunch {
  int sample, read_step, dim, npsrc, yline;
  long long slice;
  float vtst;
};

// One lane per source, as the reference.
template <class ST>
__global__ __launch_bounds__(256) void addsrc_kernel(
    const int* __restrict__ psrc, const float* __restrict__ axx, const float* __restrict__ ayy,
    const float* __restrict__ azz, const float* __restrict__ axz, const float* __restrict__ ayz,
    const float* __restrict__ axy, typename ST::T* xx, typename ST::T* yy, typename ST::T* zz,
    typename ST::T* xy, typename ST::T* yz, typename ST::T* xz, const SrcLaunch A,
    const WaveScales SC) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= A.npsrc) return;
  const int idx = psrc[j * A.dim] + 1 + AWP_NGSL;
  const int idy = psrc[j * A.dim + 1] + 1 + AWP_NGSL;
  const int idz = psrc[j * A.dim + 2] + AWP_ALIGN - 1;
  const long long p = (long long)idx * A.slice + (long long)idy * A.yline + idz;
  const long long s = (long long)j * A.read_step + A.sample;
  const float v = A.vtst;
  ST::st(xx, p, ST::ld(xx, p, SC.s[AWP_SXX]) - v * axx[s], SC.inv[AWP_SXX]);
  ST::st(yy, p, ST::ld(yy, p, SC.s[AWP_SYY]) - v * ayy[s], SC.inv[AWP_SYY]);
  ST::st(zz, p, ST::ld(zz, p, SC.s[AWP_SZZ]) - v * azz[s], SC.inv[AWP_SZZ]);
  ST::st(xz, p, ST::ld(xz, p, SC.s[AWP_SXZ]) - v * axz[s], SC.inv[AWP_SXZ]);
  ST::st(yz, p, ST::ld(yz, p, SC.s[AWP_SYZ]) - v * ayz[s], SC.inv[AWP_SYZ]);
  ST::st(xy, p, ST::ld(xy, p, SC.s[AWP_SXY]) - v * axy[s], SC.inv[AWP_SXY]);
}

struct VbLaunch {
  int NVE, nbgx, nedx, nskpx, nbgy, nedy, nskpy, nbgz, nedz, nskpz, rec_nxt, rec_nyt, rec_nzt, follow;
  int d_i, nzt, yline, bathy_ld;
  long long slice;
};

__global__ __launch_bounds__(256) void velbuffer_kernel(
    const float* __restrict__ u1, const float* __restrict__ v1, const float* __restrict__ w1,
    const float* __restrict__ neta, float* __restrict__ Bufx, float* __restrict__ Bufy,
    float* __restrict__ Bufz, float* __restrict__ Bufeta, const int* __restrict__ bathy,
    const VbLaunch V) {
  const int ti = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int tj = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  const int tk = (int)(blockIdx.z * blockDim.z + threadIdx.z);
  if (ti >= V.rec_nxt || tj >= V.rec_nyt || tk >= V.rec_nzt) return;
  const int i = 2 + AWP_NGSL + V.nbgx + ti * V.nskpx;
  const int j = 2 + AWP_NGSL + V.nbgy + tj * V.nskpy;
  const int k = V.nbgz + tk * V.nskpz;
  if (i > 2 + AWP_NGSL + V.nedx || j > 2 + AWP_NGSL + V.nedy || k > V.nedz) return;
  int ko, koz;
  if (V.d_i > 0) {
    ko = V.nzt + AWP_ALIGN - 1 - k;
    koz = ko;
  } else if (V.follow == 0) {
    ko = V.nzt + AWP_ALIGN - 1 - k;
    koz = ko - 1;
  } else {
    const int b = bathy[(long long)j * V.bathy_ld + i];
    ko = b - k;
    koz = (b == V.nzt + AWP_ALIGN - 1) ? ko - 1 : ko;
  }
  if (ko < 0 || koz < 0 || ko >= V.yline || koz >= V.yline) {
    // a bathymetry level k below the surface index that leaves the column:
    // no read outside the arrays, the sample reads NaN
    const long long t = (long long)tk * V.rec_nxt * V.rec_nyt + (long long)tj * V.rec_nxt + ti;
    Bufx[t] = Bufy[t] = Bufz[t] = __builtin_nanf("");
    if (V.NVE == 3) Bufeta[t] = __builtin_nanf("");
    return;
  }
  const long long base = (long long)i * V.slice + (long long)j * V.yline;
  const long long t = (long long)((k - V.nbgz) / V.nskpz) * V.rec_nxt * V.rec_nyt +
                      (long long)((j - 2 - AWP_NGSL - V.nbgy) / V.nskpy) * V.rec_nxt +
                      (i - 2 - AWP_NGSL - V.nbgx) / V.nskpx;
  Bufx[t] = u1[base + ko];
  Bufy[t] = v1[base + ko];
  Bufz[t] = w1[base + koz];
  if (V.NVE == 3) Bufeta[t] = neta[base + ko];
}

struct RecLaunch {
  int npts, slot, zshift, yline;
  long long slice;
};

template <class ST>
__global__ __launch_bounds__(256) void receivers_kernel(const typename ST::T* __restrict__ u1,
                                                        const typename ST::T* __restrict__ v1,
                                                        const typename ST::T* __restrict__ w1,
                                                        const int* __restrict__ idx,
                                                        float* __restrict__ out,
                                                        const RecLaunch R, const WaveScales SC) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= R.npts) return;
  const long long p =
      (long long)idx[3 * n] * R.slice + (long long)idx[3 * n + 1] * R.yline + idx[3 * n + 2];
  float* o = out + (long long)R.slot * 3 * R.npts + 3 * n;
  o[0] = ST::ld(u1, p, SC.s[AWP_SU]);
  o[1] = ST::ld(v1, p, SC.s[AWP_SV]);
  o[2] = ST::ld(w1, p - R.zshift, SC.s[AWP_SW]);
}

template <class ST>
void addsrc_common(const char* W, int i, int READ_STEP, int dim, int* psrc, int npsrc,
                   hipStream_t St, float* axx, float* ayy, float* azz, float* axz, float* ayz,
                   float* axy, void* xx, void* yy, void* zz, void* xy, void* yz, void* xz,
                   int d_i, const WaveScales& SC) {
  typedef typename ST::T T;
  if (!awp_grid_ok(d_i, W)) return;
  if (npsrc <= 0) return;
  if (i < 1 || i > READ_STEP || dim < 3) {
    awp_set_error(AWP_ERR_ARG, "%s: sample %d outside [1, READ_STEP=%d] or dim=%d < 3", W, i,
                  READ_STEP, dim);
    return;
  }
  const AwpGrid& g = awp_state().grid[d_i];
  SrcLaunch A;
  A.sample = i - 1;
  A.read_step = READ_STEP;
  A.dim = dim;
  A.npsrc = npsrc;
  A.yline = g.yline;
  A.slice = g.slice;
  A.vtst = g.DT / (g.DH * g.DH * g.DH);  // kernel.cu:1790, (_prec)d_DT/(DH*DH*DH)
  const int block = npsrc < 256 ? npsrc : 256;
  hipLaunchKernelGGL((addsrc_kernel<ST>), dim3(awp_cdiv(npsrc, block)), dim3(block), 0, St, psrc,
                     axx, ayy, azz, axz, ayz, axy, (T*)xx, (T*)yy, (T*)zz, (T*)xy, (T*)yz,
                     (T*)xz, A, SC);
  awp_check_launch(W);
}

template <class ST>
void receivers_common(const char* W, const void* u1, const void* v1, const void* w1,
                      const int* idx, int npts, float* out, int step_slot, hipStream_t St,
                      int d_i, const WaveScales& SC) {
  typedef typename ST::T T;
  if (!awp_grid_ok(d_i, W)) return;
  if (npts <= 0) return;
  if (step_slot < 0) {
    awp_set_error(AWP_ERR_ARG, "%s: negative step_slot %d", W, step_slot);
    return;
  }
  const AwpGrid& g = awp_state().grid[d_i];
  RecLaunch R;
  R.npts = npts;
  R.slot = step_slot;
  R.zshift = (d_i == 0) ? 1 : 0;
  R.yline = g.yline;
  R.slice = g.slice;
  hipLaunchKernelGGL((receivers_kernel<ST>), dim3(awp_cdiv(npts, 256)), dim3(256), 0, St,
                     (const T*)u1, (const T*)v1, (const T*)w1, idx, out, R, SC);
  awp_check_launch(W);
}

}  // namespace

extern "C" {

// Launch-shape knobs (no reference counterpart; the reference fixes its
// block shapes at compile time, include/awp/pmcl3d_cons.h:2-4).
int awp_set_tuning(const char* key, int value) {
  AwpTuning& t = g_tuning;
  if (!strcmp(key, "vel_bj") || !strcmp(key, "str_bj")) {
    if (value != 1 && value != 2 && value != 4 && value != 8 && value != 16) return -1;
    (!strcmp(key, "vel_bj") ? t.vel_bj : t.str_bj) = value;
  } else if (!strcmp(key, "vel_target_wgs") || !strcmp(key, "str_target_wgs")) {
    if (value < 1) return -1;
    (!strcmp(key, "vel_target_wgs") ? t.vel_target_wgs : t.str_target_wgs) = value;
  } else if (!strcmp(key, "vel_ichunk")) {
    t.vel_ichunk = value < 0 ? 0 : value;
  } else if (!strcmp(key, "str_ichunk")) {
    t.str_ichunk = value < 0 ? 0 : value;
  } else if (!strcmp(key, "xcd_remap")) {
    t.xcd_remap = value ? 1 : 0;
  } else if (!strcmp(key, "kalign")) {
    t.kalign = value ? 1 : 0;
  } else if (!strcmp(key, "topo_vel_ichunk") || !strcmp(key, "topo_str_ichunk")) {
    if (value < 1 || value > 4096) return -1;
    (!strcmp(key, "topo_vel_ichunk") ? t.topo_vel_ichunk : t.topo_str_ichunk) = value;
  } else if (!strcmp(key, "topo_q_ichunk")) {
    if (value < 1 || value > 4096) return -1;
    t.topo_q_ichunk = value;
  } else if (!strcmp(key, "step_ichunk")) {
    t.step_ichunk = value < 0 ? 0 : value;
  } else if (!strcmp(key, "vel_jfast")) {
    t.vel_jfast = value ? 1 : 0;
  } else if (!strcmp(key, "str_jfast")) {
    t.str_jfast = value ? 1 : 0;
  } else if (!strcmp(key, "step_jfast")) {
    t.step_jfast = value ? 1 : 0;
  } else if (!strcmp(key, "topo_vel_split")) {
    t.topo_vel_split = value ? 1 : 0;
  } else if (!strcmp(key, "topo_str_split")) {
    if (value < 0 || value > 2) return -1;
    t.topo_str_split = value;
  } else if (!strcmp(key, "topo_rj")) {
    if (value < 0 || value > 8) return -1;
    t.topo_rj = value;
  } else if (!strcmp(key, "topo_wsync")) {
    t.topo_wsync = value ? 1 : 0;
  } else if (!strcmp(key, "topo_nseg")) {
    if (value < 1 || value > 8) return -1;
    t.topo_nseg = value;
  } else if (!strcmp(key, "kw")) {
    if (value != 0 && value != 32 && value != 64) return -1;
    t.kw = value;
  } else if (!strcmp(key, "adj_jgroup")) {
    if (value < 0) return -1;
    t.adj_jgroup = value;
  } else if (!strcmp(key, "step_target_wgs")) {
    if (value < 1) return -1;
    t.step_target_wgs = value;
  } else {
    return -1;
  }
  return 0;
}

void addsrc_H(int i, int READ_STEP, int dim, int* psrc, int npsrc, hipStream_t St,
              float* axx, float* ayy, float* azz, float* axz, float* ayz, float* axy,
              float* xx, float* yy, float* zz, float* xy, float* yz, float* xz, int d_i) {
  WaveScales SC = {};
  addsrc_common<StF32>("addsrc_H", i, READ_STEP, dim, psrc, npsrc, St, axx, ayy, azz, axz, ayz,
                       axy, xx, yy, zz, xy, yz, xz, d_i, SC);
}

void awp_addsrc_s16(int i, int READ_STEP, int dim, int* psrc, int npsrc, hipStream_t St,
                    float* axx, float* ayy, float* azz, float* axz, float* ayz, float* axy,
                    uint16_t* xx, uint16_t* yy, uint16_t* zz, uint16_t* xy, uint16_t* yz,
                    uint16_t* xz, int d_i, int fmt, const float* scale) {
  const char* W = "awp_addsrc_s16";
  WaveScales SC;
  if (!awp_scales_from_host(W, fmt, scale, &SC)) return;
  if (fmt == AWP_FMT_BF16)
    addsrc_common<StBF16>(W, i, READ_STEP, dim, psrc, npsrc, St, axx, ayy, azz, axz, ayz, axy,
                          xx, yy, zz, xy, yz, xz, d_i, SC);
  else
    addsrc_common<StF16>(W, i, READ_STEP, dim, psrc, npsrc, St, axx, ayy, azz, axz, ayz, axy,
                         xx, yy, zz, xy, yz, xz, d_i, SC);
}

void velbuffer_H(const float* u1, const float* v1, const float* w1, const float* neta,
                 float* Bufx, float* Bufy, float* Bufz, float* Bufeta, int NVE, int nbgx,
                 int nedx, int nskpx, int nbgy, int nedy, int nskpy, int nbgz, int nedz,
                 int nskpz, int rec_nxt, int rec_nyt, int rec_nzt, hipStream_t St,
                 int FOLLOWBATHY, const int* bathy, int d_i) {
  const char* W = "velbuffer_H";
  if (!awp_grid_ok(d_i, W)) return;
  const AwpGrid& g = awp_state().grid[d_i];
  if (nskpx < 1 || nskpy < 1 || nskpz < 1 || rec_nxt < 1 || rec_nyt < 1 || rec_nzt < 1) return;
  if (nbgx < 0 || nedx >= g.nxt + 2 + AWP_NGSL || nbgy < 0 || nedy >= g.nyt + 2 + AWP_NGSL ||
      nbgz < 0 || (FOLLOWBATHY == 0 && nedz > g.nzt + AWP_ALIGN - 2)) {
    awp_set_error(AWP_ERR_ARG, "%s: sampling window leaves the padded grid", W);
    return;
  }
  if (d_i == 0 && FOLLOWBATHY != 0 && bathy == nullptr) {
    awp_set_error(AWP_ERR_ARG, "%s: FOLLOWBATHY on grid 0 needs the bathymetry array", W);
    return;
  }
  VbLaunch V;
  V.NVE = NVE; V.nbgx = nbgx; V.nedx = nedx; V.nskpx = nskpx;
  V.nbgy = nbgy; V.nedy = nedy; V.nskpy = nskpy;
  V.nbgz = nbgz; V.nedz = nedz; V.nskpz = nskpz;
  V.rec_nxt = rec_nxt; V.rec_nyt = rec_nyt; V.rec_nzt = rec_nzt; V.follow = FOLLOWBATHY; V.d_i = d_i;
  V.nzt = g.nzt; V.yline = g.yline; V.slice = g.slice; V.bathy_ld = g.nxt + 4 + AWP_NGSL2;
  dim3 block(16, 4, 4);
  dim3 grid(awp_cdiv(rec_nxt, 16), awp_cdiv(rec_nyt, 4), awp_cdiv(rec_nzt, 4));
  hipLaunchKernelGGL(velbuffer_kernel, grid, block, 0, St, u1, v1, w1, neta, Bufx, Bufy, Bufz,
                     Bufeta, bathy, V);
  awp_check_launch(W);
}

void awp_receivers_gather(const float* u1, const float* v1, const float* w1, const int* idx,
                          int npts, float* out, int step_slot, hipStream_t St, int d_i) {
  WaveScales SC = {};
  receivers_common<StF32>("awp_receivers_gather", u1, v1, w1, idx, npts, out, step_slot, St,
                          d_i, SC);
}

void awp_receivers_gather_s16(const uint16_t* u1, const uint16_t* v1, const uint16_t* w1,
                              const int* idx, int npts, float* out, int step_slot,
                              hipStream_t St, int d_i, int fmt, const float* scale) {
  const char* W = "awp_receivers_gather_s16";
  WaveScales SC;
  if (!awp_scales_from_host(W, fmt, scale, &SC)) return;
  if (fmt == AWP_FMT_BF16)
    receivers_common<StBF16>(W, u1, v1, w1, idx, npts, out, step_slot, St, d_i, SC);
  else
    receivers_common<StF16>(W, u1, v1, w1, idx, npts, out, step_slot, St, d_i, SC);
}

}  // extern "C"
