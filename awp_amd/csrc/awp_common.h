// Shared definitions for the gfx950 AWP-ODC kernels.
//
// Memory layout (reference src/awp/kernel.cu:87-94, src/awp/grid.c:14-38):
// every 3-D field is one flat float array with k (z) fastest, then j, then i:
//   p = i*slice + j*yline + k,  yline = nzt + 2*align,
//   slice = (nyt + 4 + 2*ngsl) * yline.
// Interior: i in [2+ngsl, nxt+ngsl+1], j in [2+ngsl, nyt+ngsl+1],
// k in [align, nzt+align-1]; k = nzt+align-1 is the free surface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define AWP_NGSL 8        // include/awp/pmcl3d_cons.h:22
#define AWP_NGSL2 16      // include/awp/pmcl3d_cons.h:23
#define AWP_ALIGN 32      // include/awp/pmcl3d_cons.h:16
#define AWP_MAXGRIDS 10   // include/awp/pmcl3d_cons.h:77
#define AWP_NXCD 8        // MI355X: 8 accelerator complex dies, each with its own L2

// Per-grid scalars, derived exactly as SetDeviceConstValue derives them
// (src/awp/kernel.cu:68-117).  They travel to kernels by value (kernarg
// segment -> scalar registers) instead of through __constant__ symbols.
struct AwpGrid {
  float c1, c2, dth, dh1, DT, DH;
  int nxt, nyt, nzt;
  int yline;      // nzt + 2*align
  long long slice;  // (nyt+4+ngsl2) * yline
};

// Plasticity constants (SetDeviceConstValue's fmajor/fminor/Rz/RzT).
struct AwpPlastConst {
  float fmajor, fminor;
  float Rz[9], RzT[9];
};

// Host-side state filled by SetDeviceConstValue; process-global like the
// reference's __constant__ block (not re-entrant, at most AWP_MAXGRIDS grids).
struct AwpState {
  int ngrids;
  int initialised;
  AwpGrid grid[AWP_MAXGRIDS];
  AwpPlastConst plast;
};

AwpState& awp_state();

// Error reporting.  The reference's launchers return void and either print or
// exit (pmcl3d_cons.h:43-53, kernel.cu:1233-1235).  These record the first
// error since the last awp_clear_error() and print it to stderr; the Python
// layer raises on it.
void awp_set_error(int code, const char* fmt, ...);
bool awp_check_launch(const char* where);
bool awp_grid_ok(int d_i, const char* where);

// XCD-aware tile order.  The hardware hands workgroup b to XCD (b % 8); this
// renumbers workgroups so that each XCD receives one contiguous run of logical
// tiles, which keeps the j/k halo lines neighbouring tiles share inside one
// L2.  The map is a bijection on [0, nblocks).
__device__ __forceinline__ int awp_xcd_remap(int b, int nblocks) {
  const int q = nblocks / AWP_NXCD, r = nblocks % AWP_NXCD;
  const int x = b % AWP_NXCD, s = b / AWP_NXCD;
  return x * q + (x < r ? x : r) + s;
}

static inline int awp_cdiv(int a, int b) { return (a + b - 1) / b; }

// k-run width of the (k, j)-tiled kernels: a wave covers 64 consecutive k of
// one j row, or 32 consecutive k of each of two j rows when that leaves at
// least 16 fewer lanes idle per column of n levels (n = 96 fills 2 x 64
// lanes to 75 % but 3 x 32 to 100 %).  forced: 32 or 64 from the tuning,
// else 0 for this rule.
static inline int awp_pick_kw(int n, int forced) {
  if (forced == 32 || forced == 64) return forced;
  const int idle64 = awp_cdiv(n, 64) * 64 - n, idle32 = awp_cdiv(n, 32) * 32 - n;
  return idle64 - idle32 >= 16 ? 32 : 64;
}

// The reference's (float)(c / (double)x) for binary32 c and x.  Rounding the
// binary64 quotient to binary32 gives the correctly rounded binary32 quotient:
// double rounding is innocuous when the wider format has at least 2p+2
// significand bits, and 53 >= 2*24+2 (S. A. Figueroa, "When is double rounding
// innocuous?", ACM SIGNUM Newsletter 30(3), 1995).  So one binary32 division
// (correctly rounded: hipcc's default) gives the same bits without binary64
// arithmetic.  A quotient in the subnormal range flushes to zero either way in
// this build.
__device__ __forceinline__ float awp_quot(float c, float x) { return c / x; }

// (k tile, j tile, i chunk) of logical block b: k tiles fastest, or with
// jfast the j tiles fastest, so that consecutive blocks (one XCD's L2 after
// awp_xcd_remap) are j neighbours and share their halo rows.
__device__ __forceinline__ void awp_tile3(int b, int nkt, int njt, int jfast, int& kt, int& jt,
                                          int& ic) {
  if (jfast) {
    jt = b % njt;
    kt = (b / njt) % nkt;
  } else {
    kt = b % nkt;
    jt = (b / nkt) % njt;
  }
  ic = b / (nkt * njt);
}

// Loads and stores through a wave-uniform row pointer and a per-lane 32-bit
// byte offset.  With one wave per j row the row pointer lives in scalar
// registers and every access of a row lowers to
// global_load/store v, v_koff, s[base:base+1]: one lane-offset VGPR shared
// by all fields instead of a 64-bit address pair per access.
__device__ __forceinline__ float awp_ld(const float* row, unsigned kbyte) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(row) + kbyte);
}
__device__ __forceinline__ int awp_ldi(const int* row, unsigned kbyte) {
  return *reinterpret_cast<const int*>(reinterpret_cast<const char*>(row) + kbyte);
}
__device__ __forceinline__ void awp_st(float* row, unsigned kbyte, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(row) + kbyte) = v;
}

