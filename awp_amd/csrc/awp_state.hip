// Process-global grid constants and error state for the C ABI.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <math.h>

#include "awp_common.h"
#include "../../include/awp_hip.h"

static AwpState g_state;
static int g_err = 0;
static char g_errmsg[512];

AwpState& awp_state() { return g_state; }

void awp_set_error(int code, const char* fmt, ...) {
  if (g_err != 0) return;  // keep the first error
  g_err = code;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_errmsg, sizeof(g_errmsg), fmt, ap);
  va_end(ap);
  fprintf(stderr, "awp_hip error: %s\n", g_errmsg);
  fflush(stderr);
}

bool awp_check_launch(const char* where) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    awp_set_error(AWP_ERR_HIP, "%s: %s", where, hipGetErrorString(e));
    return false;
  }
  return true;
}

bool awp_grid_ok(int d_i, const char* where) {
  if (!g_state.initialised) {
    awp_set_error(AWP_ERR_STATE, "%s: SetDeviceConstValue has not been called", where);
    return false;
  }
  if (d_i < 0 || d_i >= g_state.ngrids) {
    awp_set_error(AWP_ERR_ARG, "%s: grid index d_i=%d outside [0,%d)", where, d_i, g_state.ngrids);
    return false;
  }
  return true;
}

extern "C" {

int awp_last_error(void) { return g_err; }
const char* awp_last_error_string(void) { return g_err ? g_errmsg : ""; }
void awp_clear_error(void) {
  g_err = 0;
  g_errmsg[0] = 0;
}

// SetDeviceConstValue, reference src/awp/kernel.cu:68-117: the same
// derivations (dth = DT/DH in float, dh1 = 1.0/DH in double then float,
// c1 = 9/8, c2 = -1/24, slice/yline strides) stored host-side per grid.
void SetDeviceConstValue(float* DH, float DT, int* nxt, int* nyt, int* nzt, int ngrids,
                         float fmajor, float fminor, float* Rz, float* RzT) {
  if (ngrids < 1 || ngrids > AWP_MAXGRIDS) {
    awp_set_error(AWP_ERR_ARG, "SetDeviceConstValue: ngrids=%d outside [1,%d]", ngrids,
                  AWP_MAXGRIDS);
    return;
  }
  for (int p = 0; p < ngrids; ++p) {
    if (nxt[p] < 1 || nyt[p] < 1 || nzt[p] < 4) {
      awp_set_error(AWP_ERR_ARG, "SetDeviceConstValue: grid %d has nxt=%d nyt=%d nzt=%d", p,
                    nxt[p], nyt[p], nzt[p]);
      return;
    }
  }
  g_state.ngrids = ngrids;
  for (int p = 0; p < ngrids; ++p) {
    AwpGrid& g = g_state.grid[p];
    g.c1 = (float)(9.0 / 8.0);
    g.c2 = (float)(-1.0 / 24.0);
    g.dth = DT / DH[p];
    g.dh1 = (float)(1.0 / (double)DH[p]);
    g.DT = DT;
    g.DH = DH[p];
    g.nxt = nxt[p];
    g.nyt = nyt[p];
    g.nzt = nzt[p];
    g.yline = nzt[p] + 2 * AWP_ALIGN;
    g.slice = (long long)(nyt[p] + 4 + AWP_NGSL2) * g.yline;
  }
  g_state.plast.fmajor = fmajor;
  g_state.plast.fminor = fminor;
  for (int q = 0; q < 9; ++q) {
    g_state.plast.Rz[q] = Rz ? Rz[q] : 0.f;
    g_state.plast.RzT[q] = RzT ? RzT[q] : 0.f;
  }
  g_state.initialised = 1;
}

}  // extern "C"
