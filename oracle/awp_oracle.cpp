// AWP-ODC time-stepping oracle: a CPU restatement of the reference kernels.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
// awp_amd/csrc; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.  Nothing in the product path links it.
//
// Every function follows a reference kernel in operation order and literal
// type, so that for T=float the arithmetic is the reference's single-precision
// arithmetic (sub-expressions the reference writes with double literals are
// evaluated in double here too), and for T=double it is the FP64 oracle.
// Build with -ffp-contract=off: the reference's expression trees are restated
// without fused multiply-adds so that results are reproducible bit for bit.
//
// Memory layout (reference src/awp/kernel.cu:87-94, src/awp/grid.c:14-38):
//   one flat array per field, index p = i*slice + j*yline + k,
//   slice = (nyt+4+2*ngsl)*(nzt+2*align), yline = nzt+2*align, k fastest.
//   k = nzt+align-1 is the free surface; depth grows as k decreases.
//
// Pinning (what checks each function; see tests/test_oracle_*.py):
//   dvelcx  : cubic-exactness of the 4th-order staggered derivative at the
//             staggered points, constant-field zero update, 4-point density
//             means against closed forms, sponge factor application.
//   dstrqc  : elastic closed forms (harmonic 8-cell lambda/mu means, pair means)
//             for Q^-1=0, Q-branch weights against a brute-force evaluation of
//             the coarse-grained memory-variable recursion, plane-wave Q.
//   fs/fstr : closed-form ghost values on hand-built columns; antisymmetry.
//   addsrc  : closed form.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <omp.h>
#include <xmmintrin.h>

namespace {

// The reference is built with -use_fast_math (CMakeLists.txt:9-14), which
// flushes single-precision denormals to zero on input and output.  MXCSR
// FTZ|DAZ does the same on the host; MXCSR is per thread, so every parallel
// loop body sets it.
inline void ftz_daz() { _mm_setcsr(_mm_getcsr() | 0x8040); }

constexpr int NGSL = 8;    // include/awp/pmcl3d_cons.h:22
constexpr int NGSL2 = 16;  // include/awp/pmcl3d_cons.h:23
constexpr int ALIGN = 32;  // include/awp/pmcl3d_cons.h:16

// Per-grid constants as SetDeviceConstValue derives them
// (src/awp/kernel.cu:68-117): dth = DT/DH in _prec, dh1 = 1.0/DH in double
// then stored as _prec, c1 = 9/8, c2 = -1/24.
template <typename T>
struct Grid {
  T c1, c2, dth, dh1, DT, DH;
  long slice, yline;
  int nxt, nyt, nzt;
};

template <typename T>
Grid<T> make_grid(double DH, double DT, int nxt, int nyt, int nzt) {
  Grid<T> g;
  const T dh = (T)DH, dt = (T)DT;
  g.c1 = (T)(9.0 / 8.0);
  g.c2 = (T)(-1.0 / 24.0);
  g.dth = dt / dh;
  g.dh1 = (T)(1.0 / (double)dh);
  g.DT = dt;
  g.DH = dh;
  g.yline = nzt + 2 * ALIGN;
  g.slice = (long)(nyt + 4 + NGSL2) * g.yline;
  g.nxt = nxt;
  g.nyt = nyt;
  g.nzt = nzt;
  return g;
}

// dvelcx_opt, src/awp/kernel.cu:125-268 (register pipelines unrolled into
// direct indexing; offsets per the pipeline contents at iteration i).
template <typename T>
void dvelcx(T* u1, T* v1, T* w1, const T* xx, const T* yy, const T* zz,
            const T* xy, const T* xz, const T* yz, const T* dcrjx,
            const T* dcrjy, const T* dcrjz, const T* d1, const Grid<T>& g,
            int s_i, int e_i, int d_i, int ngrids) {
  const long sl = g.slice, yl = g.yline;
  const int j0 = 2 + NGSL, j1 = g.nyt + NGSL + 1;
  int k0 = ALIGN, k1 = g.nzt + ALIGN - 1;
  if (d_i > 0) k1 = g.nzt + ALIGN - 3;             // kernel.cu:146
  if (d_i < ngrids - 1) k0 = ALIGN + 3;            // kernel.cu:147
#pragma omp parallel for schedule(static)
  for (int i = s_i; i <= e_i; ++i) {
    ftz_daz();
    for (int j = j0; j <= j1; ++j) {
      for (int k = k0; k <= k1; ++k) {
        const long p = i * sl + j * yl + k;
        const T dcrj = dcrjx[i] * dcrjy[j] * dcrjz[k];
        // kernel.cu:214-220: 0.25 is a double literal (exact scaling).
        T d_1 = (T)(0.25 * (double)(d1[p] + d1[p - yl] + d1[p - 1] + d1[p - yl - 1]));
        T d_2 = (T)(0.25 * (double)(d1[p] + d1[p + sl] + d1[p - 1] + d1[p + sl - 1]));
        T d_3 = (T)(0.25 * (double)(d1[p] + d1[p + sl] + d1[p - yl] + d1[p + sl - yl]));
        d_1 = g.dth / d_1;
        d_2 = g.dth / d_2;
        d_3 = g.dth / d_3;
        const T c1 = g.c1, c2 = g.c2;
        u1[p] = (u1[p] + d_1 * (c1 * (xx[p] - xx[p - sl]) + c2 * (xx[p + sl] - xx[p - 2 * sl])
                              + c1 * (xy[p] - xy[p - yl]) + c2 * (xy[p + yl] - xy[p - 2 * yl])
                              + c1 * (xz[p] - xz[p - 1]) + c2 * (xz[p + 1] - xz[p - 2]))) * dcrj;
        v1[p] = (v1[p] + d_2 * (c1 * (xy[p + sl] - xy[p]) + c2 * (xy[p + 2 * sl] - xy[p - sl])
                              + c1 * (yy[p + yl] - yy[p]) + c2 * (yy[p + 2 * yl] - yy[p - yl])
                              + c1 * (yz[p] - yz[p - 1]) + c2 * (yz[p + 1] - yz[p - 2]))) * dcrj;
        w1[p] = (w1[p] + d_3 * (c1 * (xz[p + sl] - xz[p]) + c2 * (xz[p + 2 * sl] - xz[p - sl])
                              + c1 * (yz[p] - yz[p - yl]) + c2 * (yz[p + yl] - yz[p - 2 * yl])
                              + c1 * (zz[p + 1] - zz[p]) + c2 * (zz[p + 2] - zz[p - 1]))) * dcrj;
      }
    }
  }
}

// Coarse-grained Q weight, production dstrqc (kernel.cu:978-1045): the branch
// test and the quadratic are double expressions, the fallback is float.
template <typename T>
inline T qweight_prod(T q, const T* coeff, int ww, T wwo) {
  T w;
  if (1. / ((double)q * 2.0) <= 200.0)
    w = (T)((double)coeff[ww * 2 - 2] * (2. * (double)q) * (2. * (double)q) +
            (double)coeff[ww * 2 - 1] * (2. * (double)q));
  else
    w = (T)2.0f * wwo * q;
  return w / wwo;
}

// dstrqc_new's variants (kernel.cu:505-604): float test; qpa's quadratic stays
// double, the four shear-side quadratics are float.
template <typename T>
inline T qweight_new(T q, const T* coeff, int ww, T wwo, bool dbl_poly) {
  T w;
  if (1.0f / (q * 2.0f) <= 200.0f) {
    if (dbl_poly)
      w = (T)((double)coeff[ww * 2 - 2] * (2. * (double)q) * (2. * (double)q) +
              (double)coeff[ww * 2 - 1] * (2. * (double)q));
    else
      w = coeff[ww * 2 - 2] * (2.0f * q) * (2.0f * q) + coeff[ww * 2 - 1] * (2.0f * q);
  } else {
    w = 2.0f * wwo * q;
  }
  return w / wwo;
}

// Free-surface velocity ghosts written inside the stress kernel
// (kernel.cu:1087-1112 for dstrqc, 660-687 for dstrqc_new).  The k=top-1
// thread's writes of u1/v1 at top+1 repeat the same formula on the same
// interior values, so one write per column reproduces them.
template <typename T, bool NEWLIT>
void fs_ghosts(T* u1, T* v1, T* w1, const T* lam_mu, const Grid<T>& g, int NX,
               int NPC, int rankx, int ranky, int s_i, int e_i, int s_j, int e_j) {
  const long sl = g.slice, yl = g.yline;
  const int k = g.nzt + ALIGN - 1;
#pragma omp parallel for schedule(static)
  for (int i = s_i; i <= e_i; ++i) {
    ftz_daz();
    for (int j = s_j; j <= e_j; ++j) {
      const long p = i * sl + j * yl + k;
      u1[p + 1] = u1[p] - (w1[p] - w1[p - sl]);
      v1[p + 1] = v1[p] - (w1[p + yl] - w1[p]);
      T vs1, vs2;
      int gi = g.nxt * rankx + i - NGSL - 1;
      if (gi < NX || (NEWLIT && NPC == 2))
        vs1 = u1[p + sl] - (w1[p + sl] - w1[p]);
      else
        vs1 = (T)0.0;
      gi = g.nyt * ranky + j - NGSL - 1;
      if (gi > 1 || NPC == 2)
        vs2 = v1[p - yl] - (w1[p] - w1[p - yl]);
      else
        vs2 = (T)0.0;
      w1[p + 1] = w1[p - 1] - lam_mu[(long)i * (g.nyt + 4 + NGSL2) + j] *
                                  ((vs1 - u1[p + 1]) + (u1[p + sl] - u1[p]) +
                                   (v1[p + 1] - vs2) + (v1[p] - v1[p - yl]));
    }
  }
}

// dstrqc (kernel.cu:880-1197) when NEWLIT=false; dstrqc_new (343-859, the
// non-ELA, non-SMEM build) when NEWLIT=true.
template <typename T, bool NEWLIT>
void dstrqc(T* xx, T* yy, T* zz, T* xy, T* xz, T* yz, T* r1, T* r2, T* r3,
            T* r4, T* r5, T* r6, T* u1, T* v1, T* w1, const T* lam, const T* mu,
            const T* qp, const T* coeff, const T* qs, const T* dcrjx,
            const T* dcrjy, const T* dcrjz, const T* lam_mu, const T* vx1,
            const T* vx2, const int* ww, const T* wwo, int NX, int NPC,
            int rankx, int ranky, int s_i, int e_i, int s_j, int e_j, int d_i,
            const Grid<T>& g, int parts = 3) {
  // parts: bit 0 the free-surface velocity ghosts (grid 0), bit 1 the stress
  // sweep; the GPU runs them as two passes, the reference as one kernel.
  const long sl = g.slice, yl = g.yline;
  const int mink = ALIGN + 3;
  const int maxk = (d_i == 0) ? g.nzt + ALIGN - 1 : g.nzt + ALIGN - 3;
  if (d_i == 0 && (parts & 1))
    fs_ghosts<T, NEWLIT>(u1, v1, w1, lam_mu, g, NX, NPC, rankx, ranky, s_i, e_i, s_j, e_j);
  if (!(parts & 2)) return;
  const T c1 = g.c1, c2 = g.c2, DT = g.DT, dth = g.dth, dh1 = g.dh1;
#pragma omp parallel for schedule(static)
  for (int i = s_i; i <= e_i; ++i) {
    ftz_daz();
    for (int j = s_j; j <= e_j; ++j) {
      for (int k = mink; k <= maxk; ++k) {
        const long p = i * sl + j * yl + k;
        const long ip1 = p + sl, jm1 = p - yl, ijk = p + sl - yl, km1 = p - 1,
                   ik1 = p + sl - 1, jk1 = p - yl - 1, ijk1 = p + sl - yl - 1;
        const T f_vx1 = vx1[p], f_vx2 = vx2[p], f_wwo = wwo[p];
        const int f_ww = ww[p];
        const T f_dcrj = dcrjx[i] * dcrjy[j] * dcrjz[k];
        const T sl8 = lam[p] + lam[ip1] + lam[jm1] + lam[ijk] + lam[km1] + lam[ik1] + lam[jk1] + lam[ijk1];
        const T sm8 = mu[p] + mu[ip1] + mu[jm1] + mu[ijk] + mu[km1] + mu[ik1] + mu[jk1] + mu[ijk1];
        const T sp8 = qp[p] + qp[ip1] + qp[jm1] + qp[ijk] + qp[km1] + qp[ik1] + qp[jk1] + qp[ijk1];
        const T ss8 = qs[p] + qs[ip1] + qs[jm1] + qs[ijk] + qs[km1] + qs[ik1] + qs[jk1] + qs[ijk1];
        T xl, xm, xmu1, xmu2, xmu3, qpa, h, h1, h2, h3, qpaw, hw, h1w, h2w, h3w;
        if (!NEWLIT) {
          xl = (T)(8.0 / (double)sl8);
          xm = (T)(16.0 / (double)sm8);
          xmu1 = (T)(2.0 / (double)(mu[p] + mu[km1]));
          xmu2 = (T)(2.0 / (double)(mu[p] + mu[jm1]));
          xmu3 = (T)(2.0 / (double)(mu[p] + mu[ip1]));
          xl = xl + xm;
          qpa = (T)(0.0625 * (double)sp8);
          qpaw = qweight_prod(qpa, coeff, f_ww, f_wwo);
          h = (T)(0.0625 * (double)ss8);
          hw = qweight_prod(h, coeff, f_ww, f_wwo);
          h1 = (T)(0.250 * (double)(qs[p] + qs[km1]));
          h1w = qweight_prod(h1, coeff, f_ww, f_wwo);
          h2 = (T)(0.250 * (double)(qs[p] + qs[jm1]));
          h2w = qweight_prod(h2, coeff, f_ww, f_wwo);
          h3 = (T)(0.250 * (double)(qs[p] + qs[ip1]));
          h3w = qweight_prod(h3, coeff, f_ww, f_wwo);
        } else {
          xl = 8.0f / sl8;
          xm = 16.0f / sm8;
          xmu1 = 2.0f / (mu[p] + mu[km1]);
          xmu2 = (T)(2.0 / (double)(mu[p] + mu[jm1]));
          xmu3 = (T)(2.0 / (double)(mu[p] + mu[ip1]));
          xl = xl + xm;
          qpa = 0.0625f * sp8;
          qpaw = qweight_new(qpa, coeff, f_ww, f_wwo, true);
          h = 0.0625f * ss8;
          hw = qweight_new(h, coeff, f_ww, f_wwo, false);
          h1 = 0.250f * (qs[p] + qs[km1]);
          h1w = qweight_new(h1, coeff, f_ww, f_wwo, false);
          h2 = 0.250f * (qs[p] + qs[jm1]);
          h2w = qweight_new(h2, coeff, f_ww, f_wwo, false);
          h3 = 0.250f * (qs[p] + qs[ip1]);
          h3w = qweight_new(h3, coeff, f_ww, f_wwo, false);
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
        const T vx1p = DT * (1 + f_vx2 * f_vx1);

        T vs1 = c1 * (u1[p + sl] - u1[p]) + c2 * (u1[p + 2 * sl] - u1[p - sl]);
        T vs2 = c1 * (v1[p] - v1[p - yl]) + c2 * (v1[p + yl] - v1[p - 2 * yl]);
        T vs3 = c1 * (w1[p] - w1[p - 1]) + c2 * (w1[p + 1] - w1[p - 2]);
        T tmp = xl * (vs1 + vs2 + vs3);
        const T a1 = qpa * (vs1 + vs2 + vs3);
        tmp = tmp + DT * a1;

        T f_r, f_rtmp, s;
        // xx: sponge applied once.
        f_r = r1[p];
        f_rtmp = -h * (vs2 + vs3) + a1;
        s = xx[p] + tmp - xm * (vs2 + vs3) + vx1p * f_r;
        r1[p] = f_vx2 * f_r + f_wwo * f_rtmp;
        f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
        xx[p] = (s + DT * f_rtmp) * f_dcrj;
        // yy, zz: sponge applied twice (kernel.cu:1124-1142).
        f_r = r2[p];
        f_rtmp = -h * (vs1 + vs3) + a1;
        s = (yy[p] + tmp - xm * (vs1 + vs3) + vx1p * f_r) * f_dcrj;
        r2[p] = f_vx2 * f_r + f_wwo * f_rtmp;
        f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
        yy[p] = (s + DT * f_rtmp) * f_dcrj;
        f_r = r3[p];
        f_rtmp = -h * (vs1 + vs2) + a1;
        s = (zz[p] + tmp - xm * (vs1 + vs2) + vx1p * f_r) * f_dcrj;
        r3[p] = f_vx2 * f_r + f_wwo * f_rtmp;
        f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
        zz[p] = (s + DT * f_rtmp) * f_dcrj;
        // xy
        vs1 = c1 * (u1[p + yl] - u1[p]) + c2 * (u1[p + 2 * yl] - u1[p - yl]);
        vs2 = c1 * (v1[p] - v1[p - sl]) + c2 * (v1[p + sl] - v1[p - 2 * sl]);
        f_r = r4[p];
        f_rtmp = h1 * (vs1 + vs2);
        s = xy[p] + xmu1 * (vs1 + vs2) + vx1p * f_r;
        r4[p] = f_vx2 * f_r + f_wwo * f_rtmp;
        f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
        xy[p] = (s + DT * f_rtmp) * f_dcrj;
        // xz
        vs1 = c1 * (u1[p + 1] - u1[p]) + c2 * (u1[p + 2] - u1[p - 1]);
        vs2 = c1 * (w1[p] - w1[p - sl]) + c2 * (w1[p + sl] - w1[p - 2 * sl]);
        f_r = r5[p];
        f_rtmp = h2 * (vs1 + vs2);
        s = xz[p] + xmu2 * (vs1 + vs2) + vx1p * f_r;
        r5[p] = f_vx2 * f_r + f_wwo * f_rtmp;
        f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
        xz[p] = (s + DT * f_rtmp) * f_dcrj;
        // yz
        vs1 = c1 * (v1[p + 1] - v1[p]) + c2 * (v1[p + 2] - v1[p - 1]);
        vs2 = c1 * (w1[p + yl] - w1[p]) + c2 * (w1[p + 2 * yl] - w1[p - yl]);
        f_r = r6[p];
        f_rtmp = h3 * (vs1 + vs2);
        s = yz[p] + xmu3 * (vs1 + vs2) + vx1p * f_r;
        r6[p] = f_vx2 * f_r + f_wwo * f_rtmp;
        f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
        yz[p] = (s + DT * f_rtmp) * f_dcrj;
      }
    }
  }
}

// fstr (kernel.cu:1616-1647): antisymmetric stress imaging above the free
// surface, zero horizontal shear on it.
template <typename T>
void fstr(T* zz, T* xz, T* yz, const Grid<T>& g, int s_i, int e_i, int s_j, int e_j) {
  const long sl = g.slice, yl = g.yline;
  const int k = g.nzt + ALIGN - 1;
  for (int i = s_i; i <= e_i; ++i) {
    for (int j = s_j; j <= e_j; ++j) {
      const long p = i * sl + j * yl + k;
      zz[p + 1] = -zz[p];
      zz[p + 2] = -zz[p - 1];
      xz[p + 1] = -xz[p - 1];
      xz[p + 2] = -xz[p - 2];
      yz[p + 1] = -yz[p - 1];
      yz[p + 2] = -yz[p - 2];
      xz[p] = (T)0.0;
      yz[p] = (T)0.0;
    }
  }
}

// addsrc_cu (kernel.cu:1780-1806) / host addsrc (src/awp/source.c:271-308):
// sigma -= DT/DH^3 * Mdot at sample (i-1); sources applied in list order.
template <typename T>
void addsrc(int i, int READ_STEP, int dim, const int* psrc, int npsrc,
            const T* axx, const T* ayy, const T* azz, const T* axz,
            const T* ayz, const T* axy, T* xx, T* yy, T* zz, T* xy, T* yz,
            T* xz, const Grid<T>& g) {
  const T vtst = g.DT / (g.DH * g.DH * g.DH);
  i = i - 1;
  for (int j = 0; j < npsrc; ++j) {
    const int idx = psrc[j * dim] + 1 + NGSL;
    const int idy = psrc[j * dim + 1] + 1 + NGSL;
    const int idz = psrc[j * dim + 2] + ALIGN - 1;
    const long p = idx * g.slice + idy * g.yline + idz;
    const long s = (long)j * READ_STEP + i;
    xx[p] = xx[p] - vtst * axx[s];
    yy[p] = yy[p] - vtst * ayy[s];
    zz[p] = zz[p] - vtst * azz[s];
    xz[p] = xz[p] - vtst * axz[s];
    yz[p] = yz[p] - vtst * ayz[s];
    xy[p] = xy[p] - vtst * axy[s];
  }
}

// ---------------------------------------------------------------------------
// Drucker-Prager plasticity (kernel.cu:43-64 rotate_principal, 1253-1422
// drprecpc_calc_opt, 1650-1778 drprecpc_app).  Transcendentals (cos, sin, exp)
// are evaluated in double and rounded to T: the reference's -use_fast_math
// versions are approximations that no test pins, so a correctly rounded
// value is used on both sides.
template <typename T>
struct PlastC {
  T fmajor, fminor;
  T Rz[9], RzT[9];
};

template <typename T>
inline void matmul3(const T* a, const T* b, T* c) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      T acc = (T)0;
      for (int k = 0; k < 3; ++k) acc += a[i * 3 + k] * b[k * 3 + j];
      c[i * 3 + j] = acc;
    }
}

template <typename T>
inline void rotate_principal(T sigma2, T pfluid, const PlastC<T>& pc, T* ssp) {
  T ss[9], tmp[9];
  for (int k = 0; k < 9; ++k) ss[k] = tmp[k] = ssp[k] = (T)0;
  ss[0] = (sigma2 + pfluid) * pc.fmajor;
  ss[4] = sigma2 + pfluid;
  ss[8] = (sigma2 + pfluid) * pc.fminor;
  matmul3(pc.RzT, ss, tmp);
  matmul3(tmp, pc.Rz, ssp);
  ssp[0] -= pfluid;
  ssp[4] -= pfluid;
  ssp[8] -= pfluid;
}

template <typename T>
inline T t_cos(T x) { return (T)std::cos((double)x); }
template <typename T>
inline T t_sin(T x) { return (T)std::sin((double)x); }
template <typename T>
inline T t_exp(T x) { return (T)std::exp((double)x); }

// pfluid at depth (srfpos-k)*DH: (depth + DH/2) * 9.81e3, evaluated in double.
template <typename T>
inline T pfluid_at(int srfpos, int k, T DH) {
  const T depth = (T)(srfpos - k) * DH;
  if (depth > 0) return (T)(((double)depth + (double)DH * 0.5) * 9.81e3);
  return (T)((double)DH / 2. * 9.81e3);
}

template <typename T>
void drprecpc_calc(const T* xx, const T* yy, const T* zz, const T* xy, const T* xz,
                   const T* yz, const T* mu, const T* d1, const T* sigma2, T* yldfac,
                   const T* cohes, const T* phi, T* neta, int s_i, int e_i, int s_j, int e_j,
                   const Grid<T>& g, const PlastC<T>& pc, T Tv_fixed) {
  const long sl = g.slice, yl = g.yline;
  const int srfpos = g.nzt + ALIGN - 1;
#pragma omp parallel for schedule(static)
  for (int i = s_i; i <= e_i; ++i) {
    ftz_daz();
    for (int j = s_j; j <= e_j; ++j) {
      for (int k = ALIGN; k <= g.nzt + ALIGN - 1; ++k) {
        const long p = i * sl + j * yl + k;
        const T pfluid = pfluid_at<T>(srfpos, k, g.DH);
        T ini[9], ini_ip1[9];
        rotate_principal(sigma2[p], pfluid, pc, ini);
        rotate_principal(sigma2[p + sl], pfluid, pc, ini_ip1);
        const T iixx = 0.5f * (ini[0] + ini_ip1[0]);
        const T iiyy = 0.5f * (ini[4] + ini_ip1[4]);
        const T iizz = 0.5f * (ini[8] + ini_ip1[8]);
        T Sxx = xx[p] + iixx;
        T Syy = yy[p] + iiyy;
        T Szz = zz[p] + iizz;
        T Sxz = 0.25f * (xz[p] + xz[p + sl] + xz[p - 1] + xz[p + sl - 1]) + 0.5f * (ini[2] + ini_ip1[2]);
        T Syz = 0.25f * (yz[p] + yz[p - yl] + yz[p - 1] + yz[p - yl - 1]) + 0.5f * (ini[5] + ini_ip1[5]);
        T Sxy = 0.25f * (xy[p] + xy[p + sl] + xy[p - yl] + xy[p + sl - yl]) + 0.5f * (ini[1] + ini_ip1[1]);
        const T mu_i = mu[p], mu_ip1 = mu[p + sl];
        T Tv = g.DH / std::sqrt((T)1.0f / (mu_i * d1[p]));
        if (Tv_fixed > 0) Tv = Tv_fixed;
        const T Sxxp = Sxx, Syyp = Syy, Szzp = Szz, Sxyp = Sxy, Sxzp = Sxz, Syzp = Syz;
        const T rphi = phi[p] * (T)0.017453292519943295f;
        const T sigma_m = (Sxx + Syy + Szz) / (T)3.0f;
        const T SDxx = Sxx - sigma_m, SDyy = Syy - sigma_m, SDzz = Szz - sigma_m;
        const T secinv = 0.5f * (SDxx * SDxx + SDyy * SDyy + SDzz * SDzz) + Sxz * Sxz + Sxy * Sxy + Syz * Syz;
        const T sqrtSecinv = std::sqrt(secinv);
        const T taulim2 = cohes[p] * t_cos(rphi) - (sigma_m + pfluid) * t_sin(rphi);
        const T taulim = taulim2 > 0.0f ? taulim2 : (T)0.0f;
        T iyldfac;
        if (sqrtSecinv > taulim) {
          iyldfac = taulim / sqrtSecinv + ((T)1.0f - taulim / sqrtSecinv) * t_exp(-g.DT / Tv);
          Sxx = SDxx * iyldfac + sigma_m;
          Syy = SDyy * iyldfac + sigma_m;
          Szz = SDzz * iyldfac + sigma_m;
          Sxz = Sxz * iyldfac;
          Sxy = Sxy * iyldfac;
          Syz = Syz * iyldfac;
        } else {
          iyldfac = 1.0f;
        }
        yldfac[p] = iyldfac;
        if (iyldfac < 1.0f) {
          const T xm = 2.0f / (mu_i + mu_ip1);
          const T dxx = (Sxx - Sxxp) / xm, dyy = (Syy - Syyp) / xm, dzz = (Szz - Szzp) / xm;
          const T dxy = (Sxy - Sxyp) / xm, dxz = (Sxz - Sxzp) / xm, dyz = (Syz - Syzp) / xm;
          neta[p] = neta[p] + std::sqrt(0.5f * (dxx * dxx + dyy * dyy + dzz * dzz) +
                                        2.0f * (dxy * dxy + dxz * dxz + dyz * dyz));
        } else {
          yldfac[p] = 1.0f;
        }
      }
    }
  }
}

template <typename T>
void drprecpc_app(T* xx, T* yy, T* zz, T* xy, T* xz, T* yz, const T* sigma2,
                  const T* yldfac, int s_i, int e_i, int s_j, int e_j, const Grid<T>& g,
                  const PlastC<T>& pc) {
  const long sl = g.slice, yl = g.yline;
  const int srfpos = g.nzt + ALIGN - 1;
#pragma omp parallel for schedule(static)
  for (int i = s_i; i <= e_i; ++i) {
    ftz_daz();
    for (int j = s_j; j <= e_j; ++j) {
      for (int k = ALIGN; k <= g.nzt + ALIGN - 1; ++k) {
        const long p = i * sl + j * yl + k;
        const long im1 = p - sl, ip1 = p + sl, jp1 = p + yl, kp1 = p + 1;
        const T pf = pfluid_at<T>(srfpos, k, g.DH);
        const T pf1 = pfluid_at<T>(srfpos, k - 1, g.DH);  // depth of k+1 level
        T ini[9], inb[9], t3[9], t4[9], t5[9], t6[9], t7[9], t8[9];
        if (yldfac[p] < 1.) {
          rotate_principal(sigma2[p], pf, pc, ini);
          rotate_principal(sigma2[ip1], pf1, pc, inb);
          const T iixx = (T)(0.5 * (double)(ini[0] + inb[0]));
          const T iiyy = (T)(0.5 * (double)(ini[4] + inb[4]));
          const T iizz = (T)(0.5 * (double)(ini[8] + inb[8]));
          const T Sxx = xx[p] + iixx, Syy = yy[p] + iiyy, Szz = zz[p] + iizz;
          const T sigma_m = (T)((double)(Sxx + Syy + Szz) / 3.);
          const T SDxx = xx[p] + iixx - sigma_m;
          const T SDyy = yy[p] + iiyy - sigma_m;
          const T SDzz = zz[p] + iizz - sigma_m;
          xx[p] = SDxx * yldfac[p] + sigma_m - iixx;
          yy[p] = SDyy * yldfac[p] + sigma_m - iiyy;
          zz[p] = SDzz * yldfac[p] + sigma_m - iizz;
        }
        T iy = (T)(0.25 * (double)(yldfac[im1] + yldfac[p] + yldfac[im1 + 1] + yldfac[kp1]));
        if (iy < 1.) {
          rotate_principal(sigma2[p], pf, pc, ini);
          rotate_principal(sigma2[kp1], pf1, pc, inb);
          const T iist = (T)(0.5 * (double)(ini[2] + inb[2]));
          xz[p] = (xz[p] + iist) * iy - iist;
        }
        iy = (T)(0.25 * (double)(yldfac[p] + yldfac[jp1] + yldfac[jp1 + 1] + yldfac[kp1]));
        if (iy < 1.) {
          rotate_principal(sigma2[p], pf, pc, ini);
          rotate_principal(sigma2[ip1], pf, pc, inb);
          rotate_principal(sigma2[kp1], pf1, pc, t3);
          rotate_principal(sigma2[ip1 + 1], pf1, pc, t4);
          rotate_principal(sigma2[jp1], pf, pc, t5);
          rotate_principal(sigma2[ip1 + yl], pf, pc, t6);
          rotate_principal(sigma2[jp1 + 1], pf1, pc, t7);
          rotate_principal(sigma2[ip1 + yl + 1], pf1, pc, t8);
          const T iist = (T)(0.125 * (double)(ini[5] + inb[5] + t3[5] + t4[5] + t5[5] + t6[5] + t7[5] + t8[5]));
          yz[p] = (yz[p] + iist) * iy - iist;
        }
        iy = (T)(0.25 * (double)(yldfac[p] + yldfac[jp1] + yldfac[im1] + yldfac[im1 + yl]));
        if (iy < 1.) {
          rotate_principal(sigma2[p], pf, pc, ini);
          rotate_principal(sigma2[jp1], pf, pc, inb);
          const T iist = (T)(0.5 * (double)(ini[1] + inb[1]));
          xy[p] = (xy[p] + iist) * iy - iist;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Discontinuous mesh (kernel.cu:2289-3234).  A fine grid (d_i) sits on a
// coarse grid (d_i+1) with 3x refinement; each level's constants come from
// its own Grid.
constexpr int WWL = 2;  // include/awp/pmcl3d_cons.h:92

// dvelc2 (kernel.cu:2289-2343): 2nd-order velocity at k = align+2 of a fine grid.
template <typename T>
void dvelc2(T* u1, T* v1, T* w1, const T* xx, const T* yy, const T* zz, const T* xy,
            const T* xz, const T* yz, const T* dcrjx, const T* dcrjy, const T* dcrjz,
            const T* d1, const Grid<T>& g) {
  const long sl = g.slice, yl = g.yline;
  const int k = 2 + ALIGN;
  for (int i = 2 + NGSL; i < g.nxt + 2 + NGSL; ++i)
    for (int j = 2 + NGSL; j < g.nyt + 2 + NGSL; ++j) {
      const long p = i * sl + j * yl + k;
      T f_d3 = (T)(0.25 * (double)(d1[p] + d1[p + sl] + d1[p - yl] + d1[p + sl - yl]));
      f_d3 = g.dth / f_d3;
      const T f_dcrj = dcrjz[k] * dcrjx[i] * dcrjy[j];
      w1[p] = (w1[p] + f_d3 * ((xz[p + sl] - xz[p]) + (yz[p] - yz[p - yl]) + (zz[p + 1] - zz[p]))) * f_dcrj;
      T f_d1 = (T)(0.25 * (double)(d1[p] + d1[p - yl] + d1[p - 1] + d1[p - yl - 1]));
      T f_d2 = (T)(0.25 * (double)(d1[p] + d1[p + sl] + d1[p - 1] + d1[p + sl - 1]));
      f_d1 = g.dth / f_d1;
      f_d2 = g.dth / f_d2;
      u1[p] = (u1[p] + f_d1 * ((xx[p] - xx[p - sl]) + (xy[p] - xy[p - yl]) + (xz[p] - xz[p - 1]))) * f_dcrj;
      v1[p] = (v1[p] + f_d2 * ((xy[p + sl] - xy[p]) + (yy[p + yl] - yy[p]) + (yz[p] - yz[p - 1]))) * f_dcrj;
    }
}

// dstrqc2 (kernel.cu:2564-2761): 2nd-order stress + Q at k = align+2, with
// that kernel's Q-branch literals (qpa: float test, no factor 2 in the
// fallback; others: double test).
template <typename T>
void dstrqc2(T* xx, T* yy, T* zz, T* xy, T* xz, T* yz, T* r1, T* r2, T* r3, T* r4, T* r5,
             T* r6, const T* u1, const T* v1, const T* w1, const T* lam, const T* mu,
             const T* qp, const T* qs, const T* dcrjx, const T* dcrjy, const T* dcrjz,
             const T* coeff, const T* vx1, const T* vx2, const int* ww, const T* wwo, int s_i,
             int e_i, int s_j, int e_j, const Grid<T>& g) {
  const long sl = g.slice, yl = g.yline;
  const int k = ALIGN + 2;
  const T DT = g.DT, dth = g.dth, dh1 = g.dh1;
  for (int i = s_i; i <= e_i; ++i)
    for (int j = s_j; j <= e_j; ++j) {
      const long p = i * sl + j * yl + k;
      const long ip1 = p + sl, jm1 = p - yl, ijk = p + sl - yl, km1 = p - 1, ik1 = p + sl - 1,
                 jk1 = p - yl - 1, ijk1 = p + sl - yl - 1;
      const T f_dcrj = dcrjz[k] * dcrjx[i] * dcrjy[j];
      const T f_vx1 = vx1[p], f_vx2 = vx2[p], f_wwo = wwo[p];
      const int f_ww = ww[p];
      T xl = (T)(8.0 / (double)(lam[p] + lam[ip1] + lam[jm1] + lam[ijk] + lam[km1] + lam[ik1] + lam[jk1] + lam[ijk1]));
      T xm = (T)(16.0 / (double)(mu[p] + mu[ip1] + mu[jm1] + mu[ijk] + mu[km1] + mu[ik1] + mu[jk1] + mu[ijk1]));
      T xmu1 = (T)(2.0 / (double)(mu[p] + mu[km1]));
      T xmu2 = (T)(2.0 / (double)(mu[p] + mu[jm1]));
      T xmu3 = (T)(2.0 / (double)(mu[p] + mu[ip1]));
      xl = xl + xm;
      T qpa = (T)(0.0625 * (double)(qp[p] + qp[ip1] + qp[jm1] + qp[ijk] + qp[km1] + qp[ik1] + qp[jk1] + qp[ijk1]));
      T qpaw;
      if (1.0f / (qpa * 2.0f) <= 200.0f)
        qpaw = (T)((double)coeff[f_ww * 2 - 2] * (2. * (double)qpa) * (2. * (double)qpa) +
                   (double)coeff[f_ww * 2 - 1] * (2. * (double)qpa));
      else
        qpaw = f_wwo * qpa;
      qpaw = qpaw / f_wwo;
      auto qw2 = [&](T q) {
        T w;
        if (1. / ((double)q * 2.0) <= 200.0)
          w = (T)((double)coeff[f_ww * 2 - 2] * (2. * (double)q) * (2. * (double)q) +
                  (double)coeff[f_ww * 2 - 1] * (2. * (double)q));
        else
          w = f_wwo * q;
        return w / f_wwo;
      };
      T h = 0.0625f * (qs[p] + qs[ip1] + qs[jm1] + qs[ijk] + qs[km1] + qs[ik1] + qs[jk1] + qs[ijk1]);
      const T hw = qw2(h);
      T h1 = (T)(0.25 * (double)(qs[p] + qs[km1]));
      const T h1w = qw2(h1);
      T h2 = (T)(0.250 * (double)(qs[p] + qs[jm1]));
      const T h2w = qw2(h2);
      T h3 = (T)(0.250 * (double)(qs[p] + qs[ip1]));
      const T h3w = qw2(h3);
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
      const T vx1p = DT * (1 + f_vx2 * f_vx1);
      T f_r, f_rtmp, t, vs1, vs2, vs3;
      // xz, yz
      vs1 = u1[p + 1] - u1[p];
      vs2 = w1[p] - w1[p - sl];
      f_r = r5[p];
      f_rtmp = h2 * (vs1 + vs2);
      t = xz[p] + xmu2 * (vs1 + vs2) + vx1p * f_r;
      r5[p] = f_vx2 * f_r + f_wwo * f_rtmp;
      f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
      xz[p] = (t + DT * f_rtmp) * f_dcrj;
      vs1 = v1[p + 1] - v1[p];
      vs2 = w1[p + yl] - w1[p];
      f_r = r6[p];
      f_rtmp = h3 * (vs1 + vs2);
      t = yz[p] + xmu3 * (vs1 + vs2) + vx1p * f_r;
      r6[p] = f_vx2 * f_r + f_wwo * f_rtmp;
      f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
      yz[p] = (t + DT * f_rtmp) * f_dcrj;
      // normal
      vs1 = u1[p + sl] - u1[p];
      vs2 = v1[p] - v1[p - yl];
      vs3 = w1[p] - w1[p - 1];
      T tmp = xl * (vs1 + vs2 + vs3);
      const T a1 = qpa * (vs1 + vs2 + vs3);
      tmp = tmp + DT * a1;
      f_r = r1[p];
      f_rtmp = -h * (vs2 + vs3) + a1;
      t = xx[p] + tmp - xm * (vs2 + vs3) + vx1p * f_r;
      r1[p] = f_vx2 * f_r + f_wwo * f_rtmp;
      f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
      xx[p] = (t + DT * f_rtmp) * f_dcrj;
      f_r = r2[p];
      f_rtmp = -h * (vs1 + vs3) + a1;
      t = (yy[p] + tmp - xm * (vs1 + vs3) + vx1p * f_r) * f_dcrj;
      r2[p] = f_vx2 * f_r + f_wwo * f_rtmp;
      f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
      yy[p] = (t + DT * f_rtmp) * f_dcrj;
      f_r = r3[p];
      f_rtmp = -h * (vs1 + vs2) + a1;
      t = (zz[p] + tmp - xm * (vs1 + vs2) + vx1p * f_r) * f_dcrj;
      r3[p] = f_vx2 * f_r + f_wwo * f_rtmp;
      f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
      zz[p] = (t + DT * f_rtmp) * f_dcrj;
      // xy
      vs1 = u1[p + yl] - u1[p];
      vs2 = v1[p] - v1[p - sl];
      f_r = r4[p];
      f_rtmp = h1 * (vs1 + vs2);
      t = xy[p] + xmu1 * (vs1 + vs2) + vx1p * f_r;
      r4[p] = f_vx2 * f_r + f_wwo * f_rtmp;
      f_rtmp = f_rtmp * (f_wwo - 1) + f_vx2 * f_r * (1 - f_vx1);
      xy[p] = (t + DT * f_rtmp) * f_dcrj;
    }
}

// intp3d (kernel.cu:2359-2561): coarse (gl) -> fine (gh) bilinear 4x4 fill of
// the fine level align+1 from the coarse level nzt+align-3, weights
// 1, 2/3, 1/3, 0 in float; per-field staggered offsets (ox, oy).
template <typename T>
void intp3d(T* const* L9, T* const* H9, const Grid<T>& gl, const Grid<T>& gh) {
  // field order: u1 v1 w1 xx yy zz xy xz yz; offsets added to (ii-1, jj-1)
  static const int ox[9] = {0, 1, 1, 1, 1, 1, 0, 0, 1};
  static const int oy[9] = {0, 1, 0, 0, 0, 0, 1, 0, 1};
  const T w[4] = {(T)1., (T)(2. / 3.), (T)(1. / 3.), (T)0.};
  const long sl = gl.slice, yl = gl.yline;
  const int k = gl.nzt + ALIGN - 3;
  for (int i = NGSL; i <= gl.nxt + 3 + NGSL; ++i)
    for (int j = NGSL; j <= gl.nyt + 3 + NGSL; ++j) {
      const long p = i * sl + j * yl + k;
      const int ih = (i - 1 - 1 - NGSL) * 3 + 2 + NGSL;
      const int jh = (j - 1 - 1 - NGSL) * 3 + 2 + NGSL;
      const long ph = (long)ih * gh.slice + (long)jh * gh.yline + ALIGN + 1;
      for (int f = 0; f < 9; ++f) {
        const T* a = L9[f];
        const T v0 = a[p], v1 = a[p + sl], v2 = a[p + yl], v3 = a[p + sl + yl];
        for (int jj = 1; jj <= 4; ++jj)
          for (int ii = 1; ii <= 4; ++ii) {
            const long idx = ph + (long)(ii - 1 + ox[f]) * gh.slice + (long)(jj - 1 + oy[f]) * gh.yline;
            H9[f][idx] = v0 * w[ii - 1] * w[jj - 1] + v1 * w[4 - ii] * w[jj - 1] +
                         v2 * w[ii - 1] * w[4 - jj] + v3 * w[4 - ii] * w[4 - jj];
          }
      }
    }
}

// swap3 (kernel.cu:3032-3216): fine (gh) -> coarse (gl) triangular-weighted
// 5x5x5 averages onto the coarse levels nzt-2 and nzt-1 (+align).  The
// weights and the accumulation are double; the running sum is float.
// Fine-grid reads go through bgcaccess (kernel.cu:2809-2852), which takes a
// halo buffer only for an index outside [0, nxt+20) x [0, nyt+20); with
// WWL = 2 the index spans [2, nxt_fine+19], so every read is of the array.
template <typename T>
void swap3(T* const* L9, const T* const* H9, const Grid<T>& gl, const Grid<T>& gh) {
  // field order u1 v1 w1 xx yy zz xy xz yz; sbvpos of each (pmcl3d_cons.h:80-88)
  double ttl = 0.;
  for (int jj = -WWL; jj <= WWL; ++jj)
    for (int ii = -WWL; ii <= WWL; ++ii)
      for (int kk = -WWL; kk <= WWL; ++kk)
        ttl += (1. + WWL - std::abs(jj)) * (1. + WWL - std::abs(ii)) * (1. + WWL - std::abs(kk));
  struct Pass { int kc_off, kh, f[3], oi[3], oj[3]; };
  // (field, offset of ih/jh base: 2 or 3) per block of the reference kernel
  const Pass passes[5] = {
      {-2, ALIGN + 3, {0, 1, 6}, {2, 3, 2}, {2, 3, 3}},
      {-2, ALIGN + 3, {3, 4, 5}, {3, 3, 3}, {2, 2, 2}},
      {-2, ALIGN + 4, {7, 2, 8}, {2, 3, 3}, {2, 2, 3}},
      {-1, ALIGN + 6, {0, 1, 6}, {2, 3, 2}, {2, 3, 3}},
      {-1, ALIGN + 6, {3, 4, 5}, {3, 3, 3}, {2, 2, 2}}};
  const long sl = gl.slice, yl = gl.yline;
  for (int i = NGSL; i < gl.nxt + NGSL + 5; ++i)
    for (int j = NGSL; j < gl.nyt + NGSL + 5; ++j)
      for (const Pass& ps : passes) {
        const long p = i * sl + j * yl + ALIGN + gl.nzt + ps.kc_off;
        T sum[3] = {(T)0., (T)0., (T)0.};
        for (int jj = -WWL; jj <= WWL; ++jj)
          for (int ii = -WWL; ii <= WWL; ++ii)
            for (int kk = -WWL; kk <= WWL; ++kk)
              for (int q = 0; q < 3; ++q) {
                const int ih = (i - 1 - 1 - NGSL) * 3 + ps.oi[q] + NGSL;
                const int jh = (j - 1 - 1 - NGSL) * 3 + ps.oj[q] + NGSL;
                const T v = H9[ps.f[q]][(long)(ih + ii) * gh.slice + (long)(jh + jj) * gh.yline + ps.kh + kk];
                sum[q] = (T)((double)sum[q] + 1 / ttl * (1. + WWL - std::abs(ii)) *
                                                  (1. + WWL - std::abs(jj)) *
                                                  (1. + WWL - std::abs(kk)) * (double)v);
              }
        for (int q = 0; q < 3; ++q) L9[ps.f[q]][p] = sum[q];
      }
}

// ---------------------------------------------------------------------------
// 16-bit storage codecs: binary32 -> bfloat16 / binary16 with round to
// nearest even (binary16 with subnormals, overflow to infinity), and back.
inline uint16_t f32_to_bf16(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x0040u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

inline float bf16_to_f32(uint16_t q) {
  const uint32_t u = (uint32_t)q << 16;
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}

inline uint16_t f32_to_f16(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  const uint32_t sign = (u >> 16) & 0x8000u;
  const uint32_t a = u & 0x7fffffffu;
  if (a > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);   // NaN
  if (a >= 0x47800000u) return (uint16_t)(sign | 0x7c00u);  // >= 2^16: infinity
  if (a >= 0x38800000u) {                                   // binary16 normal range
    const uint32_t mant = a & 0x7fffffu;
    uint32_t h = ((((a >> 23) - 127u + 15u)) << 10) | (mant >> 13);
    const uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h += 1u;  // may carry into inf
    return (uint16_t)(sign | h);
  }
  if (a <= 0x33000000u) return (uint16_t)sign;  // <= 2^-25: rounds to zero
  const uint32_t mant = (a & 0x7fffffu) | 0x800000u;
  const uint32_t shift = 126u - (a >> 23);       // 14 .. 24
  uint32_t h = mant >> shift;
  const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1u);
  if (rem > half || (rem == half && (h & 1u))) h += 1u;
  return (uint16_t)(sign | h);
}

inline float f16_to_f32(uint16_t q) {
  const uint32_t sign = (uint32_t)(q & 0x8000u) << 16;
  const uint32_t e = (q >> 10) & 0x1fu, m = q & 0x3ffu;
  uint32_t u;
  if (e == 0) {
    if (m == 0) {
      u = sign;
    } else {  // subnormal: normalise
      int sh = 0;
      uint32_t mm = m;
      while (!(mm & 0x400u)) { mm <<= 1; ++sh; }
      u = sign | ((uint32_t)(127 - 15 + 1 - sh) << 23) | ((mm & 0x3ffu) << 13);
    }
  } else if (e == 31) {
    u = sign | 0x7f800000u | (m << 13);
  } else {
    u = sign | ((e - 15u + 127u) << 23) | (m << 13);
  }
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}

}  // namespace

extern "C" {

void awp_oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int awp_oracle_get_threads(void) { return omp_get_max_threads(); }

#define AWP_ORACLE_DVELCX(SUF, T)                                                   \
  void awp_oracle_dvelcx_##SUF(T* u1, T* v1, T* w1, const T* xx, const T* yy,       \
                               const T* zz, const T* xy, const T* xz, const T* yz,  \
                               const T* dcrjx, const T* dcrjy, const T* dcrjz,      \
                               const T* d1, int nxt, int nyt, int nzt, double DH,   \
                               double DT, int s_i, int e_i, int d_i, int ngrids) {  \
    ftz_daz();                                                                      \
    const Grid<T> g = make_grid<T>(DH, DT, nxt, nyt, nzt);                          \
    dvelcx<T>(u1, v1, w1, xx, yy, zz, xy, xz, yz, dcrjx, dcrjy, dcrjz, d1, g, s_i,  \
              e_i, d_i, ngrids);                                                    \
  }
AWP_ORACLE_DVELCX(f32, float)
AWP_ORACLE_DVELCX(f64, double)

#define AWP_ORACLE_DSTRQC(NAME, T, NEWLIT)                                           \
  void NAME(T* xx, T* yy, T* zz, T* xy, T* xz, T* yz, T* r1, T* r2, T* r3, T* r4,   \
            T* r5, T* r6, T* u1, T* v1, T* w1, const T* lam, const T* mu,           \
            const T* qp, const T* coeff, const T* qs, const T* dcrjx,               \
            const T* dcrjy, const T* dcrjz, const T* lam_mu, const T* vx1,          \
            const T* vx2, const int* ww, const T* wwo, int nxt, int nyt, int nzt,   \
            double DH, double DT, int NX, int NPC, int rankx, int ranky, int s_i,   \
            int e_i, int s_j, int e_j, int d_i) {                                   \
    ftz_daz();                                                                      \
    const Grid<T> g = make_grid<T>(DH, DT, nxt, nyt, nzt);                          \
    dstrqc<T, NEWLIT>(xx, yy, zz, xy, xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1,   \
                      lam, mu, qp, coeff, qs, dcrjx, dcrjy, dcrjz, lam_mu, vx1,     \
                      vx2, ww, wwo, NX, NPC, rankx, ranky, s_i, e_i, s_j, e_j, d_i, \
                      g);                                                           \
  }
AWP_ORACLE_DSTRQC(awp_oracle_dstrqc_f32, float, false)

void awp_oracle_dstrqc_parts_f32(
    float* xx, float* yy, float* zz, float* xy, float* xz, float* yz, float* r1, float* r2,
    float* r3, float* r4, float* r5, float* r6, float* u1, float* v1, float* w1,
    const float* lam, const float* mu, const float* qp, const float* coeff, const float* qs,
    const float* dcrjx, const float* dcrjy, const float* dcrjz, const float* lam_mu,
    const float* vx1, const float* vx2, const int* ww, const float* wwo, int nxt, int nyt,
    int nzt, double DH, double DT, int NX, int NPC, int rankx, int ranky, int s_i, int e_i,
    int s_j, int e_j, int d_i, int parts) {
  ftz_daz();
  const Grid<float> g = make_grid<float>(DH, DT, nxt, nyt, nzt);
  dstrqc<float, false>(xx, yy, zz, xy, xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1, lam, mu, qp,
                       coeff, qs, dcrjx, dcrjy, dcrjz, lam_mu, vx1, vx2, ww, wwo, NX, NPC,
                       rankx, ranky, s_i, e_i, s_j, e_j, d_i, g, parts);
}
AWP_ORACLE_DSTRQC(awp_oracle_dstrqc_f64, double, false)
AWP_ORACLE_DSTRQC(awp_oracle_dstrqc_new_f32, float, true)

#define AWP_ORACLE_FSTR(SUF, T)                                                     \
  void awp_oracle_fstr_##SUF(T* zz, T* xz, T* yz, int nxt, int nyt, int nzt,        \
                             int s_i, int e_i, int s_j, int e_j) {                  \
    ftz_daz();                                                                      \
    const Grid<T> g = make_grid<T>(1.0, 1.0, nxt, nyt, nzt);                        \
    fstr<T>(zz, xz, yz, g, s_i, e_i, s_j, e_j);                                     \
  }
AWP_ORACLE_FSTR(f32, float)
AWP_ORACLE_FSTR(f64, double)

#define AWP_ORACLE_ADDSRC(SUF, T)                                                    \
  void awp_oracle_addsrc_##SUF(int i, int READ_STEP, int dim, const int* psrc,       \
                               int npsrc, const T* axx, const T* ayy, const T* azz,  \
                               const T* axz, const T* ayz, const T* axy, T* xx,      \
                               T* yy, T* zz, T* xy, T* yz, T* xz, int nxt, int nyt,  \
                               int nzt, double DH, double DT) {                      \
    ftz_daz();                                                                       \
    const Grid<T> g = make_grid<T>(DH, DT, nxt, nyt, nzt);                           \
    addsrc<T>(i, READ_STEP, dim, psrc, npsrc, axx, ayy, azz, axz, ayz, axy, xx, yy,  \
              zz, xy, yz, xz, g);                                                    \
  }
AWP_ORACLE_ADDSRC(f32, float)
AWP_ORACLE_ADDSRC(f64, double)

#define AWP_ORACLE_PLAST(SUF, T)                                                        \
  void awp_oracle_drprecpc_calc_##SUF(const T* xx, const T* yy, const T* zz, const T* xy,  \
                                      const T* xz, const T* yz, const T* mu, const T* d1,  \
                                      const T* sigma2, T* yldfac, const T* cohes,          \
                                      const T* phi, T* neta, int nxt, int nyt, int nzt,    \
                                      double DH, double DT, int s_i, int e_i, int s_j,     \
                                      int e_j, const double* pl, double Tv_fixed) {        \
    ftz_daz();                                                                          \
    const Grid<T> g = make_grid<T>(DH, DT, nxt, nyt, nzt);                              \
    PlastC<T> pc;                                                                       \
    pc.fmajor = (T)pl[0];                                                               \
    pc.fminor = (T)pl[1];                                                               \
    for (int q = 0; q < 9; ++q) {                                                       \
      pc.Rz[q] = (T)pl[2 + q];                                                          \
      pc.RzT[q] = (T)pl[11 + q];                                                        \
    }                                                                                   \
    drprecpc_calc<T>(xx, yy, zz, xy, xz, yz, mu, d1, sigma2, yldfac, cohes, phi, neta,   \
                     s_i, e_i, s_j, e_j, g, pc, (T)Tv_fixed);                           \
  }                                                                                     \
  void awp_oracle_drprecpc_app_##SUF(T* xx, T* yy, T* zz, T* xy, T* xz, T* yz,          \
                                     const T* sigma2, const T* yldfac, int nxt, int nyt,   \
                                     int nzt, double DH, int s_i, int e_i, int s_j,        \
                                     int e_j, const double* pl) {                          \
    ftz_daz();                                                                          \
    const Grid<T> g = make_grid<T>(DH, 1.0, nxt, nyt, nzt);                             \
    PlastC<T> pc;                                                                       \
    pc.fmajor = (T)pl[0];                                                               \
    pc.fminor = (T)pl[1];                                                               \
    for (int q = 0; q < 9; ++q) {                                                       \
      pc.Rz[q] = (T)pl[2 + q];                                                          \
      pc.RzT[q] = (T)pl[11 + q];                                                        \
    }                                                                                   \
    drprecpc_app<T>(xx, yy, zz, xy, xz, yz, sigma2, yldfac, s_i, e_i, s_j, e_j, g, pc);  \
  }
AWP_ORACLE_PLAST(f32, float)
AWP_ORACLE_PLAST(f64, double)

void awp_oracle_rotate_principal_f64(double sigma2, double pfluid, const double* pl, double* out9) {
  PlastC<double> pc;
  pc.fmajor = pl[0];
  pc.fminor = pl[1];
  for (int q = 0; q < 9; ++q) {
    pc.Rz[q] = pl[2 + q];
    pc.RzT[q] = pl[11 + q];
  }
  rotate_principal<double>(sigma2, pfluid, pc, out9);
}

#define AWP_ORACLE_DM(SUF, T)                                                            \
  void awp_oracle_dvelc2_##SUF(T* u1, T* v1, T* w1, const T* xx, const T* yy, const T* zz,  \
                               const T* xy, const T* xz, const T* yz, const T* dcrjx,       \
                               const T* dcrjy, const T* dcrjz, const T* d1, int nxt,        \
                               int nyt, int nzt, double DH, double DT) {                    \
    ftz_daz();                                                                           \
    dvelc2<T>(u1, v1, w1, xx, yy, zz, xy, xz, yz, dcrjx, dcrjy, dcrjz, d1,               \
              make_grid<T>(DH, DT, nxt, nyt, nzt));                                      \
  }                                                                                      \
  void awp_oracle_dstrqc2_##SUF(T* xx, T* yy, T* zz, T* xy, T* xz, T* yz, T* r1, T* r2,      \
                                T* r3, T* r4, T* r5, T* r6, const T* u1, const T* v1,       \
                                const T* w1, const T* lam, const T* mu, const T* qp,        \
                                const T* qs, const T* dcrjx, const T* dcrjy,                \
                                const T* dcrjz, const T* coeff, const T* vx1,               \
                                const T* vx2, const int* ww, const T* wwo, int nxt,         \
                                int nyt, int nzt, double DH, double DT, int s_i, int e_i,   \
                                int s_j, int e_j) {                                         \
    ftz_daz();                                                                           \
    dstrqc2<T>(xx, yy, zz, xy, xz, yz, r1, r2, r3, r4, r5, r6, u1, v1, w1, lam, mu, qp,   \
               qs, dcrjx, dcrjy, dcrjz, coeff, vx1, vx2, ww, wwo, s_i, e_i, s_j, e_j,      \
               make_grid<T>(DH, DT, nxt, nyt, nzt));                                     \
  }                                                                                      \
  void awp_oracle_intp3d_##SUF(T* const* L9, T* const* H9, int nxtl, int nytl, int nztl,   \
                               int nxth, int nyth, int nzth) {                             \
    ftz_daz();                                                                           \
    intp3d<T>(L9, H9, make_grid<T>(1.0, 1.0, nxtl, nytl, nztl),                          \
              make_grid<T>(1.0, 1.0, nxth, nyth, nzth));                                  \
  }                                                                                      \
  void awp_oracle_swap3_##SUF(T* const* L9, const T* const* H9, int nxtl, int nytl,        \
                              int nztl, int nxth, int nyth, int nzth) {                    \
    ftz_daz();                                                                           \
    swap3<T>(L9, H9, make_grid<T>(1.0, 1.0, nxtl, nytl, nztl),                           \
             make_grid<T>(1.0, 1.0, nxth, nyth, nzth));                                   \
  }
AWP_ORACLE_DM(f32, float)
AWP_ORACLE_DM(f64, double)

// q = rne(x * inv) and x = dec(q) * s over n values (fmt 1 bfloat16, 2
// binary16), with the binary32 arithmetic in flush-to-zero mode like the GPU.
void awp_oracle_s16_encode(const float* x, uint16_t* q, long long n, int fmt, float inv) {
  ftz_daz();
  for (long long i = 0; i < n; ++i) {
    volatile float y = x[i] * inv;
    q[i] = fmt == 1 ? f32_to_bf16(y) : f32_to_f16(y);
  }
}

void awp_oracle_s16_decode(const uint16_t* q, float* x, long long n, int fmt, float s) {
  ftz_daz();
  for (long long i = 0; i < n; ++i) {
    volatile float y = fmt == 1 ? bf16_to_f32(q[i]) : f16_to_f32(q[i]);
    x[i] = y * s;
  }
}

}  // extern "C"
