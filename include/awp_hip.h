/* C ABI of the gfx950 (MI355X) AWP-ODC time-stepping library, libawp_hip.so.
 *
 * Entry points keep the names, argument order and meaning of the reference's
 * kernel launchers in include/awp/kernel_launch.h (SCECcode/awp), with
 * hipStream_t where the reference has cudaStream_t.  The caller allocates and
 * owns every device array; launchers never allocate fields, mutate them in
 * place and run asynchronously on the given stream.
 *
 * Layout of every 3-D field (reference src/awp/kernel.cu:87-94): one flat
 * float array of (nxt+4+16) * (nyt+4+16) * (nzt+64) elements with k (z)
 * contiguous, index i*slice + j*yline + k, yline = nzt+64,
 * slice = (nyt+20)*yline.  k = nzt+31 is the free surface.
 *
 * Errors: the launchers return void like the reference's.  An invalid
 * argument (index range outside the padded arrays, grid index without
 * SetDeviceConstValue, ...) launches nothing, prints a message to stderr and
 * records an error code readable with awp_last_error(); launch failures are
 * recorded the same way.  The reference prints or exits instead
 * (include/awp/pmcl3d_cons.h:43-53, src/awp/kernel.cu:1233-1235).
 */
#ifndef AWP_HIP_H
#define AWP_HIP_H

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AWP_ERR_ARG 1   /* invalid argument: nothing was launched */
#define AWP_ERR_STATE 2 /* SetDeviceConstValue missing or inconsistent */
#define AWP_ERR_HIP 3   /* a HIP runtime call or launch failed */

/* Error channel (no reference counterpart; the reference prints or exits). */
int awp_last_error(void);
const char* awp_last_error_string(void);
void awp_clear_error(void);

/* Launch-shape knobs (no reference counterpart), process-global; results do
 * not depend on them.  Keys: vel_bj, str_bj (rows per workgroup: 1, 2, 4, 8,
 * 16), vel_target_wgs, str_target_wgs, vel_ichunk, str_ichunk (i-march
 * chunking, 0 = automatic), xcd_remap, kalign (0/1), topo_vel_ichunk,
 * topo_str_ichunk, topo_q_ichunk (planes per march, 1..4096), step_ichunk
 * (0 = automatic), step_target_wgs (awp_step's i-march chunking).  Returns 0,
 * or -1 for an unknown key or a value out of range. */
int awp_set_tuning(const char* key, int value);

/* Replaces SetDeviceConstValue, include/awp/kernel_launch.h:24-26
 * (src/awp/kernel.cu:68-117).  Stores per-grid DT/DH-derived constants and
 * strides, plus the plasticity rotation (fmajor, fminor, Rz, RzT). */
void SetDeviceConstValue(float* DH, float DT, int* nxt, int* nyt, int* nzt, int ngrids,
                         float fmajor, float fminor, float* Rz, float* RzT);

/* Replaces dvelcx_H_opt, include/awp/kernel_launch.h:31-34
 * (src/awp/kernel.cu:298-316, kernel dvelcx_opt :125-268).  4th-order
 * staggered velocity update of u1, v1, w1 for i in [s_i, e_i], all interior
 * j, and k in [align, nzt+align-1] (narrowed by the DM guards for d_i). */
void dvelcx_H_opt(float* u1, float* v1, float* w1, float* xx, float* yy, float* zz,
                  float* xy, float* xz, float* yz, float* dcrjx, float* dcrjy,
                  float* dcrjz, float* d_1, int nyt, int nzt, hipStream_t St, int s_i,
                  int e_i, int d_i, int ngrids);

/* Replaces dstrqc_H, include/awp/kernel_launch.h:51-58 (src/awp/kernel.cu:
 * 861-878, kernel dstrqc :880-1197).  Stress + coarse-grained Q memory
 * variables for i in [s_i, e_i], j in [s_j, e_j], k in [align+3, top];
 * on grid 0 it also writes the free-surface velocity ghosts at k = top+1. */
void dstrqc_H(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz, float* r1,
              float* r2, float* r3, float* r4, float* r5, float* r6, float* u1, float* v1,
              float* w1, float* lam, float* mu, float* qp, float* coeff, float* qs,
              float* dcrjx, float* dcrjy, float* dcrjz, int nyt, int nzt, hipStream_t St,
              float* lam_mu, float* vx1, float* vx2, int* ww, float* wwo, int NX, int NPC,
              int rankx, int ranky, int s_i, int e_i, int s_j, int e_j, int d_i);

/* Replaces dstrqc_H_new, include/awp/kernel_launch.h:42-49 (src/awp/kernel.cu:
 * 1201-1239, kernel dstrqc_new :343-859): the same update with that kernel's
 * float literals and its periodic (NPC==2) rule on the x edge. */
void dstrqc_H_new(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz,
                  float* r1, float* r2, float* r3, float* r4, float* r5, float* r6,
                  float* u1, float* v1, float* w1, float* lam, float* mu, float* qp,
                  float* coeff, float* qs, float* dcrjx, float* dcrjy, float* dcrjz, int nyt,
                  int nzt, hipStream_t St, float* lam_mu, float* vx1, float* vx2, int* ww,
                  float* wwo, int NX, int NPC, int rankx, int ranky, int s_i, int e_i,
                  int s_j, int e_j, int d_i);

/* Replaces fstr_H, include/awp/kernel_launch.h:59-60 (src/awp/kernel.cu:
 * 1243-1250, kernel fstr :1616-1647).  Antisymmetric stress imaging above the
 * free surface of grid 0 and zero xz, yz on it. */
void fstr_H(float* zz, float* xz, float* yz, hipStream_t St, int s_i, int e_i, int s_j,
            int e_j);

/* Replaces addsrc_H, include/awp/kernel_launch.h:73-76 (src/awp/kernel.cu:
 * 1465-1491, kernel addsrc_cu :1780-1806).  sigma -= DT/DH^3 * a[j*READ_STEP
 * + i-1] at each source cell (x+1+ngsl, y+1+ngsl, z+align-1).  Two sources on
 * one cell race, as in the reference. */
void addsrc_H(int i, int READ_STEP, int dim, int* psrc, int npsrc, hipStream_t St,
              float* axx, float* ayy, float* azz, float* axz, float* ayz, float* axy,
              float* xx, float* yy, float* zz, float* xy, float* yz, float* xz, int d_i);

/* Replaces SetDeviceFilterParameters, include/awp/kernel_launch.h:25-26
 * (src/awp/kernel.cu:119-123): the IIR filter (order <= 19, a[0] != 0)
 * addkinsrc_H applies to every source time function; order 0 = none.  Kept
 * host-side and passed to the kernel by value. */
void SetDeviceFilterParameters(int filtorder, double* srcfilt_b, double* srcfilt_a);
/* Replaces addkinsrc_H, include/awp/kernel_launch.h:121-126 (src/awp/kernel.cu:
 * 3677-3858): kinematic subfaults.  Two parameter rows per subfault j: row 2j
 * of axx..ayz = STF type (1 Brune, 2 Liu), slip, rupture time, rise time,
 * area; row 2j+1 of axx, ayy, azz = strike, dip, rake in degrees.  Once step
 * i*DT passes the rupture time, sigma -= DT/DH^3 * stf * slip * area / mu *
 * M(strike, dip, rake) at the subfault's cell and mom[j] += stf * slip * area
 * / mu * DT; srcfilt_d holds npsrc x (order+1) doubles of filter state. */
void addkinsrc_H(int i, int dim, int* psrc, int npsrc, hipStream_t St, float* mu, float* axx,
                 float* ayy, float* azz, float* axz, float* ayz, float* axy, float* xx, float* yy,
                 float* zz, float* xy, float* yz, float* xz, float* mom, double* srcfilt_d,
                 int d_i);
/* Replaces frcvel_H, include/awp/kernel_launch.h:74-76 (src/awp/kernel.cu:
 * 1808-1929): prescribed velocities at the source cells, sampled every tskp
 * steps in axx, ayy, azz (rows 0..READ_STEP, npsrc wide: the arrays hold
 * (READ_STEP+1)*npsrc floats) and linearly interpolated in time; at the last
 * prescribed step the force that holds the velocity is stored in row
 * READ_STEP and applied from then on.  ymin/ymax (-1: none) bound the cells'
 * x index, as the reference's kernel reads them. */
void frcvel_H(int i, int READ_STEP, int dim, int* psrc, int npsrc, int tskp, hipStream_t St,
              float* axx, float* ayy, float* azz, float* axz, float* ayz, float* axy, float* u1,
              float* v1, float* w1, int ymin, int ymax, int d_i);

/* Replaces velbuffer_H, include/awp/kernel_launch.h:181-186
 * (src/awp/kernel.cu:3927-4011).  Strided sub-volume sampling: u1, v1 at
 * ko = nzt+align-1-k and w1 at ko-1 on grid 0 (FOLLOWBATHY = 0), ko on finer
 * levels; neta when NVE == 3.  FOLLOWBATHY = 1 uses bathy[j*(nxt+20)+i]. */
void velbuffer_H(const float* u1, const float* v1, const float* w1, const float* neta,
                 float* Bufx, float* Bufy, float* Bufz, float* Bufeta, int NVE, int nbgx,
                 int nedx, int nskpx, int nbgy, int nedy, int nskpy, int nbgz, int nedz,
                 int nskpz, int rec_nxt, int rec_nyt, int rec_nzt, hipStream_t St,
                 int FOLLOWBATHY, const int* bathy, int d_i);

/* Replaces drprecpc_calc_H_opt, include/awp/kernel_launch.h:62-66
 * (src/awp/kernel.cu:1425-1443, kernel drprecpc_calc_opt :1253-1422).
 * Drucker-Prager yield factor yldfac and plastic strain neta (+=) on
 * i in [xls, xre], j in [yls, yre], k in [align, nzt+align-1], with the
 * initial stress rotated from sigma2 by SetDeviceConstValue's fmajor,
 * fminor, Rz, RzT and fluid pressure 9.81e3*(depth+DH/2); viscoplastic
 * relaxation time Tv = DH/vs. */
void drprecpc_calc_H_opt(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz,
                         float* mu, float* d1, float* sigma2, float* yldfac, float* cohes,
                         float* phi, float* neta, int nzt, int xls, int xre, int yls, int yre,
                         hipStream_t St, int d_i);

/* No reference counterpart: drprecpc_calc_H_opt with a fixed viscoplastic
 * relaxation time Tv > 0 in place of DH/vs (the reference hard-codes DH/vs,
 * src/awp/kernel.cu:1362). */
void awp_drprecpc_calc_tv(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz,
                          float* mu, float* d1, float* sigma2, float* yldfac, float* cohes,
                          float* phi, float* neta, int nzt, int xls, int xre, int yls, int yre,
                          hipStream_t St, int d_i, float Tv);

/* Replaces drprecpc_app_H, include/awp/kernel_launch.h:68-71
 * (src/awp/kernel.cu:1445-1463, kernel drprecpc_app :1650-1778).  Scales the
 * deviatoric normal stresses by yldfac and the shear stresses by the 4-point
 * mean yield factor at their staggered points, about the initial stress. */
void drprecpc_app_H(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz,
                    float* mu, float* sigma2, float* yldfac, int nzt, int xls, int xre, int yls,
                    int yre, hipStream_t St, int d_i);

/* No reference counterpart: gathers u1, v1 at cell (i, j, k) and w1 at
 * (i, j, k-1) on grid 0 (k-1 -> k on d_i > 0), velbuffer's sampling rule, for
 * an arbitrary list of npts receiver cells idx[3*n + {0,1,2}] = (i, j, k) in
 * array coordinates, writing out[step_slot*3*npts + 3*n + c]. */
void awp_receivers_gather(const float* u1, const float* v1, const float* w1, const int* idx,
                          int npts, float* out, int step_slot, hipStream_t St, int d_i);

/* dvelcx_H_opt followed by dvelc2_H in one launch (no reference
 * counterpart; it combines kernel.cu:298-316 and :2289-2356): on a grid above
 * a coarser one (d_i < ngrids-1) the lane at k = align+2, which the 4th-order
 * update skips there, applies dvelc2's 2nd-order update from the operands
 * the sweep loads anyway.  s_i, e_i must be the grid's interior (dvelc2_H's
 * range); on the lowest grid it is dvelcx_H_opt.  Bit for bit the two
 * launches' result. */
void awp_dvelcx_dm(float* u1, float* v1, float* w1, float* xx, float* yy, float* zz, float* xy,
                   float* xz, float* yz, float* dcrjx, float* dcrjy, float* dcrjz, float* d_1,
                   int nyt, int nzt, hipStream_t St, int s_i, int e_i, int d_i, int ngrids);

/* Replaces dvelc2_H, include/awp/kernel_launch.h:92-95
 * (src/awp/kernel.cu:2289-2356).  Discontinuous mesh, fine grid d_i: velocity
 * at k = align+2 with 2nd-order differences, on the interior columns. */
void dvelc2_H(float* u1, float* v1, float* w1, float* xx, float* yy, float* zz, float* xy,
              float* xz, float* yz, float* dcrjx, float* dcrjy, float* dcrjz, float* d_1, int nxt,
              int nyt, hipStream_t St, int d_i);

/* Replaces dstrqc2_H, include/awp/kernel_launch.h:97-103
 * (src/awp/kernel.cu:2564-2784).  Discontinuous mesh, fine grid d_i: stress
 * and Q memory variables at k = align+2 with 2nd-order differences on
 * i in [s_i, e_i], j in [s_j, e_j]. */
void dstrqc2_H(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz, float* r1,
               float* r2, float* r3, float* r4, float* r5, float* r6, float* u1, float* v1,
               float* w1, float* lam, float* mu, float* qp, float* qs, float* dcrjx, float* dcrjy,
               float* dcrjz, int nxt, int nyt, hipStream_t St, float* coeff, float* d_vx1,
               float* d_vx2, int* d_ww, float* d_wwo, int s_i, int e_i, int s_j, int e_j,
               int d_i);

/* dstrqc2_H (reference src/awp/kernel.cu:2564-2784) with the relaxation
 * texture as the table of awp_dstrqc_tex: tex (host, 24 floats: vx1[8],
 * vx2[8], wwo[8]) indexed by ww - 1, ww read from the ww array or, with ww
 * NULL, from init_texture's parity rule (src/awp/mesh.c:1103-1137) at this
 * level.  Each fine column's level costs a memory line per array read, so
 * the table saves three (ww given) or four lines per cell.  Results equal
 * dstrqc2_H's bit for bit whenever the arrays dstrqc2_H would read hold
 * tex[ww - 1]. */
void awp_dstrqc2_tex(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz,
                     float* r1, float* r2, float* r3, float* r4, float* r5, float* r6, float* u1,
                     float* v1, float* w1, float* lam, float* mu, float* qp, float* qs,
                     float* dcrjx, float* dcrjy, float* dcrjz, int nxt, int nyt, hipStream_t St,
                     float* coeff, const float* tex, int* ww, int s_i, int e_i, int s_j, int e_j,
                     int d_i);

/* Replaces intp3d_H, include/awp/kernel_launch.h:105-109
 * (src/awp/kernel.cu:2359-2561, 2786-2804).  d_i is the coarse grid ("l"
 * arrays); the fine grid d_i-1 ("h" arrays) receives, at k = align+1, the
 * 4x4 bilinear interpolation of the coarse level k = nzt+align-3.  rank is
 * unused. */
void intp3d_H(float* u1l, float* v1l, float* w1l, float* xxl, float* yyl, float* zzl, float* xyl,
              float* xzl, float* yzl, float* u1h, float* v1h, float* w1h, float* xxh, float* yyh,
              float* zzh, float* xyh, float* xzh, float* yzh, int nxtl, int nytl, int rank,
              hipStream_t St, int d_i);

/* Replaces swap_H, include/awp/kernel_launch.h:111-116
 * (src/awp/kernel.cu:3032-3234).  d_i is the fine grid ("h" arrays); the
 * coarse grid d_i+1 ("l" arrays) receives at k = nzt+align-2 and nzt+align-1
 * the 5x5x5 triangular-weighted averages of the fine levels align+1..align+8.
 * The halo buffers buf_L/R/F/B are accepted and never read: every fine index
 * the average touches lies inside the padded fine array.  rank is unused. */
void swap_H(float* xxl, float* yyl, float* zzl, float* xyl, float* xzl, float* yzl, float* u1l,
            float* v1l, float* w1l, float* xxh, float* yyh, float* zzh, float* xyh, float* xzh,
            float* yzh, float* u1h, float* v1h, float* w1h, int nxtl, int nytl, float* buf_L,
            float* buf_R, float* buf_F, float* buf_B, int rank, hipStream_t St, int d_i);

/* ---- 16-bit wavefield storage (no reference counterpart) ----------------
 * For runs whose nine wavefield components and six memory variables are kept
 * in HBM as bfloat16 (fmt 1) or IEEE binary16 (fmt 2) with one power-of-two
 * scale per field: value = stored * scale.  Every load decodes to binary32,
 * all arithmetic is the binary32 arithmetic of the launcher named, and every
 * store rounds x / scale to nearest even.  scale is a host array of 15
 * scales in the order u1 v1 w1 xx yy zz xy xz yz r1 r2 r3 r4 r5 r6.  The
 * medium stays binary32.  Arguments otherwise mean what they mean for the
 * binary32 launcher. */
void awp_dvelcx_s16(uint16_t* u1, uint16_t* v1, uint16_t* w1, uint16_t* xx, uint16_t* yy,
                    uint16_t* zz, uint16_t* xy, uint16_t* xz, uint16_t* yz, float* dcrjx,
                    float* dcrjy, float* dcrjz, float* d_1, int nyt, int nzt, hipStream_t St,
                    int s_i, int e_i, int d_i, int ngrids, int fmt, const float* scale);
void awp_dstrqc_s16(uint16_t* xx, uint16_t* yy, uint16_t* zz, uint16_t* xy, uint16_t* xz,
                    uint16_t* yz, uint16_t* r1, uint16_t* r2, uint16_t* r3, uint16_t* r4,
                    uint16_t* r5, uint16_t* r6, uint16_t* u1, uint16_t* v1, uint16_t* w1,
                    float* lam, float* mu, float* qp, float* coeff, float* qs, float* dcrjx,
                    float* dcrjy, float* dcrjz, int nyt, int nzt, hipStream_t St, float* lam_mu,
                    float* vx1, float* vx2, int* ww, float* wwo, int NX, int NPC, int rankx,
                    int ranky, int s_i, int e_i, int s_j, int e_j, int d_i, int fmt,
                    const float* scale);
/* awp_dstrqc_s16 with the relaxation texture of awp_dstrqc_tex: tex (host,
 * 24 floats: vx1[8], vx2[8], wwo[8]) indexed by ww - 1, ww read, or with ww
 * NULL from init_texture's parity rule (16 B/cell less than the 86 B of
 * awp_dstrqc_s16).  The same results whenever the arrays awp_dstrqc_s16
 * would read hold tex[ww - 1]. */
void awp_dstrqc_tex_s16(uint16_t* xx, uint16_t* yy, uint16_t* zz, uint16_t* xy, uint16_t* xz,
                        uint16_t* yz, uint16_t* r1, uint16_t* r2, uint16_t* r3, uint16_t* r4,
                        uint16_t* r5, uint16_t* r6, uint16_t* u1, uint16_t* v1, uint16_t* w1,
                        float* lam, float* mu, float* qp, float* coeff, float* qs, float* dcrjx,
                        float* dcrjy, float* dcrjz, int nyt, int nzt, hipStream_t St,
                        float* lam_mu, const float* tex, int* ww, int NX, int NPC, int rankx,
                        int ranky, int s_i, int e_i, int s_j, int e_j, int d_i, int fmt,
                        const float* scale);
void awp_fstr_s16(uint16_t* zz, uint16_t* xz, uint16_t* yz, hipStream_t St, int s_i, int e_i,
                  int s_j, int e_j, int fmt, const float* scale);
void awp_addsrc_s16(int i, int READ_STEP, int dim, int* psrc, int npsrc, hipStream_t St,
                    float* axx, float* ayy, float* azz, float* axz, float* ayz, float* axy,
                    uint16_t* xx, uint16_t* yy, uint16_t* zz, uint16_t* xy, uint16_t* yz,
                    uint16_t* xz, int d_i, int fmt, const float* scale);
void awp_receivers_gather_s16(const uint16_t* u1, const uint16_t* v1, const uint16_t* w1,
                              const int* idx, int npts, float* out, int step_slot,
                              hipStream_t St, int d_i, int fmt, const float* scale);
/* Element-wise helpers on n stored values: encode binary32 x (q = rne(x /
 * scale)), decode (x = q * scale), rescale in place by 2^shift (exact unless
 * a value leaves the format's range), and the largest |q| (unit scale) into
 * the device float *out (NaN if any stored value is NaN). */
void awp_s16_encode(const float* x, uint16_t* q, long long n, int fmt, float scale,
                    hipStream_t St);
void awp_s16_decode(const uint16_t* q, float* x, long long n, int fmt, float scale,
                    hipStream_t St);
void awp_s16_rescale(uint16_t* q, long long n, int fmt, int shift, hipStream_t St);
void awp_s16_absmax(const uint16_t* q, long long n, int fmt, float* out, hipStream_t St);

/* ---- Adjoint and material gradient (no reference counterpart) ----------
 * For callers that need dJ/d(lambda, mu, rho) of a receiver misfit through
 * the NVE = 1 step of one grid (awp_amd/adjoint.py drives them).
 * awp_adjoint_step applies the transpose of step n to the adjoint state
 * adj[15] (xx yy zz xy xz yz r1..r6 u1 v1 w1, binary32, in place), reading
 * the forward velocities after step n (fwd_v[3]: u1 v1 w1) and the forward
 * stresses before it (fwd_s[6]: xx yy zz xy xz yz); med[9] = d1 lam mu qp qs
 * vx1 vx2 wwo lam_mu; scratch[9] are zero-initialised full-size arrays that
 * the call reuses; grad[8] (f_d1 f_d2 f_d3 xl xm xmu1 xmu2 xmu3) and
 * grad_lam_mu accumulate coefficient gradients in binary32; col holds
 * 3 x (nxt+20) x (nyt+20) floats of per-column scratch.  The last part of
 * the transposed velocity update (its gather into the six stress adjoints)
 * is applied below the free-surface band (k <= top-3) by the next call, which
 * reads those adjoints anyway; scratch[6..8] carry its weights between calls.
 * After the last step, awp_adjoint_settle applies it, if the caller needs
 * the adjoint state itself (the gradients do not). */
void awp_adjoint_step(float* const* adj, const float* const* fwd_v, const float* const* fwd_s,
                      const float* const* med, const int* ww, const float* coeff,
                      const float* dcrjx, const float* dcrjy, const float* dcrjz,
                      float* const* scratch, float* const* grad, float* grad_lam_mu,
                      float* col, hipStream_t St);
void awp_adjoint_settle(float* const* adj, float* const* scratch, hipStream_t St);
/* adds the binary64 residuals res[3*n + c] of receiver n (the awp_receivers_
 * gather sampling) to the adjoint velocities */
void awp_adjoint_inject(float* ub, float* vb, float* wb, const int* idx, int npts,
                        const double* res, hipStream_t St);
/* g64 += g32; g32 = 0 over n values */
void awp_adjoint_flush(float* g32, double* g64, long long n, hipStream_t St);
/* res = d - obs (binary64) and *J = 1/2 sum res^2 over n samples; work holds
 * AWP_MISFIT_PARTS doubles (one partial sum per block) */
#define AWP_MISFIT_PARTS 256
void awp_misfit(const float* traces, const float* obs, long long n, double* res, double* J,
                double* work, hipStream_t St);
/* awp_misfit without the work argument and without any allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The partial sums go to a library-owned device array of
 * AWP_MISFIT_PARTS doubles (2 KiB per GPU, held until the process ends),
 * looked up on the first call on each device.  Any n >= 0 fits (the scratch
 * size does not grow with n); n < 0 is rejected with AWP_ERR_ARG.  Results are
 * those of awp_misfit bit for bit.  All calls on one GPU share that array:
 * calls that overlap on one GPU (different streams, or threads) race. */
void awp_misfit_noalloc(const float* traces, const float* obs, long long n, double* res,
                        double* J, hipStream_t St);
/* maps the binary64 coefficient gradients gcoef[8] and grad_lam_mu to the
 * gradients in lam (1/lambda), mu (1/mu) and d1 (rho) of every padded cell */
void awp_adjoint_map(const double* const* gcoef, const double* grad_lam_mu,
                     const float* const* med, const int* ww, const float* coeff,
                     const float* dcrjx, const float* dcrjy, const float* dcrjz, double* glam,
                     double* gmu, double* grho, hipStream_t St);

/* Out-of-place forms of dvelcx_H_opt and dstrqc_H (production Q weights) for
 * the adjoint's segment recompute, which keeps every state of a checkpoint
 * segment: the field before the step is read from u0 v0 w0 (s0[6]: xx yy zz
 * xy xz yz) and the updated field written to u1 v1 w1 (xx..yz), which must be
 * distinct arrays; memory variables r1..r6 are updated in place.  The free-
 * surface velocity ghosts are written into u1 v1 w1 as dstrqc_H does.  Every
 * cell the in-place launcher would leave untouched is left untouched in the
 * output array, so outputs whose untouched cells hold the input's values give
 * the in-place results bit for bit. */
void awp_dvelcx_oop(float* u1, float* v1, float* w1, const float* u0, const float* v0,
                    const float* w0, float* xx, float* yy, float* zz, float* xy, float* xz,
                    float* yz, float* dcrjx, float* dcrjy, float* dcrjz, float* d_1, int nyt,
                    int nzt, hipStream_t St, int s_i, int e_i, int d_i, int ngrids);
void awp_dstrqc_oop(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz,
                    const float* const* s0, float* r1, float* r2, float* r3, float* r4,
                    float* r5, float* r6, float* u1, float* v1, float* w1, float* lam, float* mu,
                    float* qp, float* coeff, float* qs, float* dcrjx, float* dcrjy,
                    float* dcrjz, int nyt, int nzt, hipStream_t St, float* lam_mu, float* vx1,
                    float* vx2, int* ww, float* wwo, int NX, int NPC, int rankx, int ranky,
                    int s_i, int e_i, int s_j, int e_j, int d_i);

/* dstrqc_H (reference src/awp/kernel.cu:1201-1239) with the relaxation
 * texture as a table.  init_texture (src/awp/mesh.c:1103-1137) assigns every
 * cell one of 8 mechanisms ww = 1..8, and vx1, vx2 and wwo are functions of
 * the mechanism alone: tex (host, 24 floats) holds vx1[8], vx2[8], wwo[8]
 * indexed by ww - 1, replacing the three per-cell arrays (12 B/cell less).
 * ww NULL takes the mechanism from init_texture's parity rule over this
 * rank's interior (16 B/cell less).  Results equal dstrqc_H's bit for bit
 * whenever the arrays dstrqc_H would read hold tex[ww - 1]. */
void awp_dstrqc_tex(float* xx, float* yy, float* zz, float* xy, float* xz, float* yz,
                    float* r1, float* r2, float* r3, float* r4, float* r5, float* r6, float* u1,
                    float* v1, float* w1, float* lam, float* mu, float* qp, float* coeff,
                    float* qs, float* dcrjx, float* dcrjy, float* dcrjz, int nyt, int nzt,
                    hipStream_t St, float* lam_mu, const float* tex, int* ww, int NX, int NPC,
                    int rankx, int ranky, int s_i, int e_i, int s_j, int e_j, int d_i);

/* The extended entry point: one NVE = 1 step of a single Cartesian grid
 * (grid 0 of SetDeviceConstValue, ngrids = 1, one rank: NX = nxt, NPC = 0)
 * in one sweep, equal bit for bit to dvelcx_H_opt over the interior
 * (kernel.cu:298-316) followed by dstrqc_H over the interior (kernel.cu:
 * 1201-1239, its free-surface ghosts included).  It reads the state before
 * the step from a9 (u1 v1 w1 xx yy zz xy xz yz) and writes the state after
 * it to b9 (same order; distinct arrays); r6 (r1..r6) are updated in place.
 * b9 must hold a9's values in every cell the step does not write: give both
 * the same initial state, and pass the arrays of the last b9 as the next a9
 * (addsrc_H, fstr_H and the receivers act on b9 in between, as on the
 * in-place arrays).  med = d1 lam mu qp qs; tex and ww are the relaxation
 * texture as for awp_dstrqc_tex (ww NULL: init_texture's parity rule). */
void awp_step(const float* const* a9, float* const* b9, float* const* r6,
              const float* const* med, const float* coeff, const float* tex, const int* ww,
              const float* dcrjx, const float* dcrjy, const float* dcrjz, const float* lam_mu,
              int nxt, int nyt, int nzt, hipStream_t St);

/* HBM streaming probe (no reference counterpart): reads nin arrays and writes
 * nout arrays of n floats, every element once, moving 4 bytes per lane and
 * array (vec4 = 0), 16 bytes (vec4 = 1) or 8 bytes (vec4 = 2; n a multiple of
 * the width, arrays aligned to it), over `blocks` workgroups of 256 lanes.
 * batched = 0: each array's load is waited for before the next is issued;
 * batched = 1: a lane issues the loads of all nin arrays before using any.
 * bench.py --workload stream uses it to measure the bandwidth an access mix
 * reaches without stencils.  Returns 0 or an AWP_ERR_* code
 * (nin, nout <= 32; nin >= 1 when batched). */
int awp_stream_probe(const float* const* in, int nin, float* const* out, int nout, long long n,
                     int vec4, int blocks, int batched, hipStream_t St);

/* ---- The reference's structs at the topography and Lagrange boundary -----
 * The topography and Lagrange launchers take the reference's own structs by
 * pointer.  They are restated here member for member, in the reference's
 * order and with its types: _prec and prec are float, CUDA's dim3 is three
 * 32-bit unsigned integers (awp_dim3_t), cudaStream_t is hipStream_t (both
 * are opaque pointers).  A caller built against the reference's headers
 * therefore hands these launchers the layout they read; the CPU tests check
 * every member offset against the member lists of the reference's headers.
 * The launchers read only the members named in their comments; the others
 * are carried for layout. */
typedef struct {
  int x, y, z;
} int3_t; /* include/awp/definitions.h:55-58 */

typedef struct {
  uint32_t x, y, z;
} awp_dim3_t; /* CUDA's dim3 */

/* fcn_grid_t = grid3_t, include/grid/grid_3d.h:90-111.  The Lagrange path
 * reads offset1, line and slice (grid_index: offset1.z + k + (offset1.y + j)
 * * line + (offset1.x + i) * slice), size, shift and boundary2 (the
 * coordinate array lengths of grid_boundary_size, src/grid/grid_3d.c:143-160)
 * and mem (a field holds mem.x * mem.y * mem.z values). */
typedef struct {
  int3_t size;
  int3_t inner_size;
  int3_t outer_size;
  int3_t mem;
  int3_t coordinate;
  int3_t shift;
  int3_t offset1;
  int3_t offset2;
  int3_t boundary1;
  int3_t boundary2;
  int3_t alignment;
  float gridspacing;
  int3_t padding;
  int num_bytes;
  int line;
  int slice;
  int exclude_top_row;
} fcn_grid_t;
typedef fcn_grid_t grid3_t;

/* f_grid_t, include/topography/metrics/metrics.h:58-97.  The launchers read
 * the device arrays d_f .. d_f2_c: (nx+20) x (ny+84) floats, element (i, j)
 * of topographic coordinates i in [-10, nx+10), j in [-42, ny+42) at
 * (j+42) + (i+10)*(ny+84); d_f holds the geometry on [-8, nx+8) x [-8, ny+8)
 * and awp_topo_metrics_f fills the others on [-6, n+6)^2. */
typedef struct {
  int size[3];
  int mem[3];
  int offset[3];
  int bounds_x[2];
  int bounds_y[2];
  int bounds_stress_x[2];
  int bounds_stress_y[2];
  float hi;
  int line;
  int slice;
  float *f, *f_1, *f_2, *f_c, *f1_1, *f1_2, *f1_c, *f2_1, *f2_2, *f2_c; /* host */
  float *d_f, *d_f_1, *d_f_2, *d_f_c, *d_f1_1, *d_f1_2, *d_f1_c, *d_f2_1, *d_f2_2, *d_f2_c;
} f_grid_t;

/* g_grid_t, include/topography/metrics/metrics.h:106-127.  The launchers
 * read d_g, d_g_c, d_g3, d_g3_c: nz+64 floats each, level k at k+32. */
typedef struct {
  int size;
  int mem;
  int line;
  int offset;
  int bounds_z[2];
  float hi;
  float *g, *g_c, *g3, *g3_c;         /* host */
  float *d_g, *d_g_c, *d_g3, *d_g3_c; /* device */
} g_grid_t;

/* topo_t, include/topography/topography.h:136-273.  The launchers read use,
 * the neighbour ranks, nx ny nz, dth timestep gridspacing, the medium
 * (rho lami mui qpi qsi ww wwo vx1 vx2 coeff), the wavefield (xx .. yz,
 * r1 .. r6, u1 v1 w1), dcrjx dcrjy dcrjz, metrics_f, metrics_g and
 * stream_i.  Wavefield and medium arrays are the 3-D fields of the
 * Cartesian kernels: topographic (i, j, k) is array element
 * (i+10, j+10, k+32). */
typedef struct {
  int use;
  int dbg;
  int verbose;
  int rank;
  const char* topography_file;
  int x_rank_l; /* -1 on the domain boundary; on one GPU all four are -1 */
  int x_rank_r;
  int y_rank_b;
  int y_rank_f;
  int coord[2];
  int px;
  int py;
  int nx;
  int ny;
  int nz;
  int mx;
  int my;
  int mz;
  int off_x[4];
  int off_y[4];
  int off_z[4];
  int velocity_bounds_left[2];
  int velocity_bounds_right[2];
  int velocity_bounds_front[2];
  int velocity_bounds_back[2];
  int velocity_bounds_xdir[2];
  int velocity_offset_x[6];
  int velocity_offset_y[6];
  int boundary_top_size;
  int boundary_bottom_size;
  awp_dim3_t velocity_grid_interior;
  awp_dim3_t velocity_grid_front;
  awp_dim3_t velocity_grid_back;
  int stress_bounds_left[2];
  int stress_bounds_right[2];
  int stress_bounds_ydir[2];
  int stress_offset_x[6];
  int stress_offset_y[4];
  awp_dim3_t stress_grid_interior;
  awp_dim3_t stress_grid_left;
  awp_dim3_t stress_grid_right;
  int sxx[3];
  int syy[3];
  int szz[3];
  int sxy[3];
  int sxz[3];
  int syz[3];
  int su1[3];
  int sv1[3];
  int sw1[3];
  int gridsize;
  int slice;
  int slice_gl;
  int line;
  float dth;         /* dt/h */
  float timestep;    /* 1, or 0 for the spatial part only */
  float gridspacing; /* h */
  float *rho, *lami, *mui, *lam_mu;
  float *xx, *yy, *zz, *xy, *xz, *yz;
  float *r1, *r2, *r3, *r4, *r5, *r6;
  float *qpi, *qsi;
  int* ww;
  float *wwo, *vx1, *vx2, *coeff;
  float *u1, *v1, *w1;
  float *f_u1, *f_v1, *f_w1, *b_u1, *b_v1, *b_w1;
  f_grid_t metrics_f;
  g_grid_t metrics_g;
  grid3_t topography_grid;
  grid3_t stress_grid;
  grid3_t velocity_grid;
  float *x1, *y1, *z1;
  float *dcrjx, *dcrjy, *dcrjz;
  hipStream_t stream_1;
  hipStream_t stream_2;
  hipStream_t stream_i;
} topo_t;

/* ---- Curvilinear topography (grid 0) ------------------------------------- */
/* topo_set_constants, src/topography/stress_attenuation.cu:19-23: the
 * stress kernels' DT, dt/h and 1/h from T's dth and gridspacing, kept
 * process-wide as the reference keeps them in __constant__ memory; the
 * stress launchers refuse to run before it. */
void topo_set_constants(topo_t* T);
/* include/topography/velocity.cuh:16-18: curvilinear velocity on levels
 * 3 <= k < nz (interior stencils below nz-6, one-sided rows above).  front and
 * back compute the halo-send strips of a multi-GPU run: with no neighbour
 * (-1) they return, as the reference's do; a neighbour is an error here. */
void topo_velocity_interior_H(topo_t* T);
void topo_velocity_front_H(topo_t* T);
void topo_velocity_back_H(topo_t* T);
/* include/topography/stress.cuh:17-19: curvilinear stress + Q on levels
 * 3 <= k < nz of the stress bounds (topography.c:377-420) */
void topo_stress_interior_H(topo_t* T);
void topo_stress_left_H(topo_t* T);
void topo_stress_right_H(topo_t* T);
/* no reference counterpart: topo_stress_interior_H in two passes over
 * caller-owned scratch `rates` (AWP_TOPO_NRATES arrays of the 3-D field
 * size, back to back): the strain rates first, then the streaming material
 * and memory-variable update; the same results bit for bit, faster. */
#define AWP_TOPO_NRATES 6
void awp_topo_stress_interior_rates_H(topo_t* T, float* rates);
/* no reference counterpart: awp_topo_stress_interior_rates_H with the
 * relaxation texture of awp_dstrqc_tex: tex (host, 24 floats: vx1[8],
 * vx2[8], wwo[8]) indexed by ww - 1, ww read from T->ww (rule = 0) or from
 * init_texture's parity rule over grid 0 (rule = 1, mesh.c:1103-1137; the
 * vx1, vx2, wwo and ww pointers of T are then not read).  The same results
 * bit for bit when the per-cell arrays are those functions of ww. */
void awp_topo_stress_interior_tex_H(topo_t* T, float* rates, const float* tex, int rule);
/* no reference counterpart: the same updates with 16-bit wavefield storage
 * (the wavefield pointers of T address uint16_t; fmt and scale as for
 * awp_dvelcx_s16; rates NULL for one pass, or the two-pass scratch) */
void awp_topo_velocity_s16(topo_t* T, int fmt, const float* scale);
void awp_topo_stress_s16(topo_t* T, int fmt, const float* scale, float* rates);
/* geom_custom, src/topography/geometry/geometry.c:188-210: f = 1 + elevation
 * / (h (nz-2)) in place on the f grid (f holds the elevation on entry) */
void awp_topo_geom_custom(float* f, int nx, int ny, int nz, float h, hipStream_t St);
/* metrics_build_f (src/topography/metrics/metrics.c:35-41, 141-170): the
 * nine interpolants and derivatives of f */
void awp_topo_metrics_f(const f_grid_t* mf, int nx, int ny, float h, hipStream_t St);
/* geom_no_grid_stretching, geometry.c:33-42: g(k) = h k below the top ghost */
void awp_topo_geom_no_stretching(float* g, int nz, float h, hipStream_t St);
/* metrics_build_g (metrics.c:223-229, 293-318): g_c, g3, g3_c from g */
void awp_topo_metrics_g(const g_grid_t* mg, int nz, float h, hipStream_t St);

/* ---- Lagrange-interpolated receivers and sources ---------------------------
 * cu_interp_t, include/interpolation/interpolation.cuh:21-32, restated
 * member for member (prec is float).  cuinterp_htod also builds, library-
 * side, the map the source launchers gather over: the grid points the
 * queries touch, each with its (query, basis point) terms in ascending query
 * order, kept under the struct's d_ix and released by cuinterp_finalize.
 * The launchers run on the default stream, as the reference's do, unless
 * awp_interp_set_stream names another for this struct. */
typedef struct {
  int num_basis; /* degree + 1 */
  int num_query;
  size_t size_l, size_i;
  float *lx, *ly, *lz;      /* host, num_query x num_basis */
  int *ix, *iy, *iz, *ridx; /* host, num_query */
  float *d_lx, *d_ly, *d_lz;
  int *d_ix, *d_iy, *d_iz, *d_ridx;
  grid3_t grid;
} cu_interp_t;

/* no reference counterpart: the stream the launchers of this struct use
 * (after cuinterp_htod; 0 = the default stream).  Returns 0, or 400 if the
 * struct has no map yet. */
int awp_interp_set_stream(const cu_interp_t* interp, hipStream_t stream);
/* no reference counterpart: the number of grid points in the struct's
 * source map (-1 before cuinterp_htod) */
int awp_interp_num_cells(const cu_interp_t* interp);

/* cuinterp_init, src/interpolation/interpolation.cu:13-40: allocate, compute
 * the Lagrange coefficients of num_query query points (qx, qy, qz) on the
 * grid coordinates x, y, z (grid_boundary_size(grid) entries) and upload
 * them.
 * indices become ridx (receivers: output slot; sources: input column).
 * Returns 0 or the reference's error code (800: num_query <= 0,
 * 900: allocation; 400 here also when an index leaves the field). */
int cuinterp_init(cu_interp_t* out, const float* x, const float* y, const float* z,
                  grid3_t grid, const float* qx, const float* qy, const float* qz,
                  const int* indices, int num_query, int degree);
/* cuinterp_lagrange_h, interpolation.cu:42-70 (host only) */
int cuinterp_lagrange_h(cu_interp_t* host, const float* x, const float* y,
                        const float* z, grid3_t grid, const float* qx,
                        const float* qy, const float* qz);
/* cuinterp_htod / cuinterp_dtoh / cuinterp_malloc / cuinterp_finalize,
 * interpolation.cu:72-162.  htod also rebuilds the grid-point map and checks
 * every index against the field (400 if one leaves it). */
int cuinterp_htod(cu_interp_t* interp);
int cuinterp_dtoh(cu_interp_t* interp);
int cuinterp_malloc(cu_interp_t* interp);
void cuinterp_finalize(cu_interp_t* interp);
/* cuinterp_interp_H, interpolation.cu:164-175: out[ridx[q]] = sum over the
 * num_basis^3 stencil of lx ly lz in, one thread per query */
void cuinterp_interp_H(const cu_interp_t* I, float* out, const float* in);
/* cusource_add_cartesian_H, src/topography/sources/source.cu:11-22:
 * out += -dt/h^3 lx ly lz in[ridx[q]] */
void cusource_add_cartesian_H(const cu_interp_t* I, float* out, const float* in,
                              float h, float dt);
/* cusource_add_curvilinear_H, source.cu:51-63: the same divided by the
 * Jacobian f(i, j) dg(k) of the topography metrics (f: the (nx+20) x (ny+84)
 * metric layout, dg: a g-grid array) */
void cusource_add_curvilinear_H(const cu_interp_t* I, float* out, const float* in,
                                float h, float dt, const float* f, int ny, const float* dg);
/* cusource_add_force_H, source.cu:97-113: a body force on a velocity field,
 * scaled by -quad_weight / (f dg rho) with rho from d1 (a 3-D field) */
void cusource_add_force_H(const cu_interp_t* I, float* out, const float* in,
                          const float* d1, float h, float dt, float quad_weight,
                          const float* f, int nx, int ny, int nz, const float* dg);

#ifdef __cplusplus
}
#endif
#endif /* AWP_HIP_H */
