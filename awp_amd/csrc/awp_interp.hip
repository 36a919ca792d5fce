// Lagrange-interpolated receivers and sources (reference
// src/interpolation/lagrange.c, src/interpolation/interpolation.c,
// src/interpolation/interpolation.cu, src/topography/sources/source.cu).
//
// Host: the coefficients restate lagrange_basis (lagrange.c:20-52),
// lagrange_node_poly (:70-77), interp_argnearest (interpolation.c:14-38),
// interp_argnearest_range (:49-63) and interp_lagrange1_coef (:90-128) with
// the same mix of single and double precision, so they equal the reference
// host code's bit for bit.
//
// Device: receivers sample with one lane per query, as the reference (the
// queries are independent).  A source scatters num_basis^3 terms into the
// grid; the reference's lanes add into shared grid points without atomics,
// which races when two stencils overlap.  cuinterp_htod inverts the scatter
// on the host into a per-grid-point list of (query, basis point) terms in
// ascending query order, and the source launchers run one lane per grid point
// that sums its list in that order.  The result is the sequential,
// query-ordered sum, the same on every run, with no atomics.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "awp_common.h"
#include "../../include/awp_hip.h"

namespace {

constexpr double kFltol = 1e-5;  // FLTOL, include/awp/definitions.h:43
// include/awp/error.h:6-35
enum {
  kSuccess = 0,
  kErrGridOutOfBounds = 400,
  kErrLowerExtrapolation = 600,
  kErrUpperExtrapolation = 601,
  kErrOutOfBoundsLower = 700,
  kErrOutOfBoundsUpper = 701,
  kErrNonPositive = 800,
  kErrInterpMalloc = 900,
  kErrInconsistentSize = 902,
};
constexpr int kMaxBasis = 16;  // degree <= 15: keeps the term code in an int
constexpr int kThreads = 256;  // INTERP_THREADS, interpolation.cuh:35 (source launches)
// Receiver launches use 64-lane workgroups: a few thousand queries then
// spread over all 256 CUs instead of a quarter of them.
constexpr int kInterpThreads = 64;

// lagrange_node_poly: prod (x0 - x_i), each difference in single precision,
// the product in double.
double node_poly(float x0, const float* x, int n) {
  double l = 1.0;
  for (int i = 0; i < n; ++i) l *= (double)(x0 - x[i]);
  return l;
}

// lagrange_basis: l_j(x0) = L(x0) lambda_j / (x0 - x_j) with the node
// polynomial L rounded to single precision and the barycentric weight
// lambda_j = 1 / prod_{k != j} (x_j - x_k) in double; l_j = 1 at a node.
int lagrange_basis(float* l, int n, const float* x, float x0) {
  if (n == 1) {
    l[0] = 1.0f;
    return kSuccess;
  }
  if (x0 - x[0] < -kFltol) return kErrLowerExtrapolation;
  if (x0 - x[n - 1] > kFltol) return kErrUpperExtrapolation;
  const float L = (float)node_poly(x0, x, n);
  for (int j = 0; j < n; ++j) {
    l[j] = 1.0f;
    const float dj = x0 - x[j];
    if (fabs((double)dj) < kFltol) continue;
    double denom = 1.0;
    for (int k = 0; k < n; ++k)
      if (k != j) denom *= (double)(x[j] - x[k]);
    const double lambda = 1.0 / denom;
    l[j] = (float)((double)L * lambda / (double)dj);
  }
  return kSuccess;
}

// interp_argnearest: the first closest point; a query outside [p0, p_{n-1}]
// reports the side and the end point.  A NaN query, which the reference
// leaves undefined (its index would be n+1), is reported as below the grid.
int argnearest(int* nearest, const float* p, int n, float q) {
  if (q < p[0] || q != q) {
    *nearest = 0;
    return kErrOutOfBoundsLower;
  }
  if (q > p[n - 1]) {
    *nearest = n - 1;
    return kErrOutOfBoundsUpper;
  }
  int arg = n + 1;
  float dist = INFINITY;
  for (int i = 0; i < n; ++i) {
    const float d = fabsf(q - p[i]);
    if (d < dist) {
      dist = d;
      arg = i;
    }
  }
  *nearest = arg;
  return kSuccess;
}

// interp_argnearest_range: [first, last) holds `lower` points below and
// `upper` above `nearest`, shifted inside [0, n) at the ends.
void argnearest_range(int* first, int* last, int lower, int upper, int nearest, int n) {
  const int dl = nearest - lower, dr = n - nearest - upper - 1;
  *first = std::max(dl, 0) + std::min(dr, 0);
  *last = std::min(nearest + upper + 1, n) - std::min(dl, 0);
}

// interp_lagrange1_coef: stencil of deg+1 points around the nearest point
// (for odd deg the extra point goes to the query's side), first index and
// Lagrange weights at the query clamped into the grid.
int lagrange1_coef(float* xloc, float* l, int* first, const float* x, int n, float query,
                   int deg) {
  const int hi = (int)ceil((double)deg * 0.5), lo = (int)floor((double)deg * 0.5);
  int lower = hi, upper = hi, nearest = 0;
  float q = query;
  int err = argnearest(&nearest, x, n, query);
  if (err == kErrOutOfBoundsLower) q = x[0];
  if (err == kErrOutOfBoundsUpper) q = x[n - 1];
  if (deg % 2 == 1) {
    if (x[nearest] - query > 0) {
      lower = hi;
      upper = lo;
    } else {
      lower = lo;
      upper = hi;
    }
  }
  int last = 0;
  argnearest_range(&lower, &last, lower, upper, nearest, n);
  for (int j = 0; j <= deg; ++j) xloc[j] = x[lower + j];
  *first = lower;
  err |= lagrange_basis(l, deg + 1, xloc, q);
  return err;
}

// What the path reads of a grid3_t, in a compact form for kernel arguments.
struct IGrid {
  int off_x, off_y, off_z;    // offset1
  int line, slice;
  int size_x, size_y, size_z; // grid_boundary_size: coordinate array lengths
  long long numel;            // mem.x * mem.y * mem.z: values in one field
};

// grid_boundary_size (src/grid/grid_3d.c:143-160): a node grid (shift 0)
// with a closed upper boundary drops its last point.
IGrid igrid(const grid3_t& g) {
  IGrid o;
  o.off_x = g.offset1.x;
  o.off_y = g.offset1.y;
  o.off_z = g.offset1.z;
  o.line = g.line;
  o.slice = g.slice;
  o.size_x = g.size.x - (g.shift.x == 0 && g.boundary2.x == 1 ? 1 : 0);
  o.size_y = g.size.y - (g.shift.y == 0 && g.boundary2.y == 1 ? 1 : 0);
  o.size_z = g.size.z - (g.shift.z == 0 && g.boundary2.z == 1 ? 1 : 0);
  o.numel = (long long)g.mem.x * g.mem.y * g.mem.z;
  return o;
}

// grid_index (include/grid/grid_3d.h:113-115) in 64-bit arithmetic
__host__ __device__ inline long long grid_pos(const IGrid& g, int i, int j, int k) {
  return (long long)g.off_z + k + (long long)(g.off_y + j) * g.line +
         (long long)(g.off_x + i) * g.slice;
}

// The source launchers' grid-point map of one cu_interp_t, built by
// cuinterp_htod and kept library-side under the struct's d_ix (a device
// allocation unique to that struct for as long as it lives): the grid points
// the queries touch, each with its (query, basis point) terms in ascending
// query order.
struct SourceMap {
  long long* d_cell_pos = nullptr;
  int* d_cell_ptr = nullptr;   // num_cells + 1
  int* d_cell_term = nullptr;  // ((q*nb + i)*nb + j)*nb + k for basis point (i, j, k)
  int num_cells = 0;
  int ix_max = -1, iy_max = -1, iz_max = -1;  // largest index + num_basis - 1
  hipStream_t stream = nullptr;
};
std::mutex g_map_mutex;
std::map<const void*, SourceMap> g_maps;

void free_map(SourceMap& m) {
  if (m.d_cell_pos) (void)hipFree(m.d_cell_pos);
  if (m.d_cell_ptr) (void)hipFree(m.d_cell_ptr);
  if (m.d_cell_term) (void)hipFree(m.d_cell_term);
  m = SourceMap{};
}

// A copy of the map of I (an empty one with num_cells = -1 if there is none).
SourceMap find_map(const cu_interp_t* I) {
  std::lock_guard<std::mutex> lock(g_map_mutex);
  auto it = g_maps.find(I->d_ix);
  if (it == g_maps.end()) {
    SourceMap none;
    none.num_cells = -1;
    return none;
  }
  return it->second;
}

// Field layouts the curvilinear and force sources read (source.cu:79-80, 131-140):
// metric f(i, j) at (j + align + 2) + (i + 2)*(2 align + 2 ngsl + ny + 4),
// dg(k) at k + align, density (i, j, k) at
// (k + align) + (2 align + nz)*((i + ngsl + 2)*(2 ngsl + ny + 4) + (j + ngsl + 2)),
// with i, j, k the stencil's coordinate-array indices.
struct Jac {
  const float* f;
  const float* dg;
  const float* rho;
  int fline;          // 2*align + 2*ngsl + ny + 4
  int rline;          // 2*align + nz
  long long rslice;   // (2*align + nz) * (2*ngsl + ny + 4)
  float quad_weight;
};

enum { kCartesian = 0, kCurvilinear = 1, kForce = 2 };

// One lane per query.  NB > 0 fixes the number of basis points at compile
// time (the common degrees 0-3), so the whole stencil's loads are in flight
// at once; the sum keeps the reference's i, j, k order either way.
template <int NB>
__global__ __launch_bounds__(kInterpThreads) void interp_kernel(
    float* __restrict__ out, const float* __restrict__ in, const float* __restrict__ lx,
    const float* __restrict__ ly, const float* __restrict__ lz, const int nb_arg,
    const int* __restrict__ ix, const int* __restrict__ iy, const int* __restrict__ iz,
    const int* __restrict__ ridx, const int nq, const IGrid g) {
  const int nb = NB > 0 ? NB : nb_arg;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  float Lx[NB > 0 ? NB : 1], Ly[NB > 0 ? NB : 1], Lz[NB > 0 ? NB : 1];
  if (NB == 4) {
    // 16-byte rows of the coefficient arrays (hipMalloc bases are 256-byte aligned)
    const float4 a = reinterpret_cast<const float4*>(lx)[q];
    const float4 b = reinterpret_cast<const float4*>(ly)[q];
    const float4 c = reinterpret_cast<const float4*>(lz)[q];
    Lx[0] = a.x; Lx[1] = a.y; Lx[2] = a.z; Lx[3] = a.w;
    Ly[0] = b.x; Ly[1] = b.y; Ly[2] = b.z; Ly[3] = b.w;
    Lz[0] = c.x; Lz[1] = c.y; Lz[2] = c.z; Lz[3] = c.w;
  } else if (NB > 0) {
#pragma unroll
    for (int m = 0; m < (NB > 0 ? NB : 1); ++m) {
      Lx[m] = lx[q * NB + m];
      Ly[m] = ly[q * NB + m];
      Lz[m] = lz[q * NB + m];
    }
  }
  const float* Px = lx + (long long)q * nb;
  const float* Py = ly + (long long)q * nb;
  const float* Pz = lz + (long long)q * nb;
  const long long base = grid_pos(g, ix[q], iy[q], iz[q]);
  float acc = 0.0f;
#pragma unroll
  for (int i = 0; i < nb; ++i) {
#pragma unroll
    for (int j = 0; j < nb; ++j) {
      const float* row = in + base + (long long)i * g.slice + (long long)j * g.line;
#pragma unroll
      for (int k = 0; k < nb; ++k) {
        const float w = NB > 0 ? Lx[i] * Ly[j] * Lz[k] : Px[i] * Py[j] * Pz[k];
        acc = acc + w * row[k];
      }
    }
  }
  out[ridx[q]] = acc;
}

// One lane per grid point touched; its terms arrive in ascending query order.
template <int KIND, int NB>
__global__ __launch_bounds__(kThreads) void source_kernel(
    float* __restrict__ out, const float* __restrict__ in, const float* __restrict__ lx,
    const float* __restrict__ ly, const float* __restrict__ lz, const int nb_arg,
    const int* __restrict__ ix, const int* __restrict__ iy, const int* __restrict__ iz,
    const int* __restrict__ lidx, const long long* __restrict__ cell_pos,
    const int* __restrict__ cell_ptr, const int* __restrict__ cell_term, const int ncells,
    const float mdth, const Jac J) {
  const int nb = NB > 0 ? NB : nb_arg;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const long long p = cell_pos[c];
  float acc = out[p];
  const int t1 = cell_ptr[c + 1];
  for (int t = cell_ptr[c]; t < t1; ++t) {
    int e = cell_term[t];
    const int k = e % nb;
    e /= nb;
    const int j = e % nb;
    e /= nb;
    const int i = e % nb;
    const int q = e / nb;
    const long long qb = (long long)q * nb;
    float v = mdth * lx[qb + i] * ly[qb + j] * lz[qb + k] * in[lidx[q]];
    if (KIND != kCartesian) {
      const int gi = ix[q] + i, gj = iy[q] + j, gk = iz[q] + k;
      const float fv = J.f[(long long)gj + AWP_ALIGN + 2 + (long long)(gi + 2) * J.fline];
      const float gv = J.dg[gk + AWP_ALIGN];
      float Ji;
      if (KIND == kCurvilinear) {
        Ji = (float)(1.0 / (double)(fv * gv));
      } else {
        const float rv = J.rho[(long long)gk + AWP_ALIGN +
                               (long long)(gi + AWP_NGSL + 2) * J.rslice +
                               (long long)J.rline * (gj + AWP_NGSL + 2)];
        Ji = -J.quad_weight / (fv * gv * rv);
      }
      v = v * Ji;
    }
    acc = acc + v;
  }
  out[p] = acc;
}

bool launch_ok(const cu_interp_t* I, const void* out, const void* in, const char* where) {
  if (!I || !out || !in) {
    awp_set_error(AWP_ERR_ARG, "%s: null argument", where);
    return false;
  }
  if (I->num_basis < 1 || I->num_basis > kMaxBasis || I->num_query < 1 || !I->d_lx ||
      !I->d_ly || !I->d_lz || !I->d_ix || !I->d_iy || !I->d_iz || !I->d_ridx) {
    awp_set_error(AWP_ERR_STATE, "%s: interpolation not initialised (cuinterp_init)", where);
    return false;
  }
  return true;
}

// Indices of the source map must address the metric, g and density arrays.
bool jac_ok(const SourceMap& M, long long f_rows, int fline, int g_len, const Jac* J, int nx,
            int ny, int nz, const char* where) {
  const bool f_in = M.ix_max + 2 < f_rows && M.iy_max + AWP_ALIGN + 2 < fline;
  const bool g_in = M.iz_max + AWP_ALIGN < g_len;
  bool r_in = true;
  if (J && J->rho)
    r_in = M.ix_max + AWP_NGSL + 2 < nx + 4 + AWP_NGSL2 &&
           M.iy_max + AWP_NGSL + 2 < ny + 4 + AWP_NGSL2 && M.iz_max + AWP_ALIGN < nz + 2 * AWP_ALIGN;
  if (!f_in || !g_in || !r_in) {
    awp_set_error(AWP_ERR_ARG,
                  "%s: stencil indices (up to %d, %d, %d) leave the metric/density arrays", where,
                  M.ix_max, M.iy_max, M.iz_max);
    return false;
  }
  return true;
}

bool map_ok(const SourceMap& M, const char* where) {
  if (M.num_cells < 0 || (M.num_cells > 0 && (!M.d_cell_pos || !M.d_cell_ptr || !M.d_cell_term))) {
    awp_set_error(AWP_ERR_STATE, "%s: no grid-point map (cuinterp_htod)", where);
    return false;
  }
  return true;
}

template <int KIND>
void source_launch(const cu_interp_t* I, const SourceMap& M, float* out, const float* in,
                   float h, float dt, const Jac& J, const char* where) {
  if (M.num_cells == 0) return;
  const float dth = dt / (h * h * h);  // source.cu:38
  const int blocks = (M.num_cells + kThreads - 1) / kThreads;
  auto kern = source_kernel<KIND, 0>;
  switch (I->num_basis) {
    case 1: kern = source_kernel<KIND, 1>; break;
    case 2: kern = source_kernel<KIND, 2>; break;
    case 3: kern = source_kernel<KIND, 3>; break;
    case 4: kern = source_kernel<KIND, 4>; break;
    default: break;
  }
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(kThreads), 0, M.stream, out, in, I->d_lx, I->d_ly,
                     I->d_lz, I->num_basis, I->d_ix, I->d_iy, I->d_iz, I->d_ridx, M.d_cell_pos,
                     M.d_cell_ptr, M.d_cell_term, M.num_cells, -dth, J);
  awp_check_launch(where);
}

}  // namespace

extern "C" {

int cuinterp_malloc(cu_interp_t* I) {
  I->lx = (float*)calloc(I->size_l ? I->size_l : 1, sizeof(float));
  I->ly = (float*)calloc(I->size_l ? I->size_l : 1, sizeof(float));
  I->lz = (float*)calloc(I->size_l ? I->size_l : 1, sizeof(float));
  I->ix = (int*)calloc(I->size_i ? I->size_i : 1, sizeof(int));
  I->iy = (int*)calloc(I->size_i ? I->size_i : 1, sizeof(int));
  I->iz = (int*)calloc(I->size_i ? I->size_i : 1, sizeof(int));
  I->ridx = (int*)calloc(I->size_i ? I->size_i : 1, sizeof(int));
  I->d_lx = I->d_ly = I->d_lz = nullptr;
  I->d_ix = I->d_iy = I->d_iz = I->d_ridx = nullptr;
  if (!I->lx || !I->ly || !I->lz || !I->ix || !I->iy || !I->iz || !I->ridx)
    return kErrInterpMalloc;
  const size_t bl = sizeof(float) * I->size_l, bi = sizeof(int) * I->size_i;
  bool ok = hipMalloc((void**)&I->d_lx, bl) == hipSuccess &&
            hipMalloc((void**)&I->d_ly, bl) == hipSuccess &&
            hipMalloc((void**)&I->d_lz, bl) == hipSuccess &&
            hipMalloc((void**)&I->d_ix, bi) == hipSuccess &&
            hipMalloc((void**)&I->d_iy, bi) == hipSuccess &&
            hipMalloc((void**)&I->d_iz, bi) == hipSuccess &&
            hipMalloc((void**)&I->d_ridx, bi) == hipSuccess;
  if (!ok) {
    awp_set_error(AWP_ERR_HIP, "cuinterp_malloc: device allocation failed");
    return kErrInterpMalloc;
  }
  return kSuccess;
}

int cuinterp_lagrange_h(cu_interp_t* H, const float* x, const float* y, const float* z,
                        grid3_t grid3, const float* qx, const float* qy, const float* qz) {
  const int nb = H->num_basis, deg = nb - 1;
  const IGrid grid = igrid(grid3);
  if (nb < 1 || nb > kMaxBasis) return kErrInconsistentSize;
  if (grid.size_x < nb || grid.size_y < nb || grid.size_z < nb) return kErrInconsistentSize;
  float loc[kMaxBasis];
  int err = kSuccess;
  for (int q = 0; q < H->num_query; ++q) {
    const long long o = (long long)q * nb;
    err = lagrange1_coef(loc, H->lx + o, H->ix + q, x, grid.size_x, qx[q], deg);
    err = lagrange1_coef(loc, H->ly + o, H->iy + q, y, grid.size_y, qy[q], deg);
    err = lagrange1_coef(loc, H->lz + o, H->iz + q, z, grid.size_z, qz[q], deg);
  }
  return err;
}

int cuinterp_htod(cu_interp_t* I) {
  const int nb = I->num_basis, nq = I->num_query;
  const IGrid grid = igrid(I->grid);
  if (!I->d_ix) return kErrInterpMalloc;
  if (nb < 1 || nb > kMaxBasis || nq < 1) return kErrInconsistentSize;
  const long long nb3 = (long long)nb * nb * nb, nterm = nb3 * nq;
  if (nterm > INT_MAX) return kErrInconsistentSize;
  // Every stencil point must lie in the field; build the grid-point map.
  std::vector<std::pair<long long, int>> pt((size_t)nterm);
  int xm = -1, ym = -1, zm = -1;
  for (int q = 0; q < nq; ++q) {
    if (I->ix[q] < 0 || I->iy[q] < 0 || I->iz[q] < 0) return kErrGridOutOfBounds;
    xm = std::max(xm, I->ix[q] + nb - 1);
    ym = std::max(ym, I->iy[q] + nb - 1);
    zm = std::max(zm, I->iz[q] + nb - 1);
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j)
        for (int k = 0; k < nb; ++k) {
          const long long p = grid_pos(grid, I->ix[q] + i, I->iy[q] + j, I->iz[q] + k);
          if (p < 0 || p >= grid.numel) return kErrGridOutOfBounds;
          const int t = ((q * nb + i) * nb + j) * nb + k;
          pt[(size_t)t] = {p, t};
        }
  }
  std::sort(pt.begin(), pt.end());
  std::vector<long long> pos;
  std::vector<int> ptr, term((size_t)nterm);
  for (size_t t = 0; t < pt.size(); ++t) {
    if (t == 0 || pt[t].first != pt[t - 1].first) {
      pos.push_back(pt[t].first);
      ptr.push_back((int)t);
    }
    term[t] = pt[t].second;
  }
  ptr.push_back((int)nterm);

  const size_t bl = sizeof(float) * (size_t)nq * nb, bi = sizeof(int) * (size_t)nq;
  SourceMap M;
  {
    std::lock_guard<std::mutex> lock(g_map_mutex);
    auto it = g_maps.find(I->d_ix);
    if (it != g_maps.end()) {
      M.stream = it->second.stream;
      free_map(it->second);
      g_maps.erase(it);
    }
  }
  const hipStream_t keep_stream = M.stream;
  bool ok = hipMemcpy(I->d_lx, I->lx, bl, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(I->d_ly, I->ly, bl, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(I->d_lz, I->lz, bl, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(I->d_ix, I->ix, bi, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(I->d_iy, I->iy, bi, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(I->d_iz, I->iz, bi, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(I->d_ridx, I->ridx, bi, hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMalloc((void**)&M.d_cell_pos, sizeof(long long) * pos.size()) == hipSuccess &&
       hipMalloc((void**)&M.d_cell_ptr, sizeof(int) * ptr.size()) == hipSuccess &&
       hipMalloc((void**)&M.d_cell_term, sizeof(int) * term.size()) == hipSuccess;
  ok = ok &&
       hipMemcpy(M.d_cell_pos, pos.data(), sizeof(long long) * pos.size(),
                 hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(M.d_cell_ptr, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice) ==
           hipSuccess &&
       hipMemcpy(M.d_cell_term, term.data(), sizeof(int) * term.size(),
                 hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) {
    free_map(M);
    awp_set_error(AWP_ERR_HIP, "cuinterp_htod: device copy failed");
    return kErrInterpMalloc;
  }
  M.num_cells = (int)pos.size();
  M.ix_max = xm;
  M.iy_max = ym;
  M.iz_max = zm;
  M.stream = keep_stream;
  std::lock_guard<std::mutex> lock(g_map_mutex);
  g_maps[I->d_ix] = M;
  return kSuccess;
}

int cuinterp_dtoh(cu_interp_t* I) {
  const size_t bl = sizeof(float) * (size_t)I->num_query * I->num_basis;
  const size_t bi = sizeof(int) * (size_t)I->num_query;
  const bool ok = hipMemcpy(I->lx, I->d_lx, bl, hipMemcpyDeviceToHost) == hipSuccess &&
                  hipMemcpy(I->ly, I->d_ly, bl, hipMemcpyDeviceToHost) == hipSuccess &&
                  hipMemcpy(I->lz, I->d_lz, bl, hipMemcpyDeviceToHost) == hipSuccess &&
                  hipMemcpy(I->ix, I->d_ix, bi, hipMemcpyDeviceToHost) == hipSuccess &&
                  hipMemcpy(I->iy, I->d_iy, bi, hipMemcpyDeviceToHost) == hipSuccess &&
                  hipMemcpy(I->iz, I->d_iz, bi, hipMemcpyDeviceToHost) == hipSuccess &&
                  hipMemcpy(I->ridx, I->d_ridx, bi, hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) {
    awp_set_error(AWP_ERR_HIP, "cuinterp_dtoh: device copy failed");
    return kErrInterpMalloc;
  }
  return kSuccess;
}

int cuinterp_init(cu_interp_t* out, const float* x, const float* y, const float* z,
                  grid3_t grid, const float* qx, const float* qy, const float* qz,
                  const int* indices, int num_query, int degree) {
  out->num_basis = degree + 1;
  out->num_query = num_query;
  out->size_l = (size_t)(num_query > 0 ? num_query : 0) * (size_t)(degree + 1);
  out->size_i = (size_t)(num_query > 0 ? num_query : 0);
  out->grid = grid;
  out->lx = out->ly = out->lz = nullptr;
  out->ix = out->iy = out->iz = out->ridx = nullptr;
  out->d_lx = out->d_ly = out->d_lz = nullptr;
  out->d_ix = out->d_iy = out->d_iz = out->d_ridx = nullptr;
  if (num_query <= 0) return kErrNonPositive;
  if (degree < 0 || degree + 1 > kMaxBasis) return kErrInconsistentSize;
  int err = cuinterp_malloc(out);
  if (err != kSuccess) return err;
  for (int i = 0; i < num_query; ++i) out->ridx[i] = indices[i];
  if (cuinterp_lagrange_h(out, x, y, z, grid, qx, qy, qz) == kErrInconsistentSize)
    return kErrInconsistentSize;
  return cuinterp_htod(out);
}

void cuinterp_finalize(cu_interp_t* I) {
  if (I->d_ix) {
    std::lock_guard<std::mutex> lock(g_map_mutex);
    auto it = g_maps.find(I->d_ix);
    if (it != g_maps.end()) {
      free_map(it->second);
      g_maps.erase(it);
    }
  }
  free(I->lx);
  free(I->ly);
  free(I->lz);
  free(I->ix);
  free(I->iy);
  free(I->iz);
  free(I->ridx);
  I->lx = I->ly = I->lz = nullptr;
  I->ix = I->iy = I->iz = I->ridx = nullptr;
  float* fd[3] = {I->d_lx, I->d_ly, I->d_lz};
  int* id[4] = {I->d_ix, I->d_iy, I->d_iz, I->d_ridx};
  for (float* p : fd)
    if (p) (void)hipFree(p);
  for (int* p : id)
    if (p) (void)hipFree(p);
  I->d_lx = I->d_ly = I->d_lz = nullptr;
  I->d_ix = I->d_iy = I->d_iz = I->d_ridx = nullptr;
}

int awp_interp_num_cells(const cu_interp_t* I) {
  if (!I) return -1;
  return find_map(I).num_cells;
}

int awp_interp_set_stream(const cu_interp_t* I, hipStream_t stream) {
  if (!I) return kErrGridOutOfBounds;
  std::lock_guard<std::mutex> lock(g_map_mutex);
  auto it = g_maps.find(I->d_ix);
  if (it == g_maps.end()) return kErrGridOutOfBounds;
  it->second.stream = stream;
  return kSuccess;
}

void cuinterp_interp_H(const cu_interp_t* I, float* out, const float* in) {
  if (!launch_ok(I, out, in, "cuinterp_interp_H")) return;
  const SourceMap M = find_map(I);
  const int blocks = (I->num_query + kInterpThreads - 1) / kInterpThreads;
  auto kern = interp_kernel<0>;
  switch (I->num_basis) {
    case 1: kern = interp_kernel<1>; break;
    case 2: kern = interp_kernel<2>; break;
    case 3: kern = interp_kernel<3>; break;
    case 4: kern = interp_kernel<4>; break;
    default: break;
  }
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(kInterpThreads), 0, M.stream, out, in, I->d_lx,
                     I->d_ly, I->d_lz, I->num_basis, I->d_ix, I->d_iy, I->d_iz, I->d_ridx,
                     I->num_query, igrid(I->grid));
  awp_check_launch("cuinterp_interp_H");
}

void cusource_add_cartesian_H(const cu_interp_t* I, float* out, const float* in, float h,
                              float dt) {
  const char* where = "cusource_add_cartesian_H";
  if (!launch_ok(I, out, in, where)) return;
  const SourceMap M = find_map(I);
  if (!map_ok(M, where)) return;
  const Jac J = {nullptr, nullptr, nullptr, 0, 0, 0, 0.0f};
  source_launch<kCartesian>(I, M, out, in, h, dt, J, where);
}

void cusource_add_curvilinear_H(const cu_interp_t* I, float* out, const float* in, float h,
                                float dt, const float* f, int ny, const float* dg) {
  const char* where = "cusource_add_curvilinear_H";
  if (!launch_ok(I, out, in, where)) return;
  const SourceMap M = find_map(I);
  if (!map_ok(M, where)) return;
  if (!f || !dg) {
    awp_set_error(AWP_ERR_ARG, "%s: null metric array", where);
    return;
  }
  const int fline = 2 * AWP_ALIGN + AWP_NGSL2 + ny + 4;
  // The metric array has as many rows as the field has slices (nx + 20) and
  // dg as many levels as a field line (nz + 64).
  const IGrid G = igrid(I->grid);
  const long long f_rows = G.slice > 0 ? G.numel / G.slice : 0;
  if (!jac_ok(M, f_rows, fline, G.line, nullptr, 0, 0, 0, where)) return;
  const Jac J = {f, dg, nullptr, fline, 0, 0, 0.0f};
  source_launch<kCurvilinear>(I, M, out, in, h, dt, J, where);
}

void cusource_add_force_H(const cu_interp_t* I, float* out, const float* in, const float* d1,
                          float h, float dt, float quad_weight, const float* f, int nx, int ny,
                          int nz, const float* dg) {
  const char* where = "cusource_add_force_H";
  if (!launch_ok(I, out, in, where)) return;
  const SourceMap M = find_map(I);
  if (!map_ok(M, where)) return;
  if (!f || !dg || !d1) {
    awp_set_error(AWP_ERR_ARG, "%s: null metric or density array", where);
    return;
  }
  const int fline = 2 * AWP_ALIGN + AWP_NGSL2 + ny + 4;
  const int rline = 2 * AWP_ALIGN + nz;
  const Jac J = {f, dg, d1, fline, rline, (long long)rline * (AWP_NGSL2 + ny + 4), quad_weight};
  if (!jac_ok(M, (long long)nx + 4 + AWP_NGSL2, fline, nz + 2 * AWP_ALIGN, &J, nx, ny, nz, where))
    return;
  source_launch<kForce>(I, M, out, in, h, dt, J, where);
}

}  // extern "C"
