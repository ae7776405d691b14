// Point location, least-squares row builders and evaluation on a structured
// grid.  This translation unit is compiled with -ffp-contract=off (and the
// pragma below) so that every product, quotient and sum rounds exactly like
// the reference's numpy expressions: cell ids, trilinear weights, shape
// gradients, row norms and evaluated values are bit-identical.
//
// Reference: LC/supports/_3d_base_structured.py:251-293 (cell lookup,
// inside test), :386-412 (corner ids), LC/supports/_3d_structured_grid.py:
// 90-146 (trilinear, local coordinates), :295-391 (evaluation), :393-454
// (shape gradients); LI/_discrete_interpolator.py:361-366 (row
// normalisation); LI/_finite_difference_interpolator.py:236-269 (value),
// :342-401 (gradient), :403-476 (normal), :478-539 (gradient orthogonal);
// LC/math/_maths.py:126-160,276-332 (strike/dip basis).
#pragma clang fp contract(off)

#include <math.h>

#include "lsamd_common.hpp"

namespace lsamd {
namespace {

constexpr double kEps32 = 2.220446049250313e-16 * 32.0;  // np.finfo(float).eps * 32

// numpy's np.clip(v, 0, 1): keeps -0.0 and NaN.
__device__ inline double clip01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

// numpy pairwise order for a contiguous 8-term sum.
__device__ inline double sum8(const double* a) {
  return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

struct Located {
  bool inside;
  int cx, cy, cz;
  double lx, ly, lz;
  double hx, hy, hz;  // size of the cell along each axis
};

__device__ inline bool inside_axis(double p, double o, double h, int nn) {
  const double mx = o + (double)(nn - 1) * h;
  double ext = fabs(mx - o);
  ext = ext > 1.0 ? ext : 1.0;  // np.maximum(|max-origin|, 1.0)
  const double tol = kEps32 * ext;
  return (p >= (o - tol)) && (p <= (mx + tol));
}

__device__ inline int cell_axis(double p, double o, double h, int nn) {
  const double x = p - o;
  long long c = (long long)floor(x / h);
  const long long hi = nn - 2;  // ncells - 1
  if (c < 0) c = 0;
  if (c > hi) c = hi;
  return (int)c;
}

// Rectilinear axis (LC/supports/_3d_rectilinear_grid.py:115-159): cell =
// clip(searchsorted(nodes, p, 'right') - 1, 0, ncells - 1), local coordinate
// (p - x0) / h unclipped with h = x1 - x0 (1 where 0), inside strictly
// between the first and last node.
__device__ inline bool rect_axis(const double* nodes, int nn, double p, int* cell, double* loc,
                                 double* h) {
  int lo = 0, hi = nn;
  while (lo < hi) {  // number of nodes <= p
    const int mid = (lo + hi) >> 1;
    if (nodes[mid] <= p)
      lo = mid + 1;
    else
      hi = mid;
  }
  int c = lo - 1;
  if (c < 0) c = 0;
  if (c > nn - 2) c = nn - 2;
  const double x0 = nodes[c];
  const double x1 = nodes[c + 1 < nn ? c + 1 : nn - 1];
  double w = x1 - x0;
  if (w == 0.0) w = 1.0;
  *cell = c;
  *loc = (p - x0) / w;
  *h = w;
  return (p > nodes[0]) && (p < nodes[nn - 1]);
}

__device__ inline Located locate(const GridDev& g, double px, double py, double pz) {
  Located L;
  if (g.xn) {
    const bool ix = rect_axis(g.xn, g.nx, px, &L.cx, &L.lx, &L.hx);
    const bool iy = rect_axis(g.yn, g.ny, py, &L.cy, &L.ly, &L.hy);
    const bool iz = rect_axis(g.zn, g.nz, pz, &L.cz, &L.lz, &L.hz);
    L.inside = ix && iy && iz;
    return L;
  }
  L.hx = g.hx;
  L.hy = g.hy;
  L.hz = g.hz;
  L.inside = inside_axis(px, g.ox, g.hx, g.nx) && inside_axis(py, g.oy, g.hy, g.ny) &&
             inside_axis(pz, g.oz, g.hz, g.nz);
  L.cx = L.cy = L.cz = 0;
  if (L.inside) {
    L.cx = cell_axis(px, g.ox, g.hx, g.nx);
    L.cy = cell_axis(py, g.oy, g.hy, g.ny);
    L.cz = cell_axis(pz, g.oz, g.hz, g.nz);
  }
  // local coordinates: clip((p - (o + h*cell)) / h, 0, 1)
  L.lx = clip01((px - (g.ox + g.hx * (double)L.cx)) / g.hx);
  L.ly = clip01((py - (g.oy + g.hy * (double)L.cy)) / g.hy);
  L.lz = clip01((pz - (g.oz + g.hz * (double)L.cz)) / g.hz);
  return L;
}

__device__ inline void corner_ids(const GridDev& g, const Located& L, int64_t* id) {
  const int64_t nx = g.nx, nxy = (int64_t)g.nx * g.ny;
  const int64_t g0 = (int64_t)L.cx + nx * L.cy + nxy * L.cz;
  id[0] = g0;
  id[1] = g0 + 1;
  id[2] = g0 + nx;
  id[3] = g0 + nx + 1;
  id[4] = g0 + nxy;
  id[5] = g0 + nxy + 1;
  id[6] = g0 + nxy + nx;
  id[7] = g0 + nxy + nx + 1;
}

// _3d_structured_grid.py:114-123 (same association: left to right)
__device__ inline void trilinear(double x, double y, double z, double* w) {
  w[0] = (1 - x) * (1 - y) * (1 - z);
  w[1] = x * (1 - y) * (1 - z);
  w[2] = (1 - x) * y * (1 - z);
  w[4] = (1 - x) * (1 - y) * z;
  w[5] = x * (1 - y) * z;
  w[6] = (1 - x) * y * z;
  w[3] = x * y * (1 - z);
  w[7] = x * y * z;
}

// _3d_structured_grid.py:425-453, T[d*8 + k]
__device__ inline void shape_grad(double x, double y, double z, double hx, double hy, double hz,
                                  double* T) {
  T[0] = (1 - z) * (y - 1);
  T[1] = (1 - y) * (1 - z);
  T[2] = -y * (1 - z);
  T[4] = -(1 - y) * z;
  T[5] = (1 - y) * z;
  T[6] = -y * z;
  T[3] = y * (1 - z);
  T[7] = y * z;
  T[8 + 0] = (x - 1) * (1 - z);
  T[8 + 1] = -x * (1 - z);
  T[8 + 2] = (1 - x) * (1 - z);
  T[8 + 4] = -(1 - x) * z;
  T[8 + 5] = -x * z;
  T[8 + 6] = (1 - x) * z;
  T[8 + 3] = x * (1 - z);
  T[8 + 7] = x * z;
  T[16 + 0] = -(1 - x) * (1 - y);
  T[16 + 1] = -x * (1 - y);
  T[16 + 2] = -(1 - x) * y;
  T[16 + 4] = (1 - x) * (1 - y);
  T[16 + 5] = x * (1 - y);
  T[16 + 6] = (1 - x) * y;
  T[16 + 3] = -x * y;
  T[16 + 7] = x * y;
#pragma unroll
  for (int k = 0; k < 8; ++k) T[k] /= hx;
#pragma unroll
  for (int k = 0; k < 8; ++k) T[8 + k] /= hy;
#pragma unroll
  for (int k = 0; k < 8; ++k) T[16 + k] /= hz;
}

__global__ void k_locate(GridDev g, const double* __restrict__ pts, int64_t n,
                         int64_t* __restrict__ idc, double* __restrict__ w, double* __restrict__ T,
                         uint8_t* __restrict__ ins) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
  Located L = locate(g, px, py, pz);
  if (idc) {
    int64_t id[8];
    corner_ids(g, L, id);
#pragma unroll
    for (int k = 0; k < 8; ++k) idc[8 * i + k] = L.inside ? id[k] : -1;
  }
  if (w) {
    double a[8];
    trilinear(L.lx, L.ly, L.lz, a);
#pragma unroll
    for (int k = 0; k < 8; ++k) w[8 * i + k] = L.inside ? a[k] : 0.0;
  }
  if (T) {
    double t[24];
    shape_grad(L.lx, L.ly, L.lz, L.hx, L.hy, L.hz, t);
#pragma unroll
    for (int k = 0; k < 24; ++k) T[24 * i + k] = L.inside ? t[k] : 0.0;
  }
  if (ins) ins[i] = L.inside ? 1 : 0;
}

// np.linalg.norm of a 3-vector: sequential sum of squares (n < 8).
__device__ inline double norm3(double a, double b, double c) {
  return sqrt((a * a + b * b) + c * c);
}

// Row normalisation, LI/_discrete_interpolator.py:361-366.
__device__ inline double normalise8(double* a, double* b) {
  double sq[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) sq[k] = a[k] * a[k];
  const double len = sqrt(sum8(sq));
  if (len > 0) {
    *b /= len;
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] /= len;
  }
  return len;
}

__device__ inline void emit_row(int64_t row, const int32_t* id, const double* a, double bv,
                                double wv, double len, int32_t* cols, double* coef, double* b,
                                double* w, double* norm) {
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    cols[8 * row + k] = id[k];
    coef[8 * row + k] = a[k];
  }
  b[row] = bv;
  w[row] = wv;
  if (norm) norm[row] = len;
}

// get_vectors (LC/math/_maths.py:276-298) for an already unit vector v:
// strike = [sin(-s), -cos(-s), 0] with s = deg2rad(rad2deg(-atan2(ny, nx))),
// dip = strike x v; both normalised, then scaled by |v|.
__device__ inline void strike_dip(double* v, double* s, double* d) {
  const double len = norm3(v[0], v[1], v[2]);
  v[0] /= len;
  v[1] /= len;
  v[2] /= len;
  double nv0 = v[0], nv1 = v[1], nv2 = v[2];
  const double l2 = norm3(nv0, nv1, nv2);
  nv0 /= l2;
  nv1 /= l2;
  nv2 /= l2;
  (void)nv2;
  const double strike_deg = (-atan2(nv1, nv0)) * (180.0 / M_PI);
  const double st = strike_deg * (M_PI / 180.0);
  s[0] = sin(-st);
  s[1] = -cos(-st);
  s[2] = 0.0;
  const double ls = norm3(s[0], s[1], s[2]);
  s[0] /= ls;
  s[1] /= ls;
  s[2] /= ls;
  d[0] = s[1] * v[2] - s[2] * v[1];
  d[1] = s[2] * v[0] - s[0] * v[2];
  d[2] = s[0] * v[1] - s[1] * v[0];
  const double ld = norm3(d[0], d[1], d[2]);
  d[0] /= ld;
  d[1] /= ld;
  d[2] /= ld;
  for (int k = 0; k < 3; ++k) {
    s[k] *= len;
    d[k] *= len;
  }
}

__global__ void k_build_rows(GridDev g, int kind, double* __restrict__ pts, int64_t n, int64_t ld,
                             int wcol, double wscale, int32_t* __restrict__ cols,
                             double* __restrict__ coef, double* __restrict__ b,
                             double* __restrict__ w, double* __restrict__ norm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* P = pts + i * ld;
  const Located L = locate(g, P[0], P[1], P[2]);
  int64_t id64[8];
  corner_ids(g, L, id64);
  int32_t id[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) id[k] = (int32_t)id64[k];
  const double pw = wcol >= 0 ? P[wcol] : 1.0;
  if (kind == LSAMD_ROWS_VALUE) {
    double a[8];
    trilinear(L.lx, L.ly, L.lz, a);
    double bv = P[3];
    const double len = normalise8(a, &bv);
    emit_row(i, id, a, bv, wscale * pw, len, cols, coef, b, w, norm);
    return;
  }
  double T[24];
  shape_grad(L.lx, L.ly, L.lz, L.hx, L.hy, L.hz, T);
  if (kind == LSAMD_ROWS_NORMAL) {
    for (int d = 0; d < 3; ++d) {
      double a[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] = T[8 * d + k];
      double bv = P[3 + d];
      const double len = normalise8(a, &bv);
      emit_row((int64_t)d * n + i, id, a, bv, wscale * pw, len, cols, coef, b, w, norm);
    }
    return;
  }
  // gradient / tangent: normalise the stored vector in place, scale T by 1/|v|
  double v[3] = {P[3], P[4], P[5]};
  const double vn = norm3(v[0], v[1], v[2]);
  if (kind == LSAMD_ROWS_GRADIENT || vn > 0) {
    for (int k = 0; k < 3; ++k) v[k] /= vn;
    for (int k = 0; k < 24; ++k) T[k] /= vn;
  }
  P[3] = v[0];
  P[4] = v[1];
  P[5] = v[2];
  if (kind == LSAMD_ROWS_TANGENT) {
    double a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = (v[0] * T[k] + v[1] * T[8 + k]) + v[2] * T[16 + k];
    double bv = 0.0;
    const double len = normalise8(a, &bv);
    emit_row(i, id, a, bv, wscale, len, cols, coef, b, w, norm);
    return;
  }
  double s[3], dv[3];
  double vv[3] = {v[0], v[1], v[2]};
  strike_dip(vv, s, dv);
  double a[8], c2[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    a[k] = (s[0] * T[k] + s[1] * T[8 + k]) + s[2] * T[16 + k];
    c2[k] = (dv[0] * T[k] + dv[1] * T[8 + k]) + dv[2] * T[16 + k];
  }
  double b0 = 0.0, b1 = 0.0;
  const double l0 = normalise8(a, &b0);
  const double l1 = normalise8(c2, &b1);
  emit_row(i, id, a, b0, wscale, l0, cols, coef, b, w, norm);
  emit_row(n + i, id, c2, b1, wscale, l1, cols, coef, b, w, norm);
}

__global__ void k_evaluate(GridDev g, const double* __restrict__ c, const double* __restrict__ pts,
                           int64_t n, double* __restrict__ vals, double* __restrict__ grads) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
  const Located L = locate(g, px, py, pz);
  int64_t id[8];
  corner_ids(g, L, id);
  double cv[8];
  if (L.inside) {
#pragma unroll
    for (int k = 0; k < 8; ++k) cv[k] = c[id[k]];
  }
  if (vals) {
    double out;
    if (isnan(px) || isnan(py) || isnan(pz)) {
      out = 0.0;
    } else if (!L.inside) {
      out = __builtin_nan("");
    } else {
      double a[8];
      trilinear(L.lx, L.ly, L.lz, a);
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] *= cv[k];
      out = sum8(a);
    }
    vals[i] = out;
  }
  if (grads) {
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (L.inside) {
      double T[24];
      shape_grad(L.lx, L.ly, L.lz, L.hx, L.hy, L.hz, T);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        T[k] *= cv[k];
        T[8 + k] *= cv[k];
        T[16 + k] *= cv[k];
      }
      gx = sum8(T);
      gy = sum8(T + 8);
      gz = sum8(T + 16);
    }
    grads[3 * i] = gx;
    grads[3 * i + 1] = gy;
    grads[3 * i + 2] = gz;
  }
}

// numpy's pairwise summation of a contiguous run of n <= 128 doubles
// (n < 8: sequential from 0; else eight interleaved accumulators).
__device__ inline double np_sum(const double* a, int n) {
  if (n < 8) {
    double r = 0.0;
    for (int i = 0; i < n; ++i) r += a[i];
    return r;
  }
  double r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = a[j];
  int i = 8;
  for (; i < n - (n % 8); i += 8)
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] += a[i + j];
  double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; ++i) res += a[i];
  return res;
}

__global__ void k_rows_normalise(int64_t nrows, int k, double* __restrict__ coef,
                                 double* __restrict__ b, double* __restrict__ norm) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  double sq[128];
  double* a = coef + r * k;
  for (int j = 0; j < k; ++j) sq[j] = a[j] * a[j];
  const double len = sqrt(np_sum(sq, k));
  if (len > 0) {
    b[r] /= len;
    for (int j = 0; j < k; ++j) a[j] /= len;
  }
  if (norm) norm[r] = len;
}

// (node, point entry, w_k * gv) for the 8 corners of every point; outside
// points give zero contributions to node 0.
__global__ void k_eval_entries(GridDev g, const double* __restrict__ pts,
                               const double* __restrict__ gv, int64_t n,
                               int32_t* __restrict__ key, int32_t* __restrict__ idx,
                               double* __restrict__ val) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Located L = locate(g, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  int64_t id[8];
  corner_ids(g, L, id);
  double a[8];
  trilinear(L.lx, L.ly, L.lz, a);
  for (int k = 0; k < 8; ++k) {
    const int64_t e = 8 * i + k;
    key[e] = L.inside ? (int32_t)id[k] : 0;
    idx[e] = (int32_t)e;
    val[e] = L.inside ? a[k] * gv[i] : 0.0;
  }
}

}  // namespace

void launch_eval_entries(const GridDev& g, const double* pts, const double* gv, int64_t n,
                         int32_t* key, int32_t* idx, double* val, hipStream_t st) {
  hipLaunchKernelGGL(k_eval_entries, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, g, pts,
                     gv, n, key, idx, val);
}

}  // namespace lsamd

using namespace lsamd;

extern "C" int lsamd_rows_normalise(int64_t nrows, int32_t k, double* coef, double* b,
                                    double* norm, void* stream) {
  if (int rc = injected("rows_normalise")) return rc;
  if (nrows < 0 || k < 1 || k > 128)
    return fail(LSAMD_ERR_ARG, "lsamd_rows_normalise: need nrows >= 0 and 1 <= k <= 128");
  if (nrows == 0) return LSAMD_OK;
  if (!coef || !b) return fail(LSAMD_ERR_ARG, "lsamd_rows_normalise: NULL input");
  const int bs = 64;
  hipLaunchKernelGGL(k_rows_normalise, dim3((unsigned)ceil_div(nrows, bs)), dim3(bs), 0,
                     (hipStream_t)stream, nrows, (int)k, coef, b, norm);
  LSAMD_CHECK_LAUNCH("k_rows_normalise");
  return LSAMD_OK;
}

extern "C" int lsamd_fd_locate(const lsamd_grid* g, const double* pts, int64_t n, int64_t* idc,
                               double* w, double* T, uint8_t* inside, void* stream) {
  if (int rc = injected("locate")) return rc;
  if (int rc = check_grid(g)) return rc;
  if (n < 0) return fail(LSAMD_ERR_ARG, "lsamd_fd_locate: n < 0");
  if (n == 0) return LSAMD_OK;
  if (!pts) return fail(LSAMD_ERR_ARG, "lsamd_fd_locate: pts is NULL");
  const int bs = 256;
  hipLaunchKernelGGL(k_locate, dim3((unsigned)ceil_div(n, bs)), dim3(bs), 0, (hipStream_t)stream,
                     to_dev(*g), pts, n, idc, w, T, inside);
  LSAMD_CHECK_LAUNCH("k_locate");
  return LSAMD_OK;
}

extern "C" int lsamd_fd_build_rows(const lsamd_grid* g, int32_t kind, double* pts, int64_t n,
                                   int64_t ld, int32_t weight_col, double wscale, int32_t* cols,
                                   double* coef, double* b, double* w, double* norm,
                                   void* stream) {
  if (int rc = injected("build_rows")) return rc;
  if (int rc = check_grid(g)) return rc;
  if (kind < LSAMD_ROWS_VALUE || kind > LSAMD_ROWS_TANGENT)
    return fail(LSAMD_ERR_ARG, "lsamd_fd_build_rows: unknown row kind");
  const int need = (kind == LSAMD_ROWS_VALUE) ? 4 : 6;
  if (ld < need) return fail(LSAMD_ERR_ARG, "lsamd_fd_build_rows: ld too small for row kind");
  if (weight_col >= ld) return fail(LSAMD_ERR_ARG, "lsamd_fd_build_rows: weight_col >= ld");
  if (n < 0) return fail(LSAMD_ERR_ARG, "lsamd_fd_build_rows: n < 0");
  if (n == 0) return LSAMD_OK;
  if (!pts || !cols || !coef || !b || !w)
    return fail(LSAMD_ERR_ARG, "lsamd_fd_build_rows: NULL output");
  const int64_t nn = g->nnodes[0] * g->nnodes[1] * g->nnodes[2];
  if (nn > 2147483647LL)
    return fail(LSAMD_ERR_ARG, "lsamd_fd_build_rows: grid exceeds int32 node ids");
  const int bs = 128;
  hipLaunchKernelGGL(k_build_rows, dim3((unsigned)ceil_div(n, bs)), dim3(bs), 0,
                     (hipStream_t)stream, to_dev(*g), (int)kind, pts, n, ld, (int)weight_col,
                     wscale, cols, coef, b, w, norm);
  LSAMD_CHECK_LAUNCH("k_build_rows");
  return LSAMD_OK;
}

extern "C" int lsamd_fd_evaluate(const lsamd_grid* g, const double* c, const double* pts,
                                 int64_t n, double* values, double* grads, void* stream) {
  if (int rc = injected("evaluate")) return rc;
  if (int rc = check_grid(g)) return rc;
  if (n < 0) return fail(LSAMD_ERR_ARG, "lsamd_fd_evaluate: n < 0");
  if (n == 0) return LSAMD_OK;
  if (!c || !pts) return fail(LSAMD_ERR_ARG, "lsamd_fd_evaluate: NULL input");
  const int bs = 256;
  hipLaunchKernelGGL(k_evaluate, dim3((unsigned)ceil_div(n, bs)), dim3(bs), 0,
                     (hipStream_t)stream, to_dev(*g), c, pts, n, values, grads);
  LSAMD_CHECK_LAUNCH("k_evaluate");
  return LSAMD_OK;
}
