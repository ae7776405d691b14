// Matrix-free normal-equations operator, data-row apply/transpose, Jacobi
// diagonal and the preconditioned CG driver (FP64 plain form and FP32
// operator + FP64 refinement form) for the finite-difference interpolator.
//
// Reference path replaced: build_matrix + preprocess_main_system + solve_with_cg
// (LI/_discrete_interpolator.py:684-755,815-833,1221-1319;
// LI/_solver_pipeline.py:86-124; LI/_solver_strategy.py:48-71).  CG on the
// column-scaled system (S A^T A S) y = S A^T b, c = S y, S = diag(A^T A)^-1/2
// is Jacobi-preconditioned CG on N c = A^T b with M^-1 = S^2; the iterates
// and scipy's stopping test |S r| < max(atol, rtol |S A^T b|) are the same.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <hipcub/hipcub.hpp>
#include <mutex>
#include <string>
#include <vector>

#include "fd_operator.hpp"

namespace lsamd {

// ---------------------------------------------------------------- errors --
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}
int injected(const char* site) {
  const char* env = getenv("LSAMD_FAULT_INJECT");
  if (!env || !*env) return 0;
  std::string s(env), key(site);
  size_t pos = 0;
  while (pos <= s.size()) {
    size_t e = s.find(',', pos);
    if (e == std::string::npos) e = s.size();
    if (s.compare(pos, e - pos, key) == 0 || s.compare(pos, e - pos, "all") == 0)
      return fail(LSAMD_ERR_INJECTED, std::string("fault injected at ") + site);
    pos = e + 1;
  }
  return 0;
}
int check_grid(const lsamd_grid* g) {
  if (!g) return fail(LSAMD_ERR_ARG, "grid descriptor is NULL");
  for (int a = 0; a < 3; ++a) {
    if (g->nnodes[a] < 2) return fail(LSAMD_ERR_ARG, "grid needs >= 2 nodes per axis");
    if (!(g->step[a] != 0.0)) return fail(LSAMD_ERR_ARG, "grid step must be non-zero");
  }
  if (g->nnodes[0] * g->nnodes[1] * g->nnodes[2] > (int64_t)1 << 40)
    return fail(LSAMD_ERR_ARG, "grid too large");
  return LSAMD_OK;
}

// ------------------------------------------------------- CG control block --
struct CgCtl {
  double rho, rho_prev, pq_st, pq_rows, thr, bnorm, acc;
  int iter, done, maxiter, converged;
  int cnt_st, cnt_rows, cnt_upd, cnt_misc;
};

// ------------------------------------------------------------- stencil ----
constexpr int TX = 64, TY = 8, NT = TX * TY;
constexpr int PX = TX + 4, PY = TY + 4, PSZ = PX * PY;

template <typename T>
struct StencilArgs {
  GridDev g;
  Composite<T> k;
  RegW<T> w;
  const double* scale2;
  Shell sh;
  int nbx, nby, zc;
  int64_t nbulk;  // number of bulk blocks (shell blocks follow)
};

// y = (R^T W R + ridge2 I) x ; per-block partial of x.y, reduced by the last
// block into ctl->pq_st.  Bulk blocks: 64x8 xy tile marching over zc planes
// with the 25-tap composite, stores only nodes at distance >= 3 from every
// face.  Shell blocks: one node per thread, exact two-pass formula.
template <typename T>
__global__ __launch_bounds__(NT) void k_stencil(const T* __restrict__ x, T* __restrict__ y,
                                                StencilArgs<T> a, double* __restrict__ part,
                                                CgCtl* ctl) {
  if (coherent_load(&ctl->done)) return;
  __shared__ T pl[2][PSZ];
  __shared__ double sh[NT / 64];
  __shared__ int flag;
  const GridDev& g = a.g;
  double dot = 0.0;
  const int64_t bid = blockIdx.x;
  if (bid < a.nbulk) {
    const int bx = (int)(bid % a.nbx);
    const int by = (int)((bid / a.nbx) % a.nby);
    const int bz = (int)(bid / ((int64_t)a.nbx * a.nby));
    const int x0 = bx * TX, y0 = by * TY, z0 = bz * a.zc;
    const int lx = threadIdx.x & (TX - 1), ly = threadIdx.x / TX;
    const int ix = x0 + lx, iy = y0 + ly;
    const bool col_bulk = ix >= 3 && ix <= g.nx - 4 && iy >= 3 && iy <= g.ny - 4;
    const int zend = min(z0 + a.zc, g.nz);
    const Composite<T> K = a.k;
    T xc3 = 0, xc2 = 0, xc1 = 0, xc0 = 0, v2 = 0, v1 = 0, v0 = 0;
    T u4 = 0, u3 = 0, u2 = 0, u1 = 0, u0 = 0;
    for (int p = z0 - 2; p < zend + 2; ++p) {
      T* buf = pl[p & 1];
      const bool pin = p >= 0 && p < g.nz;
      for (int e = threadIdx.x; e < PSZ; e += NT) {
        const int ey = e / PX, ex = e - ey * PX;
        const int gx = x0 - 2 + ex, gy = y0 - 2 + ey;
        T val = (T)0;
        if (pin && gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny) val = x[nid(g, gx, gy, p)];
        buf[e] = val;
      }
      __syncthreads();
      xc3 = xc2;
      xc2 = xc1;
      xc1 = xc0;
      v2 = v1;
      v1 = v0;
      u4 = u3;
      u3 = u2;
      u2 = u1;
      u1 = u0;
      const T* b = buf + (ly + 2) * PX + (lx + 2);
      const T c = b[0];
      const T sx1 = b[1] + b[-1];
      const T sy1 = b[PX] + b[-PX];
      const T sx2 = b[2] + b[-2];
      const T sy2 = b[2 * PX] + b[-2 * PX];
      const T sd = (b[2 * PX + 2] + b[2 * PX - 2]) + (b[-2 * PX + 2] + b[-2 * PX - 2]);
      xc0 = c;
      v0 = K.c0 * c + K.cx1 * sx1 + K.cy1 * sy1 + K.cx2 * sx2 + K.cy2 * sy2 + K.cxy * sd;
      u0 = K.cz2 * c + K.cxz * sx2 + K.cyz * sy2;
      const int kz = p - 2;
      if (col_bulk && kz >= z0 && kz < zend && kz >= 3 && kz <= g.nz - 4) {
        const T yv = v2 + K.cz1 * (xc3 + xc1) + u4 + u0;
        y[nid(g, ix, iy, kz)] = yv;
        dot += (double)xc2 * (double)yv;
      }
    }
  } else {
    const int64_t t = (bid - a.nbulk) * NT + threadIdx.x;
    if (t < a.sh.count) {
      int i = 0, j = 0, k = 0;
      shell_decode(g, a.sh, t, i, j, k);
      const T yv = general_node(x, g, a.w, a.scale2, i, j, k);
      const int64_t id = nid(g, i, j, k);
      y[id] = yv;
      dot += (double)x[id] * (double)yv;
    }
  }
  const double bs = block_sum(dot, sh);
  if (threadIdx.x == 0) part[bid] = bs;
  if (last_block(&ctl->cnt_st, (int)gridDim.x, &flag)) {
    const double tot = reduce_partials(part, (int)gridDim.x, sh);
    if (threadIdx.x == 0) {
      ctl->pq_st = tot;
      ctl->cnt_st = 0;
    }
  }
}

// ----------------------------------------------------------- data rows ----
struct RowsDev {
  int64_t nrows;
  const int64_t* row_ptr;
  const int32_t* cols;
  const void* coef;  // T
  const void* w2;    // T
  int64_t ndn;       // data nodes
  const int64_t* seg_ptr;
  const int32_t* seg_entry;  // entry ids sorted by node (stable)
  const int32_t* seg_row;    // row of each sorted entry
  const int32_t* dnode;      // node id of each segment
};

// t_r = w_r^2 (a_r . x); partial sum of t_r (a_r . x) -> ctl->pq_rows
template <typename T>
__global__ __launch_bounds__(256) void k_rows_t(const T* __restrict__ x, RowsDev R,
                                                T* __restrict__ t, double* __restrict__ part,
                                                CgCtl* ctl) {
  if (coherent_load(&ctl->done)) return;
  __shared__ double sh[4];
  __shared__ int flag;
  const T* coef = (const T*)R.coef;
  const T* w2 = (const T*)R.w2;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double dot = 0.0;
  if (r < R.nrows) {
    T s = 0;
    for (int64_t e = R.row_ptr[r]; e < R.row_ptr[r + 1]; ++e) s += coef[e] * x[R.cols[e]];
    const T tv = w2[r] * s;
    t[r] = tv;
    dot = (double)tv * (double)s;
  }
  const double bs = block_sum(dot, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
  if (last_block(&ctl->cnt_rows, (int)gridDim.x, &flag)) {
    const double tot = reduce_partials(part, (int)gridDim.x, sh);
    if (threadIdx.x == 0) {
      ctl->pq_rows = tot;
      ctl->cnt_rows = 0;
    }
  }
}

// y[node] += sum over the node's entries of coef_e t[row_e]  (node-major,
// fixed order: deterministic, no atomics).  accumulate=0 overwrites.
template <typename T, typename TO>
__global__ __launch_bounds__(256) void k_rows_scatter(RowsDev R, const T* __restrict__ t,
                                                      TO* __restrict__ y, int accumulate,
                                                      CgCtl* ctl) {
  if (ctl && coherent_load(&ctl->done)) return;
  const T* coef = (const T*)R.coef;
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= R.ndn) return;
  T acc = 0;
  for (int64_t s = R.seg_ptr[d]; s < R.seg_ptr[d + 1]; ++s)
    acc += coef[R.seg_entry[s]] * t[R.seg_row[s]];
  const int64_t node = R.dnode[d];
  if (accumulate)
    y[node] += (TO)acc;
  else
    y[node] = (TO)acc;
}

// data-term diagonal: sum over rows of w_r^2 (merged coefficient of node)^2
__global__ void k_rows_diag(RowsDev R, const double* __restrict__ coef,
                            const double* __restrict__ w2, double* __restrict__ diag) {
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= R.ndn) return;
  double acc = 0.0;
  int64_t s = R.seg_ptr[d];
  const int64_t e = R.seg_ptr[d + 1];
  while (s < e) {
    const int32_t row = R.seg_row[s];
    double m = 0.0;
    while (s < e && R.seg_row[s] == row) m += coef[R.seg_entry[s++]];
    acc += w2[row] * m * m;
  }
  diag[R.dnode[d]] += acc;
}

__global__ void k_stencil_diag(GridDev g, RegW<double> w, const double* scale2, double* diag) {
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int64_t nxy = (int64_t)g.nx * g.ny;
  const int k = (int)(id / nxy);
  const int64_t r = id - (int64_t)k * nxy;
  const int j = (int)(r / g.nx);
  const int i = (int)(r - (int64_t)j * g.nx);
  diag[id] = general_diag(g, w, scale2, i, j, k);
}

// per-entry helpers used while building the transpose
__global__ void k_entry_rows(const int64_t* __restrict__ row_ptr, int64_t nrows,
                             int32_t* __restrict__ erow) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  for (int64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) erow[e] = (int32_t)r;
}
__global__ void k_iota(int32_t* v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (int32_t)i;
}
__global__ void k_gather_i32(const int32_t* __restrict__ src, const int32_t* __restrict__ idx,
                             int32_t* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}
__global__ void k_w2(const double* __restrict__ w, double* __restrict__ w2d,
                     float* __restrict__ w2f, const double* __restrict__ coef,
                     float* __restrict__ coeff, int64_t nrows, int64_t nnz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nrows) {
    const double v = w[i] * w[i];
    w2d[i] = v;
    w2f[i] = (float)v;
  }
  if (i < nnz) coeff[i] = (float)coef[i];
}
__global__ void k_w2b(const double* __restrict__ w2, const double* __restrict__ b,
                      double* __restrict__ t, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) t[i] = w2[i] * b[i];
}
__global__ void k_minv(const double* __restrict__ diag, double* __restrict__ m64,
                       float* __restrict__ m32, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = diag[i];
  const double m = d > 0.0 ? 1.0 / d : 1.0;  // S_jj^2, 1 where the column is empty
  m64[i] = m;
  m32[i] = (float)m;
}
__global__ void k_zero_d(double* v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = 0.0;
}

__global__ void k_adjoint(RowsDev R, const double* __restrict__ coef, const double* __restrict__ w,
                          const double* __restrict__ b, const double* __restrict__ c,
                          const double* __restrict__ lam, double* __restrict__ gb,
                          double* __restrict__ gw) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R.nrows) return;
  double u = 0.0, ac = 0.0;
  for (int64_t e = R.row_ptr[r]; e < R.row_ptr[r + 1]; ++e) {
    u += coef[e] * lam[R.cols[e]];
    ac += coef[e] * c[R.cols[e]];
  }
  const double wr = w[r];
  if (gb) gb[r] = wr * wr * u;
  if (gw) gw[r] = 2.0 * wr * u * (b[r] - ac);
}

// ------------------------------------------------------------------ PCG ----
constexpr int VB = 256;  // vector kernel block size

// x = 0 ; r = rhs ; rho = r.(minv r) ; thr = max(atol, rtol sqrt(rho))
template <typename T>
__global__ __launch_bounds__(VB) void k_cg_init(const double* __restrict__ rhs,
                                                const T* __restrict__ minv, T* __restrict__ x,
                                                T* __restrict__ r, int64_t n, double rtol,
                                                double atol, int maxiter, double* part,
                                                CgCtl* ctl) {
  __shared__ double sh[VB / 64];
  __shared__ int flag;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
    const T rv = (T)rhs[i];
    x[i] = (T)0;
    r[i] = rv;
    acc += (double)rv * (double)(minv[i] * rv);
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
  if (last_block(&ctl->cnt_upd, (int)gridDim.x, &flag)) {
    const double tot = reduce_partials(part, (int)gridDim.x, sh);
    if (threadIdx.x == 0) {
      const double bn = sqrt(tot);
      const double thr = fmax(atol, rtol * bn);
      ctl->rho = tot;
      ctl->rho_prev = 1.0;
      ctl->bnorm = bn;
      ctl->thr = thr;
      ctl->iter = 0;
      ctl->maxiter = maxiter;
      ctl->converged = (sqrt(tot) < thr) || bn == 0.0;
      ctl->done = ctl->converged;
      ctl->cnt_upd = 0;
    }
  }
}

// p = minv r + beta p  (beta = rho/rho_prev, p = minv r on the first step)
template <typename T>
__global__ __launch_bounds__(VB) void k_cg_dir(const T* __restrict__ r, const T* __restrict__ minv,
                                               T* __restrict__ p, int64_t n, CgCtl* ctl) {
  if (coherent_load(&ctl->done)) return;
  const int first = coherent_load(&ctl->iter) == 0;
  const T beta = first ? (T)0 : (T)(coherent_load(&ctl->rho) / coherent_load(&ctl->rho_prev));
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
    const T z = minv[i] * r[i];
    p[i] = first ? z : z + beta * p[i];
  }
}

// alpha = rho / p.q ; x += alpha p ; r -= alpha q ; rho' = r.(minv r)
template <typename T>
__global__ __launch_bounds__(VB) void k_cg_upd(T* __restrict__ x, T* __restrict__ r,
                                               const T* __restrict__ p, const T* __restrict__ q,
                                               const T* __restrict__ minv, int64_t n,
                                               double* part, CgCtl* ctl) {
  if (coherent_load(&ctl->done)) return;
  __shared__ double sh[VB / 64];
  __shared__ int flag;
  const double pq = coherent_load(&ctl->pq_st) + coherent_load(&ctl->pq_rows);
  const double rho = coherent_load(&ctl->rho);
  const T alpha = (T)(rho / pq);
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
    x[i] += alpha * p[i];
    const T rv = r[i] - alpha * q[i];
    r[i] = rv;
    acc += (double)rv * (double)(minv[i] * rv);
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
  if (last_block(&ctl->cnt_upd, (int)gridDim.x, &flag)) {
    const double tot = reduce_partials(part, (int)gridDim.x, sh);
    if (threadIdx.x == 0) {
      ctl->rho_prev = rho;
      ctl->rho = tot;
      const int it = ctl->iter + 1;
      ctl->iter = it;
      const int conv = sqrt(tot) < ctl->thr;
      ctl->converged = conv;
      ctl->done = conv || it >= ctl->maxiter;
      ctl->cnt_upd = 0;
    }
  }
}

// r64 = rhs - q64 (q64 = N c) ; |r|_M^2 -> ctl->acc
__global__ __launch_bounds__(VB) void k_residual(const double* __restrict__ rhs,
                                                 const double* __restrict__ q,
                                                 const double* __restrict__ minv,
                                                 double* __restrict__ r, int64_t n,
                                                 double* part, CgCtl* ctl) {
  __shared__ double sh[VB / 64];
  __shared__ int flag;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
    const double rv = rhs[i] - q[i];
    r[i] = rv;
    acc += rv * (minv[i] * rv);
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
  if (last_block(&ctl->cnt_misc, (int)gridDim.x, &flag)) {
    const double tot = reduce_partials(part, (int)gridDim.x, sh);
    if (threadIdx.x == 0) {
      ctl->acc = tot;
      ctl->cnt_misc = 0;
    }
  }
}

__global__ void k_axpy_f32_to_f64(const float* __restrict__ d, double* __restrict__ c, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] += (double)d[i];
}
__global__ void k_copy_d(const double* __restrict__ s, double* __restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = s[i];
}

}  // namespace lsamd

// ================================================================ host side ==
using namespace lsamd;

struct lsamd_fd_system {
  lsamd_grid grid;
  GridDev gd;
  lsamd_fd_reg reg;
  int64_t n = 0;
  double* scale2 = nullptr;
  Shell shell;
  int zc = 8, nbx = 0, nby = 0;
  int64_t nbulk = 0, nst_blocks = 0;
  // data rows
  int64_t nrows = 0, nnz = 0, ndn = 0;
  int64_t* row_ptr = nullptr;
  int32_t* cols = nullptr;
  double* coef64 = nullptr;
  float* coef32 = nullptr;
  double* b = nullptr;
  double* w = nullptr;
  double* w2_64 = nullptr;
  float* w2_32 = nullptr;
  int64_t* seg_ptr = nullptr;
  int32_t* seg_entry = nullptr;
  int32_t* seg_row = nullptr;
  int32_t* dnode = nullptr;
  // derived
  double* rhs = nullptr;
  double* diag = nullptr;
  double* minv64 = nullptr;
  float* minv32 = nullptr;
  // work
  double *x64 = nullptr, *r64 = nullptr, *p64 = nullptr, *q64 = nullptr, *t64 = nullptr;
  float *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *t32 = nullptr;
  double* part = nullptr;
  int64_t part_n = 0;
  double* part_rows = nullptr;
  CgCtl* ctl = nullptr;       // solver control
  CgCtl* ctl_apply = nullptr; // standalone apply (never done)
  CgCtl* ctl_host = nullptr;  // pinned mirror
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int graph_chunk[2] = {0, 0};
  std::vector<void*> owned;
};

namespace {

int dmalloc(lsamd_fd_system* s, void** p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(LSAMD_ERR_ALLOC, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
  s->owned.push_back(*p);
  return LSAMD_OK;
}

template <typename P>
int dalloc(lsamd_fd_system* s, P** p, int64_t count) {
  return dmalloc(s, (void**)p, (size_t)count * sizeof(P));
}

void dfree(lsamd_fd_system* s, void* p) {
  if (!p) return;
  for (size_t i = 0; i < s->owned.size(); ++i)
    if (s->owned[i] == p) {
      hipFree(p);
      s->owned[i] = s->owned.back();
      s->owned.pop_back();
      return;
    }
}

RowsDev rows_dev(const lsamd_fd_system* s, bool fp32) {
  RowsDev R;
  R.nrows = s->nrows;
  R.row_ptr = s->row_ptr;
  R.cols = s->cols;
  R.coef = fp32 ? (const void*)s->coef32 : (const void*)s->coef64;
  R.w2 = fp32 ? (const void*)s->w2_32 : (const void*)s->w2_64;
  R.ndn = s->ndn;
  R.seg_ptr = s->seg_ptr;
  R.seg_entry = s->seg_entry;
  R.seg_row = s->seg_row;
  R.dnode = s->dnode;
  return R;
}

int vec_blocks(int64_t n) {
  int64_t b = ceil_div(n, VB);
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (int)b;
}

template <typename T>
StencilArgs<T> stencil_args(const lsamd_fd_system* s) {
  StencilArgs<T> a;
  a.g = s->gd;
  a.k = make_composite<T>(s->reg);
  a.w = make_regw<T>(s->reg);
  a.scale2 = s->scale2;
  a.sh = s->shell;
  a.nbx = s->nbx;
  a.nby = s->nby;
  a.zc = s->zc;
  a.nbulk = s->nbulk;
  return a;
}

// q = N x (stencil + rows), partial dots into ctl->pq_st / pq_rows
template <typename T>
int launch_apply(lsamd_fd_system* s, const T* x, T* y, T* t, CgCtl* ctl, hipStream_t st) {
  const bool f32 = sizeof(T) == 4;
  RowsDev R = rows_dev(s, f32);
  if (s->nrows > 0) {
    hipLaunchKernelGGL(k_rows_t<T>, dim3((unsigned)ceil_div(s->nrows, 256)), dim3(256), 0, st, x,
                       R, t, s->part_rows, ctl);
  }
  hipLaunchKernelGGL(k_stencil<T>, dim3((unsigned)s->nst_blocks), dim3(NT), 0, st, x, y,
                     stencil_args<T>(s), s->part, ctl);
  if (s->ndn > 0) {
    hipLaunchKernelGGL((k_rows_scatter<T, T>), dim3((unsigned)ceil_div(s->ndn, 256)), dim3(256), 0,
                       st, R, t, y, 1, ctl);
  }
  LSAMD_CHECK_LAUNCH("apply");
  return LSAMD_OK;
}

template <typename T>
int launch_iteration(lsamd_fd_system* s, T* x, T* r, T* p, T* q, T* t, const T* minv, CgCtl* ctl,
                     hipStream_t st) {
  const int vb = vec_blocks(s->n);
  hipLaunchKernelGGL(k_cg_dir<T>, dim3(vb), dim3(VB), 0, st, r, minv, p, s->n, ctl);
  if (int rc = launch_apply<T>(s, p, q, t, ctl, st)) return rc;
  hipLaunchKernelGGL(k_cg_upd<T>, dim3(vb), dim3(VB), 0, st, x, r, p, q, minv, s->n, s->part, ctl);
  LSAMD_CHECK_LAUNCH("cg iteration");
  return LSAMD_OK;
}

int ensure_work(lsamd_fd_system* s, bool f32) {
  const int64_t n = s->n;
  if (!f32 && !s->x64) {
    if (int rc = dalloc(s, &s->x64, n)) return rc;
    if (int rc = dalloc(s, &s->r64, n)) return rc;
    if (int rc = dalloc(s, &s->p64, n)) return rc;
    if (int rc = dalloc(s, &s->q64, n)) return rc;
  }
  if (f32 && !s->x32) {
    if (int rc = dalloc(s, &s->x32, n)) return rc;
    if (int rc = dalloc(s, &s->r32, n)) return rc;
    if (int rc = dalloc(s, &s->p32, n)) return rc;
    if (int rc = dalloc(s, &s->q32, n)) return rc;
  }
  return LSAMD_OK;
}

// Run chunks of `chunk` PCG iterations (captured once into a graph) until
// the device control block reports done.
template <typename T>
int run_pcg(lsamd_fd_system* s, int chunk, T* x, T* r, T* p, T* q, T* t, const T* minv,
            hipStream_t st) {
  const int gi = sizeof(T) == 4 ? 1 : 0;
  if (s->graph[gi] && s->graph_chunk[gi] != chunk) {
    hipGraphExecDestroy(s->graph[gi]);
    s->graph[gi] = nullptr;
  }
  if (!s->graph[gi]) {
    hipStream_t cs;
    LSAMD_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t graph;
    LSAMD_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    int rc = LSAMD_OK;
    for (int i = 0; i < chunk && rc == LSAMD_OK; ++i)
      rc = launch_iteration<T>(s, x, r, p, q, t, minv, s->ctl, cs);
    hipError_t e = hipStreamEndCapture(cs, &graph);
    hipStreamDestroy(cs);
    if (rc) return rc;
    if (e != hipSuccess) return fail(LSAMD_ERR_HIP, "graph capture failed");
    e = hipGraphInstantiate(&s->graph[gi], graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(LSAMD_ERR_HIP, "graph instantiate failed");
    s->graph_chunk[gi] = chunk;
  }
  for (;;) {
    LSAMD_HIP(hipMemcpyAsync(s->ctl_host, s->ctl, sizeof(CgCtl), hipMemcpyDeviceToHost, st));
    LSAMD_HIP(hipStreamSynchronize(st));
    if (s->ctl_host->done) break;
    LSAMD_HIP(hipGraphLaunch(s->graph[gi], st));
  }
  return LSAMD_OK;
}

}  // namespace

extern "C" {

const char* lsamd_last_error(void) { return g_last_error.c_str(); }
int lsamd_version(void) { return 1; }
int lsamd_device_count(int* count) {
  if (!count) return fail(LSAMD_ERR_ARG, "count is NULL");
  LSAMD_HIP(hipGetDeviceCount(count));
  return LSAMD_OK;
}

int lsamd_fd_system_create(const lsamd_grid* g, const lsamd_fd_reg* reg, const double* scale2,
                           lsamd_fd_system** out) {
  if (int rc = injected("system_create")) return rc;
  if (!out || !reg) return fail(LSAMD_ERR_ARG, "lsamd_fd_system_create: NULL argument");
  *out = nullptr;
  if (int rc = check_grid(g)) return rc;
  lsamd_fd_system* s = new lsamd_fd_system();
  s->grid = *g;
  s->gd = to_dev(*g);
  s->reg = *reg;
  s->n = g->nnodes[0] * g->nnodes[1] * g->nnodes[2];
  auto bail = [&](int rc) {
    lsamd_fd_system_destroy(s);
    return rc;
  };
  if (scale2) {
    if (int rc = dalloc(s, &s->scale2, s->n)) return bail(rc);
    LSAMD_HIP(hipMemcpy(s->scale2, scale2, s->n * sizeof(double), hipMemcpyDefault));
  }
  s->shell = make_shell(s->gd, scale2 != nullptr);
  if (!s->shell.all) {
    s->nbx = (int)ceil_div(s->gd.nx, TX);
    s->nby = (int)ceil_div(s->gd.ny, TY);
    s->nbulk = (int64_t)s->nbx * s->nby * ceil_div(s->gd.nz, s->zc);
  }
  s->nst_blocks = s->nbulk + ceil_div(s->shell.count, NT);
  s->part_n = std::max<int64_t>(s->nst_blocks, 8192);
  if (int rc = dalloc(s, &s->part, s->part_n)) return bail(rc);
  if (int rc = dalloc(s, &s->ctl, 1)) return bail(rc);
  if (int rc = dalloc(s, &s->ctl_apply, 1)) return bail(rc);
  LSAMD_HIP(hipMemset(s->ctl, 0, sizeof(CgCtl)));
  LSAMD_HIP(hipMemset(s->ctl_apply, 0, sizeof(CgCtl)));
  if (hipHostMalloc((void**)&s->ctl_host, sizeof(CgCtl), hipHostMallocDefault) != hipSuccess)
    return bail(fail(LSAMD_ERR_ALLOC, "hipHostMalloc failed"));
  if (int rc = dalloc(s, &s->rhs, s->n)) return bail(rc);
  if (int rc = dalloc(s, &s->diag, s->n)) return bail(rc);
  if (int rc = dalloc(s, &s->minv64, s->n)) return bail(rc);
  if (int rc = dalloc(s, &s->minv32, s->n)) return bail(rc);
  // regulariser-only system until rows are set
  hipLaunchKernelGGL(k_stencil_diag, dim3((unsigned)ceil_div(s->n, 256)), dim3(256), 0, 0, s->gd,
                     make_regw<double>(s->reg), s->scale2, s->diag);
  hipLaunchKernelGGL(k_minv, dim3((unsigned)ceil_div(s->n, 256)), dim3(256), 0, 0, s->diag,
                     s->minv64, s->minv32, s->n);
  hipLaunchKernelGGL(k_zero_d, dim3((unsigned)ceil_div(s->n, 256)), dim3(256), 0, 0, s->rhs, s->n);
  LSAMD_CHECK_LAUNCH("system_create");
  LSAMD_HIP(hipDeviceSynchronize());
  *out = s;
  return LSAMD_OK;
}

int lsamd_fd_system_destroy(lsamd_fd_system* s) {
  if (!s) return LSAMD_OK;
  for (int i = 0; i < 2; ++i)
    if (s->graph[i]) hipGraphExecDestroy(s->graph[i]);
  for (void* p : s->owned) hipFree(p);
  if (s->ctl_host) hipHostFree(s->ctl_host);
  delete s;
  return LSAMD_OK;
}

int lsamd_fd_system_set_rows(lsamd_fd_system* s, int64_t nrows, int64_t nnz,
                             const int64_t* row_ptr, const int32_t* cols, const double* coef,
                             const double* b, const double* w, void* stream) {
  if (int rc = injected("set_rows")) return rc;
  if (!s) return fail(LSAMD_ERR_ARG, "lsamd_fd_system_set_rows: NULL system");
  if (nrows < 0 || nnz < 0) return fail(LSAMD_ERR_ARG, "negative row/entry count");
  if (nnz >= 2147483647LL || nrows >= 2147483647LL)
    return fail(LSAMD_ERR_ARG, "too many rows/entries for int32 ids");
  if (nrows > 0 && (!row_ptr || !cols || !coef || !b || !w))
    return fail(LSAMD_ERR_ARG, "lsamd_fd_system_set_rows: NULL row array");
  hipStream_t st = (hipStream_t)stream;
  LSAMD_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < 2; ++i)
    if (s->graph[i]) {
      hipGraphExecDestroy(s->graph[i]);
      s->graph[i] = nullptr;
    }
  void* olds[] = {s->row_ptr, s->cols,  s->coef64,   s->coef32,  s->b,     s->w,
                  s->w2_64,   s->w2_32, s->seg_ptr,  s->seg_entry, s->seg_row, s->dnode,
                  s->t64,     s->t32,   s->part_rows};
  for (void* p : olds) dfree(s, p);
  s->row_ptr = nullptr;
  s->cols = nullptr;
  s->coef64 = nullptr;
  s->coef32 = nullptr;
  s->b = s->w = s->w2_64 = nullptr;
  s->w2_32 = nullptr;
  s->seg_ptr = nullptr;
  s->seg_entry = s->seg_row = s->dnode = nullptr;
  s->t64 = nullptr;
  s->t32 = nullptr;
  s->part_rows = nullptr;
  LSAMD_HIP(hipMemsetAsync(s->ctl, 0, sizeof(CgCtl), st));
  LSAMD_HIP(hipMemsetAsync(s->ctl_apply, 0, sizeof(CgCtl), st));
  s->nrows = nrows;
  s->nnz = nnz;
  s->ndn = 0;
  const int64_t n = s->n;
  // regulariser part of the diagonal, rhs = 0
  hipLaunchKernelGGL(k_stencil_diag, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, s->gd,
                     make_regw<double>(s->reg), s->scale2, s->diag);
  hipLaunchKernelGGL(k_zero_d, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, s->rhs, n);
  if (nrows > 0) {
    if (int rc = dalloc(s, &s->row_ptr, nrows + 1)) return rc;
    if (int rc = dalloc(s, &s->cols, nnz)) return rc;
    if (int rc = dalloc(s, &s->coef64, nnz)) return rc;
    if (int rc = dalloc(s, &s->coef32, nnz)) return rc;
    if (int rc = dalloc(s, &s->b, nrows)) return rc;
    if (int rc = dalloc(s, &s->w, nrows)) return rc;
    if (int rc = dalloc(s, &s->w2_64, nrows)) return rc;
    if (int rc = dalloc(s, &s->w2_32, nrows)) return rc;
    if (int rc = dalloc(s, &s->t64, nrows)) return rc;
    if (int rc = dalloc(s, &s->t32, nrows)) return rc;
    if (int rc = dalloc(s, &s->part_rows, ceil_div(nrows, 256))) return rc;
    LSAMD_HIP(hipMemcpyAsync(s->row_ptr, row_ptr, (nrows + 1) * 8, hipMemcpyDefault, st));
    LSAMD_HIP(hipMemcpyAsync(s->cols, cols, nnz * 4, hipMemcpyDefault, st));
    LSAMD_HIP(hipMemcpyAsync(s->coef64, coef, nnz * 8, hipMemcpyDefault, st));
    LSAMD_HIP(hipMemcpyAsync(s->b, b, nrows * 8, hipMemcpyDefault, st));
    LSAMD_HIP(hipMemcpyAsync(s->w, w, nrows * 8, hipMemcpyDefault, st));
    const int64_t m = std::max(nrows, nnz);
    hipLaunchKernelGGL(k_w2, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, st, s->w, s->w2_64,
                       s->w2_32, s->coef64, s->coef32, nrows, nnz);
    // transpose: stable radix sort of entries by column
    int32_t *erow = nullptr, *keys_out = nullptr, *vals_in = nullptr;
    int32_t* uniq = nullptr;
    int32_t* counts = nullptr;
    int32_t* nrun = nullptr;
    if (int rc = dalloc(s, &erow, nnz)) return rc;
    if (int rc = dalloc(s, &keys_out, nnz)) return rc;
    if (int rc = dalloc(s, &vals_in, nnz)) return rc;
    if (int rc = dalloc(s, &s->seg_entry, nnz)) return rc;
    if (int rc = dalloc(s, &s->seg_row, nnz)) return rc;
    if (int rc = dalloc(s, &uniq, nnz)) return rc;
    if (int rc = dalloc(s, &counts, nnz)) return rc;
    if (int rc = dalloc(s, &nrun, 1)) return rc;
    hipLaunchKernelGGL(k_entry_rows, dim3((unsigned)ceil_div(nrows, 256)), dim3(256), 0, st,
                       s->row_ptr, nrows, erow);
    hipLaunchKernelGGL(k_iota, dim3((unsigned)ceil_div(nnz, 256)), dim3(256), 0, st, vals_in, nnz);
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < n) ++bits;
    size_t tb = 0;
    LSAMD_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, s->cols, keys_out, vals_in,
                                                 s->seg_entry, (int)nnz, 0, bits, st));
    void* tmp = nullptr;
    size_t tb2 = 0;
    LSAMD_HIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, tb2, keys_out, uniq, counts, nrun,
                                                    (int)nnz, st));
    size_t tb3 = 0;
    LSAMD_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, counts, s->seg_row, (int)nnz, st));
    const size_t tbytes = std::max(tb, std::max(tb2, tb3));
    if (int rc = dmalloc(s, &tmp, tbytes)) return rc;
    tb = tbytes;
    LSAMD_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, s->cols, keys_out, vals_in,
                                                 s->seg_entry, (int)nnz, 0, bits, st));
    tb = tbytes;
    LSAMD_HIP(hipcub::DeviceRunLengthEncode::Encode(tmp, tb, keys_out, uniq, counts, nrun,
                                                    (int)nnz, st));
    int32_t nr = 0;
    LSAMD_HIP(hipMemcpyAsync(&nr, nrun, 4, hipMemcpyDeviceToHost, st));
    LSAMD_HIP(hipStreamSynchronize(st));
    s->ndn = nr;
    if (int rc = dalloc(s, &s->seg_ptr, nr + 1)) return rc;
    if (int rc = dalloc(s, &s->dnode, nr)) return rc;
    // segment offsets (int32 scan, then widened)
    int32_t* off32 = nullptr;
    if (int rc = dalloc(s, &off32, nr + 1)) return rc;
    tb = tbytes;
    LSAMD_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, counts, off32, (int)nr, st));
    std::vector<int32_t> hoff(nr + 1);
    LSAMD_HIP(hipMemcpyAsync(hoff.data(), off32, nr * 4, hipMemcpyDeviceToHost, st));
    LSAMD_HIP(hipStreamSynchronize(st));
    hoff[nr] = (int32_t)nnz;
    std::vector<int64_t> hoff64(hoff.begin(), hoff.end());
    LSAMD_HIP(hipMemcpy(s->seg_ptr, hoff64.data(), (nr + 1) * 8, hipMemcpyHostToDevice));
    LSAMD_HIP(hipMemcpyAsync(s->dnode, uniq, nr * 4, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_gather_i32, dim3((unsigned)ceil_div(nnz, 256)), dim3(256), 0, st, erow,
                       s->seg_entry, s->seg_row, nnz);
    LSAMD_CHECK_LAUNCH("transpose");
    LSAMD_HIP(hipStreamSynchronize(st));
    void* tmps[] = {erow, keys_out, vals_in, uniq, counts, nrun, tmp, off32};
    for (void* p : tmps) dfree(s, p);
    // rhs = A^T W^2 b, data diagonal
    RowsDev R = rows_dev(s, false);
    hipLaunchKernelGGL(k_w2b, dim3((unsigned)ceil_div(nrows, 256)), dim3(256), 0, st, s->w2_64,
                       s->b, s->t64, nrows);
    hipLaunchKernelGGL((k_rows_scatter<double, double>), dim3((unsigned)ceil_div(s->ndn, 256)),
                       dim3(256), 0, st, R, s->t64, s->rhs, 1, (CgCtl*)nullptr);
    hipLaunchKernelGGL(k_rows_diag, dim3((unsigned)ceil_div(s->ndn, 256)), dim3(256), 0, st, R,
                       s->coef64, s->w2_64, s->diag);
  }
  hipLaunchKernelGGL(k_minv, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, s->diag,
                     s->minv64, s->minv32, n);
  LSAMD_CHECK_LAUNCH("set_rows");
  LSAMD_HIP(hipStreamSynchronize(st));
  return LSAMD_OK;
}

int lsamd_fd_system_apply(lsamd_fd_system* s, int32_t precision, const void* x, void* y,
                          void* stream) {
  if (int rc = injected("apply")) return rc;
  if (!s || !x || !y) return fail(LSAMD_ERR_ARG, "lsamd_fd_system_apply: NULL argument");
  hipStream_t st = (hipStream_t)stream;
  if (precision == LSAMD_FP64) {
    if (int rc = launch_apply<double>(s, (const double*)x, (double*)y, s->t64, s->ctl_apply, st))
      return rc;
  } else if (precision == LSAMD_FP32) {
    if (int rc = launch_apply<float>(s, (const float*)x, (float*)y, s->t32, s->ctl_apply, st))
      return rc;
  } else {
    return fail(LSAMD_ERR_ARG, "unknown precision");
  }
  return LSAMD_OK;
}

int lsamd_fd_system_rhs(lsamd_fd_system* s, double* rhs, void* stream) {
  if (!s || !rhs) return fail(LSAMD_ERR_ARG, "lsamd_fd_system_rhs: NULL argument");
  LSAMD_HIP(hipMemcpyAsync(rhs, s->rhs, s->n * 8, hipMemcpyDefault, (hipStream_t)stream));
  return LSAMD_OK;
}

int lsamd_fd_system_diag(lsamd_fd_system* s, double* diag, void* stream) {
  if (!s || !diag) return fail(LSAMD_ERR_ARG, "lsamd_fd_system_diag: NULL argument");
  LSAMD_HIP(hipMemcpyAsync(diag, s->diag, s->n * 8, hipMemcpyDefault, (hipStream_t)stream));
  return LSAMD_OK;
}

int lsamd_fd_system_solve(lsamd_fd_system* s, const lsamd_solve_opts* o, const double* rhs,
                          double* c, lsamd_solve_info* info, void* stream) {
  if (int rc = injected("solve")) return rc;
  if (!s || !o || !c) return fail(LSAMD_ERR_ARG, "lsamd_fd_system_solve: NULL argument");
  if (!(o->rtol >= 0.0) || !(o->atol >= 0.0))
    return fail(LSAMD_ERR_ARG, "rtol/atol must be non-negative");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = s->n;
  const double* b = rhs ? rhs : s->rhs;
  const int64_t maxiter64 = o->maxiter > 0 ? o->maxiter : 10 * n;
  const int maxiter = (int)std::min<int64_t>(maxiter64, 2000000000LL);
  const int chunk = o->check_every > 0 ? o->check_every : 64;
  const int vb = vec_blocks(n);
  lsamd_solve_info inf;
  memset(&inf, 0, sizeof(inf));
  if (o->precision == LSAMD_FP64) {
    if (int rc = ensure_work(s, false)) return rc;
    hipLaunchKernelGGL(k_cg_init<double>, dim3(vb), dim3(VB), 0, st, b, s->minv64, s->x64,
                       s->r64, n, o->rtol, o->atol, maxiter, s->part, s->ctl);
    LSAMD_CHECK_LAUNCH("cg init");
    if (int rc = run_pcg<double>(s, chunk, s->x64, s->r64, s->p64, s->q64, s->t64, s->minv64, st))
      return rc;
    hipLaunchKernelGGL(k_copy_d, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, s->x64, c, n);
    LSAMD_HIP(hipStreamSynchronize(st));
    inf.iterations = s->ctl_host->iter;
    inf.converged = s->ctl_host->converged;
    inf.residual = sqrt(s->ctl_host->rho);
    inf.threshold = s->ctl_host->thr;
  } else if (o->precision == LSAMD_FP32) {
    // FP64 outer refinement around FP32 inner PCG solves of N d = r.
    if (int rc = ensure_work(s, false)) return rc;
    if (int rc = ensure_work(s, true)) return rc;
    const double inner = o->inner_rtol > 0 ? o->inner_rtol : 1e-4;
    const int64_t max_outer = o->max_outer > 0 ? o->max_outer : 50;
    hipLaunchKernelGGL(k_zero_d, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, c, n);
    // threshold from |S rhs| (same test as the FP64 form)
    hipLaunchKernelGGL(k_cg_init<double>, dim3(vb), dim3(VB), 0, st, b, s->minv64, s->x64, s->r64,
                       n, o->rtol, o->atol, maxiter, s->part, s->ctl);
    LSAMD_HIP(hipMemcpyAsync(s->ctl_host, s->ctl, sizeof(CgCtl), hipMemcpyDeviceToHost, st));
    LSAMD_HIP(hipStreamSynchronize(st));
    const double thr = s->ctl_host->thr;
    double res = sqrt(s->ctl_host->rho);
    int64_t total = 0, outer = 0;
    int conv = res < thr || s->ctl_host->bnorm == 0.0;
    while (!conv && outer < max_outer && total < maxiter64) {
      // inner: r64 already holds the residual; run FP32 PCG on it
      const int left = (int)std::min<int64_t>(maxiter64 - total, 2000000000LL);
      const double irtol = inner;
      hipLaunchKernelGGL(k_cg_init<float>, dim3(vb), dim3(VB), 0, st, s->r64, s->minv32, s->x32,
                         s->r32, n, irtol, 0.0, left, s->part, s->ctl);
      LSAMD_CHECK_LAUNCH("inner init");
      if (int rc = run_pcg<float>(s, chunk, s->x32, s->r32, s->p32, s->q32, s->t32, s->minv32, st))
        return rc;
      total += s->ctl_host->iter;
      ++outer;
      hipLaunchKernelGGL(k_axpy_f32_to_f64, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st,
                         s->x32, c, n);
      // r64 = rhs - N c in FP64
      if (int rc = launch_apply<double>(s, c, s->q64, s->t64, s->ctl_apply, st)) return rc;
      hipLaunchKernelGGL(k_residual, dim3(vb), dim3(VB), 0, st, b, s->q64, s->minv64, s->r64, n,
                         s->part, s->ctl_apply);
      LSAMD_HIP(hipMemcpyAsync(s->ctl_host, s->ctl_apply, sizeof(CgCtl), hipMemcpyDeviceToHost,
                               st));
      LSAMD_HIP(hipStreamSynchronize(st));
      res = sqrt(s->ctl_host->acc);
      conv = res < thr;
    }
    inf.iterations = total;
    inf.outer = outer;
    inf.converged = conv;
    inf.residual = res;
    inf.threshold = thr;
  } else {
    return fail(LSAMD_ERR_ARG, "unknown precision");
  }
  if (info) *info = inf;
  return LSAMD_OK;
}

int lsamd_rows_adjoint(lsamd_fd_system* s, const double* c, const double* lambda, double* g_b,
                       double* g_w, void* stream) {
  if (int rc = injected("adjoint")) return rc;
  if (!s || !c || !lambda) return fail(LSAMD_ERR_ARG, "lsamd_rows_adjoint: NULL argument");
  if (s->nrows == 0) return LSAMD_OK;
  RowsDev R = rows_dev(s, false);
  hipLaunchKernelGGL(k_adjoint, dim3((unsigned)ceil_div(s->nrows, 256)), dim3(256), 0,
                     (hipStream_t)stream, R, s->coef64, s->w, s->b, c, lambda, g_b, g_w);
  LSAMD_CHECK_LAUNCH("k_adjoint");
  return LSAMD_OK;
}

}  // extern "C"
