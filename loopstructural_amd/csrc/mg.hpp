// Geometric multigrid preconditioner and flexible PCG for one FD system.
//
// Hierarchy: level 0 is the system itself; level l+1 is the grid with every
// second node of level l (cells halved, rounded up, so the coarse grid may
// reach one fine cell beyond the fine one), step doubled.  Prolongation P is
// trilinear ([1/2, 1, 1/2] per axis), restriction is P^T.  Coarse operators:
// the data term is the exact Galerkin product (A_d P): each data row is
// mapped through P and its repeated coarse columns merged; the regulariser is
// re-discretised on the coarse grid with every squared weight scaled by
// 2^-l (a smooth field's second-difference energy halves per coarsening).
// The coarsest level (<= 5 nodes per axis) is solved with a dense inverse.
// Smoother: Chebyshev polynomial (degree 2 on the fine level, 4 on the
// coarse ones) in M N on
// [lambda_max/30, 1.15 lambda_max], lambda_max by power iteration, with M
// additive Schwarz: the exact inverse of N on each block, summed over
// blocks, point Jacobi on the nodes no block covers.  Blocks are the 8 nodes
// of every cell that holds a data row, and the boundary layers: per axis the
// first three and the last three nodes form one group (a short axis of < 6
// nodes is one group: an axis of 6 gives two groups of 3, whose blocks of
// 27 invert in LDS where one block of 216 took 35 ms in global memory),
// every other index is its own group, and each product
// of groups with at least one boundary group is a block.  Point Jacobi alone
// leaves the modes that those strong local couplings (data rows of weight
// ~1 against regulariser rows of ~0.1; border rows of weight dx) do not see:
// they are oscillatory and of tiny Rayleigh quotient, so neither the
// smoother nor the coarse grid reduces them.  The V-cycle runs in FP32; the
// outer flexible CG (Polak-Ribiere beta) runs in FP64 with the same FP64
// stopping test as the Jacobi form.
#pragma once

namespace lsamd {

constexpr int kMgDegMax = 8;  // largest Chebyshev degree (LSAMD_MG_DEGREE)
constexpr int kMgCoarse = 5;  // stop coarsening at <= 5 nodes per axis

// --------------------------------------------------------------- kernels ----
__global__ void k_mg_to_f32(const double* __restrict__ a, float* __restrict__ b, int64_t n,
                            const CgCtl* ctl) {
  if (ctl->done) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
// Smoother preconditioner M applied to r = b - q (r = b when from_zero) at
// node i.  Tag sm[i]: >= 0, the Jacobi scale of an uncovered node; -1, a
// node covered by Schwarz blocks, whose row of summed block inverses is
// sval[sptr[i] .. sptr[i+1]) over the columns scol.
struct MgSm {
  const float* sm;
  const int64_t* sptr;
  const int32_t* scol;
  const float* sval;
};

__device__ inline float mg_r(const float* __restrict__ b, const float* __restrict__ q,
                             int64_t j, int from_zero) {
  return from_zero ? b[j] : b[j] - q[j];
}

__device__ inline float mg_precond(int64_t i, const MgSm& S, const float* __restrict__ b,
                                   const float* __restrict__ q, int from_zero) {
  const float t = S.sm[i];
  if (t >= 0.f) return t * mg_r(b, q, i, from_zero);
  // four entries per step: independent loads in flight hide the latency of
  // the column-indexed gathers
  float acc = 0.f;
  int64_t e = S.sptr[i];
  const int64_t e1 = S.sptr[i + 1];
  for (; e + 4 <= e1; e += 4) {
    const int32_t j0 = S.scol[e], j1 = S.scol[e + 1], j2 = S.scol[e + 2], j3 = S.scol[e + 3];
    acc += S.sval[e] * mg_r(b, q, j0, from_zero) + S.sval[e + 1] * mg_r(b, q, j1, from_zero) +
           S.sval[e + 2] * mg_r(b, q, j2, from_zero) + S.sval[e + 3] * mg_r(b, q, j3, from_zero);
  }
  for (; e < e1; ++e) acc += S.sval[e] * mg_r(b, q, S.scol[e], from_zero);
  return acc;
}

// skip_cov: leave the covered nodes to k_mg_cheb_cov (large levels, where
// their gathers would stall every wavefront that holds one)
__global__ void k_mg_cheb_first(const float* __restrict__ b, const float* __restrict__ q,
                                MgSm S, float* __restrict__ x, float* __restrict__ d, int64_t n,
                                float c0, int from_zero, int skip_cov, const CgCtl* ctl) {
  if (ctl->done) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (skip_cov && S.sm[i] < 0.f) return;
  const float dv = c0 * mg_precond(i, S, b, q, from_zero);
  d[i] = dv;
  x[i] = from_zero ? dv : x[i] + dv;
}

// later steps: d = c1 d + c2 M (b - q), x += d; last: the smoothing's final
// step, whose d is dead and not stored
__global__ void k_mg_cheb_step(const float* __restrict__ b, const float* __restrict__ q,
                               MgSm S, float* __restrict__ x, float* __restrict__ d, int64_t n,
                               float c1, float c2, int skip_cov, int last, const CgCtl* ctl) {
  if (ctl->done) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (skip_cov && S.sm[i] < 0.f) return;
  const float dv = c1 * d[i] + c2 * mg_precond(i, S, b, q, 0);
  if (!last) d[i] = dv;
  x[i] += dv;
}

// the covered nodes of a level (compact list): mode 0 first step from zero,
// 1 first step, 2 later step (d = c1 d + c2 M r); last as k_mg_cheb_step
__global__ void k_mg_cheb_cov(const int32_t* __restrict__ cov, int64_t ncov,
                              const float* __restrict__ b, const float* __restrict__ q, MgSm S,
                              float* __restrict__ x, float* __restrict__ d, int mode, float c0,
                              float c1, float c2, int last, const CgCtl* ctl) {
  if (ctl->done) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncov) return;
  const int64_t i = cov[t];
  const float z = mg_precond(i, S, b, q, mode == 0);
  const float dv = mode == 2 ? c1 * d[i] + c2 * z : c0 * z;
  if (!last) d[i] = dv;
  x[i] = mode == 0 ? dv : x[i] + dv;
}

__host__ __device__ inline void mg_ijk(int64_t id, int nx, int ny, int& i, int& j, int& k) {
  const int64_t nxy = (int64_t)nx * ny;
  k = (int)(id / nxy);
  const int64_t r = id - (int64_t)k * nxy;
  j = (int)(r / nx);
  i = (int)(r - (int64_t)j * nx);
}

// bc = P^T (bf - qf): per coarse node, the fine nodes 2J-1, 2J, 2J+1 per axis
// with weights 1/2, 1, 1/2 (those inside the fine grid)
// P^T (bf - qf) at coarse node id: the fine nodes 2J-1, 2J, 2J+1 per axis
// with weights 1/2, 1, 1/2 (those inside the fine grid)
__device__ inline float mg_restrict_node(const float* __restrict__ bf,
                                         const float* __restrict__ qf, const GridDev& gf,
                                         const GridDev& gc, int64_t id) {
  int I, J, K;
  mg_ijk(id, gc.nx, gc.ny, I, J, K);
  const int64_t fnxy = (int64_t)gf.nx * gf.ny;
  float acc = 0.f;
  for (int dk = -1; dk <= 1; ++dk) {
    const int k = 2 * K + dk;
    if (k < 0 || k >= gf.nz) continue;
    const float wk = dk == 0 ? 1.f : 0.5f;
    for (int dj = -1; dj <= 1; ++dj) {
      const int j = 2 * J + dj;
      if (j < 0 || j >= gf.ny) continue;
      const float wj = wk * (dj == 0 ? 1.f : 0.5f);
      for (int di = -1; di <= 1; ++di) {
        const int i = 2 * I + di;
        if (i < 0 || i >= gf.nx) continue;
        const float w = wj * (di == 0 ? 1.f : 0.5f);
        const int64_t f = (int64_t)i + (int64_t)j * gf.nx + (int64_t)k * fnxy;
        acc += w * (bf[f] - qf[f]);
      }
    }
  }
  return acc;
}

// (P xc) at fine node id: trilinear, coarse node c = fine node 2c
__device__ inline float mg_prolong_node(const float* __restrict__ xc, const GridDev& gc,
                                        const GridDev& gf, int64_t id) {
  int i, j, k;
  mg_ijk(id, gf.nx, gf.ny, i, j, k);
  const int i0 = i >> 1, j0 = j >> 1, k0 = k >> 1;
  const int ni = (i & 1) ? 2 : 1, nj = (j & 1) ? 2 : 1, nk = (k & 1) ? 2 : 1;
  const float wi = ni == 2 ? 0.5f : 1.f, wj = nj == 2 ? 0.5f : 1.f, wk = nk == 2 ? 0.5f : 1.f;
  const int64_t cnxy = (int64_t)gc.nx * gc.ny;
  float acc = 0.f;
  for (int c = 0; c < nk; ++c)
    for (int b = 0; b < nj; ++b)
      for (int a = 0; a < ni; ++a)
        acc += xc[(int64_t)(i0 + a) + (int64_t)(j0 + b) * gc.nx + (int64_t)(k0 + c) * cnxy];
  return wi * wj * wk * acc;
}

__global__ void k_mg_restrict(const float* __restrict__ bf, const float* __restrict__ qf,
                              GridDev gf, float* __restrict__ bc, GridDev gc, const CgCtl* ctl) {
  if (ctl->done) return;
  const int64_t nc = (int64_t)gc.nx * gc.ny * gc.nz;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nc) return;
  bc[id] = mg_restrict_node(bf, qf, gf, gc, id);
}
__global__ void k_mg_prolong(const float* __restrict__ xc, GridDev gc, float* __restrict__ xf,
                             GridDev gf, const CgCtl* ctl) {
  if (ctl->done) return;
  const int64_t nf = (int64_t)gf.nx * gf.ny * gf.nz;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nf) return;
  xf[id] += mg_prolong_node(xc, gc, gf, id);
}

__global__ __launch_bounds__(64) void k_mg_dense(const float* __restrict__ cinv,
                                                 const float* __restrict__ b,
                                                 float* __restrict__ x, int n, const CgCtl* ctl) {
  if (ctl->done) return;
  const int r = blockIdx.x;
  float acc = 0.f;
  for (int c = threadIdx.x; c < n; c += 64) acc += cinv[(int64_t)r * n + c] * b[c];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (threadIdx.x == 0) x[r] = acc;
}

// Gram matrix of the data rows of a coarse level, G = A^T W^2 A, as sorted
// (row << 32 | col, value) pairs: every row of width wc emits wc^2 pairs.
__global__ void k_gram_pairs(const int32_t* __restrict__ cols, const double* __restrict__ coef,
                             const double* __restrict__ w2, int64_t nrows, int wc,
                             uint64_t* __restrict__ key, double* __restrict__ val) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const int64_t e0 = r * wc;
  const double w = w2[r];
  for (int a = 0; a < wc; ++a)
    for (int b = 0; b < wc; ++b) {
      const int64_t o = e0 * wc + (int64_t)a * wc + b;
      key[o] = ((uint64_t)(uint32_t)cols[e0 + a] << 32) | (uint32_t)cols[e0 + b];
      val[o] = w * coef[e0 + a] * coef[e0 + b];
    }
}

// CSR row pointers of the reduced Gram entries (keys sorted): first entry of
// each node row by binary search
__global__ void k_gram_ptr(const uint64_t* __restrict__ ukey, int64_t m, int64_t n,
                           int64_t* __restrict__ ptr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)(ukey[mid] >> 32) < i)
      lo = mid + 1;
    else
      hi = mid;
  }
  ptr[i] = lo;
}

__global__ void k_gram_cols(const uint64_t* __restrict__ ukey, const double* __restrict__ v,
                            int64_t m, int32_t* __restrict__ col, float* __restrict__ v32,
                            double* __restrict__ v64) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  col[e] = (int32_t)(ukey[e] & 0xffffffffu);
  v32[e] = (float)v[e];
  v64[e] = v[e];
}

// y += G x, one node per thread
template <typename T>
__global__ void k_gram_spmv(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                            const T* __restrict__ val, const T* __restrict__ x,
                            T* __restrict__ y, int64_t n, const CgCtl* ctl) {
  if (ctl->done) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T acc = 0;
  int64_t e = ptr[i];
  const int64_t e1 = ptr[i + 1];
  for (; e + 4 <= e1; e += 4) {
    const int32_t c0 = col[e], c1 = col[e + 1], c2 = col[e + 2], c3 = col[e + 3];
    acc += val[e] * x[c0] + val[e + 1] * x[c1] + val[e + 2] * x[c2] + val[e + 3] * x[c3];
  }
  for (; e < e1; ++e) acc += val[e] * x[col[e]];
  y[i] += acc;
}

// diag += diagonal of G; then the Jacobi inverses
__global__ void k_gram_diag(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                            const double* __restrict__ val, double* __restrict__ diag,
                            double* __restrict__ m64, float* __restrict__ m32, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = diag[i];
  for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e)
    if (col[e] == (int32_t)i) d += val[e];
  diag[i] = d;
  const double m = d > 0.0 ? 1.0 / d : 1.0;
  m64[i] = m;
  m32[i] = (float)m;
}

// Coarse data rows A P: row r of width wf (columns on grid gf) mapped through
// P onto grid gc, repeated coarse columns merged.  count_only: write the
// number of distinct coarse columns; else write them (padded to wc with zero
// coefficients on the first column).
__global__ void k_mg_rows(const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ cols,
                          const double* __restrict__ coef, int64_t nrows, GridDev gf, GridDev gc,
                          int count_only, int wc, int32_t* __restrict__ ccount,
                          int32_t* __restrict__ ccols, double* __restrict__ ccoef) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  constexpr int kMax = 64;
  int32_t uc[kMax];
  double ua[kMax];
  int nu = 0;
  bool overflow = false;
  const int64_t cnxy = (int64_t)gc.nx * gc.ny;
  for (int64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
    int i, j, k;
    mg_ijk(cols[e], gf.nx, gf.ny, i, j, k);
    const int ni = (i & 1) ? 2 : 1, nj = (j & 1) ? 2 : 1, nk = (k & 1) ? 2 : 1;
    const double w = coef[e] * (ni == 2 ? 0.5 : 1.0) * (nj == 2 ? 0.5 : 1.0) * (nk == 2 ? 0.5 : 1.0);
    for (int c = 0; c < nk; ++c)
      for (int b = 0; b < nj; ++b)
        for (int a = 0; a < ni; ++a) {
          const int32_t col = (int32_t)((i >> 1) + a + (int64_t)((j >> 1) + b) * gc.nx +
                                        (int64_t)((k >> 1) + c) * cnxy);
          int u = 0;
          while (u < nu && uc[u] != col) ++u;
          if (u == nu) {
            if (nu == kMax) {
              overflow = true;
              continue;
            }
            uc[nu] = col;
            ua[nu++] = 0.0;
          }
          ua[u] += w;
        }
  }
  if (count_only) {
    ccount[r] = overflow ? kMax + 1 : nu;
    return;
  }
  for (int u = 0; u < wc; ++u) {
    ccols[r * wc + u] = u < nu ? uc[u] : uc[0];
    ccoef[r * wc + u] = u < nu ? ua[u] : 0.0;
  }
}

// power iteration step: w = M y (y = N v), partial |w|^2 and v.w
__global__ __launch_bounds__(256) void k_mg_power(const double* __restrict__ y,
                                                  const double* __restrict__ minv,
                                                  const float* __restrict__ sm,
                                                  const int64_t* __restrict__ sptr,
                                                  const int32_t* __restrict__ scol,
                                                  const double* __restrict__ sval,
                                                  const double* __restrict__ v,
                                                  double* __restrict__ w, int64_t n,
                                                  double* __restrict__ part) {
  __shared__ double sh[4];
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double wv;
    if (sm[i] >= 0.f) {
      wv = minv[i] * y[i];
    } else {
      wv = 0.0;
      for (int64_t e = sptr[i]; e < sptr[i + 1]; ++e) wv += sval[e] * y[scol[e]];
    }
    w[i] = wv;
    a += wv * wv;
    b += v[i] * wv;
  }
  const double sa = block_sum(a, sh);
  const double sb = block_sum(b, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sa;
    part[2 * blockIdx.x + 1] = sb;
  }
}

// one block: scal[0] = |w| (the eigenvalue estimate), scal[1] = 1/|w|
__global__ __launch_bounds__(256) void k_mg_norm_fin(const double* __restrict__ part, int np,
                                                     double* __restrict__ scal) {
  __shared__ double sh[4];
  double a = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) a += part[2 * i];
  const double sa = block_sum(a, sh);
  if (threadIdx.x == 0) {
    const double wn = sqrt(sa);
    scal[0] = wn;
    scal[1] = wn > 0.0 ? 1.0 / wn : 0.0;
  }
}

__global__ void k_mg_scale(const double* __restrict__ w, double* __restrict__ v, int64_t n,
                           const double* __restrict__ scal) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = w[i] * scal[1];
}

// deterministic positive start vector of the power iteration
__global__ void k_mg_start(double* __restrict__ v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL;
  z ^= z >> 31;
  z *= 0xBF58476D1CE4E5B9ULL;
  z ^= z >> 29;
  v[i] = 0.5 + (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// rptr[r] = r * wc: row pointers of fixed-width rows
__global__ void k_iota_rows(int64_t* __restrict__ rptr, int64_t n, int64_t wc) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) rptr[r] = r * wc;
}

// ------------------------------------------------------ flexible PCG ----
// partials of z.r, z.r_prev (z: the FP32 V-cycle output, widened)
__global__ __launch_bounds__(256) void k_fcg_dots(const float* __restrict__ z,
                                                  const double* __restrict__ r,
                                                  const double* __restrict__ rp, int64_t n,
                                                  double* __restrict__ part, const CgCtl* ctl) {
  if (ctl->done) return;
  __shared__ double sh[4];
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double zi = (double)z[i];
    a += zi * r[i];
    b += zi * rp[i];
  }
  const double sa = block_sum(a, sh);
  const double sb = block_sum(b, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sa;
    part[2 * blockIdx.x + 1] = sb;
  }
}

// one block: rho = z.r; beta = (z.r - z.r_prev) / rho_prev (0 on the first step)
__global__ __launch_bounds__(256) void k_fcg_dots_fin(const double* __restrict__ part, int np,
                                                      CgCtl* ctl) {
  if (ctl->done) return;
  __shared__ double sh[4];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  const double sa = block_sum(a, sh);
  const double sb = block_sum(b, sh);
  if (threadIdx.x == 0) {
    const double beta = ctl->iter == 0 ? 0.0 : (sa - sb) / ctl->rho_prev;
    ctl->rho = sa;
    ctl->acc = beta;
  }
}

// p = z + beta p
__global__ void k_fcg_dir(const float* __restrict__ z, double* __restrict__ p, int64_t n,
                          const CgCtl* ctl) {
  if (ctl->done) return;
  const double beta = ctl->acc;
  const int first = ctl->iter == 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    p[i] = first ? (double)z[i] : (double)z[i] + beta * p[i];
}

// x += alpha p ; r_prev = r ; r -= alpha q ; rf = (float) r, the next
// V-cycle's input ; partials of r.(minv r)
__global__ __launch_bounds__(256) void k_fcg_upd(double* __restrict__ x, double* __restrict__ r,
                                                 double* __restrict__ rp,
                                                 const double* __restrict__ p,
                                                 const double* __restrict__ q,
                                                 const double* __restrict__ minv,
                                                 float* __restrict__ rf, int64_t n,
                                                 double* __restrict__ part, const CgCtl* ctl) {
  if (ctl->done) return;
  __shared__ double sh[4];
  const double alpha = ctl->alpha;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    x[i] += alpha * p[i];
    const double rv = r[i];
    rp[i] = rv;
    const double rn = rv - alpha * q[i];
    r[i] = rn;
    rf[i] = (float)rn;
    acc += rn * (minv[i] * rn);
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// one block: rho_prev = rho ; iter++ ; stopping test on |S r|
__global__ __launch_bounds__(256) void k_fcg_upd_fin(const double* __restrict__ part, int np,
                                                     CgCtl* ctl) {
  if (ctl->done) return;
  __shared__ double sh[4];
  const double tot = reduce_partials(part, np, sh);
  if (threadIdx.x == 0) {
    ctl->rho_prev = ctl->rho;
    const int it = ctl->iter + 1;
    ctl->iter = it;
    ctl->res2 = tot;  // |S r|^2 of the current iterate
    const int conv = sqrt(tot) < ctl->thr;
    ctl->converged = conv;
    ctl->done = conv || it >= ctl->maxiter;
  }
}


// ----------------------------------------------- coarse levels, fused ----
// Device view of one coarse level's operator: regulariser (class table, or
// the exact per-node formula when the table does not apply) + data Gram
// matrix.
struct MgDev {
  GridDev g;
  int64_t n;
  const float* tab;  // FP32 class table, nullptr: per-node formula with w
  RegW<float> w;
  const int64_t* gptr;
  const int32_t* gcol;
  const float* gval;
};

__device__ inline float mg_apply_node(const MgDev& L, const float* __restrict__ x, int64_t id) {
  int i, j, k;
  mg_ijk(id, L.g.nx, L.g.ny, i, j, k);
  float acc;
  if (L.tab) {
    const int nx = L.g.nx, ny = L.g.ny, nz = L.g.nz;
    const int64_t nxy = (int64_t)nx * ny;
    const float* c = L.tab + ((cls7(k, nz) * 7 + cls7(j, ny)) * 7 + cls7(i, nx)) * kTaps;
    acc = c[0] * x[id];
#pragma unroll
    for (int t = 1; t < kTaps; ++t) {
      const int dx = kBTap[t][0], dy = kBTap[t][1], dz = kBTap[t][2];
      const bool ok = (unsigned)(i + dx) < (unsigned)nx && (unsigned)(j + dy) < (unsigned)ny &&
                      (unsigned)(k + dz) < (unsigned)nz;
      const float xv = x[ok ? id + dx + (int64_t)nx * dy + nxy * dz : id];
      acc += c[t] * (ok ? xv : 0.f);
    }
  } else {
    acc = general_node<float>(x, L.g, L.w, nullptr, i, j, k);
  }
  if (L.gptr) {
    int64_t e = L.gptr[id];
    const int64_t e1 = L.gptr[id + 1];
    for (; e + 4 <= e1; e += 4) {
      const int32_t c0 = L.gcol[e], c1 = L.gcol[e + 1], c2 = L.gcol[e + 2], c3 = L.gcol[e + 3];
      acc += L.gval[e] * x[c0] + L.gval[e + 1] * x[c1] + L.gval[e + 2] * x[c2] +
             L.gval[e + 3] * x[c3];
    }
    for (; e < e1; ++e) acc += L.gval[e] * x[L.gcol[e]];
  }
  return acc;
}

// y = N x of a coarse level in one launch (stencil and Gram row per node)
__global__ __launch_bounds__(256) void k_mg_apply(MgDev L, const float* __restrict__ x,
                                                  float* __restrict__ y, const CgCtl* ctl) {
  if (ctl->done) return;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id < L.n) y[id] = mg_apply_node(L, x, id);
}

constexpr int64_t kSplitMin = 1 << 20;  // levels this large smooth covered nodes apart
// levels this large keep the compact list of their covered nodes (the
// multi-rhs smoother splits them, the single-system one from kSplitMin)
constexpr int64_t kCovListMin = 1 << 17;

// ------------------------------------------------ Schwarz smoother blocks ----
// Entry N(a, b) of a level's operator: regulariser rows (from their
// definition, ridge included) plus the data Gram matrix G (CSR, columns
// sorted per row; may be absent).
__host__ __device__ inline double reg_entry(const GridDev& g, const RegW<double>& w, const double* scale2,
                                   int64_t ia, int64_t ib) {
  int ai, aj, ak, bi, bj, bk;
  mg_ijk(ia, g.nx, g.ny, ai, aj, ak);
  mg_ijk(ib, g.nx, g.ny, bi, bj, bk);
  double v = ia == ib ? w.ridge : 0.0;
  // axial masks [1, -2, 1] / sqrt6: product of two taps / 6
  auto m = [](int o) { return o == 0 ? -2.0 : 1.0; };
  if (aj == bj && ak == bk)
    for (int c = (ai > bi ? ai : bi) - 1; c <= (ai < bi ? ai : bi) + 1; ++c)
      v += row_weight(g, w, scale2, F_XX, c, aj, ak) * m(ai - c) * m(bi - c) / 6.0;
  if (ai == bi && ak == bk)
    for (int c = (aj > bj ? aj : bj) - 1; c <= (aj < bj ? aj : bj) + 1; ++c)
      v += row_weight(g, w, scale2, F_YY, ai, c, ak) * m(aj - c) * m(bj - c) / 6.0;
  if (ai == bi && aj == bj)
    for (int c = (ak > bk ? ak : bk) - 1; c <= (ak < bk ? ak : bk) + 1; ++c)
      v += row_weight(g, w, scale2, F_ZZ, ai, aj, c) * m(ak - c) * m(bk - c) / 6.0;
  // cross masks: tap -1/2 * du * dv at (du, dv) in {-1, 1}^2
  for (int du = -1; du <= 1; du += 2)
    for (int dv = -1; dv <= 1; dv += 2) {
      // a sits at centre + (du, dv); b at centre + (eu, ev)
      if (ak == bk) {
        const int cx = ai - du, cy = aj - dv, eu = bi - cx, ev = bj - cy;
        if ((eu == 1 || eu == -1) && (ev == 1 || ev == -1))
          v += row_weight(g, w, scale2, F_XY, cx, cy, ak) * 0.25 * du * dv * eu * ev;
      }
      if (aj == bj) {
        const int cx = ai - du, cz = ak - dv, eu = bi - cx, ev = bk - cz;
        if ((eu == 1 || eu == -1) && (ev == 1 || ev == -1))
          v += row_weight(g, w, scale2, F_XZ, cx, aj, cz) * 0.25 * du * dv * eu * ev;
      }
      if (ai == bi) {
        const int cy = aj - du, cz = ak - dv, eu = bj - cy, ev = bk - cz;
        if ((eu == 1 || eu == -1) && (ev == 1 || ev == -1))
          v += row_weight(g, w, scale2, F_YZ, ai, cy, cz) * 0.25 * du * dv * eu * ev;
      }
    }
  return v;
}

__device__ inline double gram_entry(const int64_t* __restrict__ gptr,
                                    const int32_t* __restrict__ gcol,
                                    const double* __restrict__ gval, int64_t a, int64_t b) {
  if (!gptr) return 0.0;
  int64_t lo = gptr[a], hi = gptr[a + 1];
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (gcol[mid] < (int32_t)b)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < gptr[a + 1] && gcol[lo] == (int32_t)b) ? gval[lo] : 0.0;
}

// dense N restricted to every block: entry e of block q is (a, c) of the
// block's s x s matrix, nodes bnodes[boff[q] + a], bnodes[boff[q] + c]
__global__ void k_blk_fill(const int64_t* __restrict__ boff, const int32_t* __restrict__ bnodes,
                           const int64_t* __restrict__ moff, int nb, int64_t ntot, GridDev g,
                           RegW<double> w, const double* scale2, const int64_t* gptr,
                           const int32_t* gcol, const double* gval, double* __restrict__ dense) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ntot) return;
  int lo = 0, hi = nb - 1;
  while (lo < hi) {  // last block with moff <= e
    const int mid = (lo + hi + 1) >> 1;
    if (moff[mid] <= e)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int64_t s = boff[lo + 1] - boff[lo];
  const int64_t r = e - moff[lo];
  const int64_t ia = bnodes[boff[lo] + r / s], ib = bnodes[boff[lo] + r % s];
  dense[e] = reg_entry(g, w, scale2, ia, ib) + gram_entry(gptr, gcol, gval, ia, ib);
}

// in-place Gauss-Jordan inverse of every (symmetric positive definite)
// block, one workgroup of 64 per block (the matrix in LDS up to kBlkLds
// rows, in global memory beyond); *bad set when a pivot is not positive
constexpr int kBlkLds = 40;
__global__ __launch_bounds__(64) void k_blk_inv(const int64_t* __restrict__ boff,
                                                const int64_t* __restrict__ moff, int nb,
                                                double* __restrict__ dense,
                                                int* __restrict__ bad) {
  __shared__ double lds[kBlkLds * kBlkLds];
  __shared__ double fcol[kBlkLds];
  const int q = blockIdx.x;
  if (q >= nb) return;
  const int s = (int)(boff[q + 1] - boff[q]);
  double* G = dense + moff[q];
  const bool in_lds = s <= kBlkLds;
  double* A = in_lds ? lds : G;
  const int t = threadIdx.x;
  if (in_lds)
    for (int e = t; e < s * s; e += 64) lds[e] = G[e];
  __syncthreads();
  for (int k = 0; k < s; ++k) {
    const double p = A[k * s + k];
    if (!(p > 0.0)) {
      if (t == 0) *bad = 1;
      return;  // uniform: every thread read the same pivot
    }
    const double ip = 1.0 / p;
    __syncthreads();
    for (int j = t; j < s; j += 64) A[k * s + j] = (j == k) ? ip : A[k * s + j] * ip;
    for (int i = t; i < s; i += 64)
      if (i != k && in_lds) fcol[i] = A[i * s + k];
    __syncthreads();
    if (in_lds) {
      for (int e = t; e < s * s; e += 64) {
        const int i = e / s, j = e - (e / s) * s;
        if (i == k) continue;
        const double f = fcol[i];
        A[e] = (j == k) ? -f * ip : A[e] - f * A[k * s + j];
      }
    } else {
      // global path: one row per thread, the factor read before the row
      for (int i = t; i < s; i += 64) {
        if (i == k) continue;
        const double f = A[i * s + k];
        for (int j = 0; j < s; ++j)
          A[i * s + j] = (j == k) ? -f * ip : A[i * s + j] - f * A[k * s + j];
      }
    }
    __syncthreads();
  }
  if (in_lds)
    for (int e = t; e < s * s; e += 64) G[e] = lds[e];
}

// (row << 32 | col, value) pairs of every block inverse
__global__ void k_blk_pairs(const int64_t* __restrict__ boff, const int32_t* __restrict__ bnodes,
                            const int64_t* __restrict__ moff, int nb, int64_t ntot,
                            const double* __restrict__ dense, uint64_t* __restrict__ key,
                            double* __restrict__ val) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ntot) return;
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (moff[mid] <= e)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int64_t s = boff[lo + 1] - boff[lo];
  const int64_t r = e - moff[lo];
  const int32_t ia = bnodes[boff[lo] + r / s], ib = bnodes[boff[lo] + r % s];
  key[e] = ((uint64_t)(uint32_t)ia << 32) | (uint32_t)ib;
  val[e] = dense[e];
}

// flags of the covered nodes (negative tags)
__global__ void k_blk_flags(const float* __restrict__ sm, uint8_t* __restrict__ flag, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = sm[i] < 0.f ? 1 : 0;
}

// smoother tags: -1 on nodes with a Schwarz row, the Jacobi scale elsewhere
__global__ void k_blk_tags(const int64_t* __restrict__ sptr, const float* __restrict__ minv,
                           float* __restrict__ sm, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sm[i] = sptr[i + 1] > sptr[i] ? -1.f : minv[i];
}

// ------------------------------------------------- sliced rows (SELL-64) ----
// A sparse row set in slices of 64 rows, one wavefront each: entry k of the
// slice's row t at off[s] + 64 k + t (columns and values), slice width
// width[s] (the longest row of the slice, rounded up to a multiple of 4).
// The 64 threads of a wavefront read one entry index of their 64 rows as one
// contiguous 256-byte line, where CSR with a thread per row touches 64
// lines.  Padding entries point at the row's own node with value 0.  Row r
// is node node[r] (node == nullptr: node r).
struct Sell {
  const int64_t* off;
  const int32_t* width;
  const int32_t* col;
  const float* val;
  const int32_t* node;
  int64_t nrows;
};

// row r of a CSR (ptr, with tags sm when given: a row whose tag is >= 0 is
// the single entry (node, tag), the smoother's point-Jacobi nodes)
__device__ inline int64_t sell_src_len(const int64_t* __restrict__ ptr,
                                       const float* __restrict__ sm, int64_t i) {
  if (sm && sm[i] >= 0.f) return 1;
  return ptr[i + 1] - ptr[i];
}

__global__ void k_sell_width(const int64_t* __restrict__ ptr, const float* __restrict__ sm,
                             const int32_t* __restrict__ node, int64_t nrows,
                             int32_t* __restrict__ width) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s * 64 >= nrows) return;
  int64_t w = 0;
  const int64_t r1 = s * 64 + 64 < nrows ? s * 64 + 64 : nrows;
  for (int64_t r = s * 64; r < r1; ++r) {
    const int64_t l = sell_src_len(ptr, sm, node ? node[r] : r);
    w = l > w ? l : w;
  }
  width[s] = (int32_t)((w + 3) & ~(int64_t)3);
}

__global__ void k_sell_fill(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                            const float* __restrict__ val, const float* __restrict__ sm,
                            const int32_t* __restrict__ node, int64_t nrows,
                            const int64_t* __restrict__ off, const int32_t* __restrict__ width,
                            int32_t* __restrict__ ocol, float* __restrict__ oval) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t s = r >> 6;
  if (s * 64 >= nrows) return;
  const int t = (int)(r & 63);
  const int w = width[s];
  const int64_t o = off[s] + t;
  if (r >= nrows) {
    for (int k = 0; k < w; ++k) {
      ocol[o + 64 * (int64_t)k] = 0;
      oval[o + 64 * (int64_t)k] = 0.f;
    }
    return;
  }
  const int64_t i = node ? node[r] : r;
  int k = 0;
  if (sm && sm[i] >= 0.f) {
    if (w > 0) {
      ocol[o] = (int32_t)i;
      oval[o] = sm[i];
      k = 1;
    }
  } else {
    const int64_t e0 = ptr[i], len = ptr[i + 1] - e0;
    for (; k < len; ++k) {
      ocol[o + 64 * (int64_t)k] = col[e0 + k];
      oval[o + 64 * (int64_t)k] = val[e0 + k];
    }
  }
  for (; k < w; ++k) {
    ocol[o + 64 * (int64_t)k] = (int32_t)i;
    oval[o + 64 * (int64_t)k] = 0.f;
  }
}

// flags of the nodes with a nonempty CSR row
__global__ void k_rows_nonempty(const int64_t* __restrict__ ptr, uint8_t* __restrict__ flag,
                                int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = ptr[i + 1] > ptr[i] ? 1 : 0;
}

// pos[cov[t]] = t (pos initialised to -1)
__global__ void k_cov_pos(const int32_t* __restrict__ cov, int64_t ncov, int32_t* __restrict__ pos) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ncov) pos[cov[t]] = (int32_t)t;
}
// sliced columns (node ids) to covered-list positions; *bad = 1 for a node
// outside the list
__global__ void k_sell_to_pos(int32_t* __restrict__ col, int64_t nent,
                              const int32_t* __restrict__ pos, int* __restrict__ bad) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nent) return;
  const int32_t p = pos[col[e]];
  if (p < 0) {
    *bad = 1;
    col[e] = 0;
  } else {
    col[e] = p;
  }
}

// data rows of level 0 in fixed width wc (repeated columns merged, padded
// with zero coefficients on the first column), for the level's Gram matrix
__global__ void k_rows_fixed(const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ cols,
                             const double* __restrict__ coef, int64_t nrows, int wc,
                             int32_t* __restrict__ ccols, double* __restrict__ ccoef) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int nu = 0;
  const int64_t o = r * wc;
  for (int64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
    int u = 0;
    while (u < nu && ccols[o + u] != cols[e]) ++u;
    if (u == nu) {
      if (nu == wc) continue;  // cannot happen: wc >= row width
      ccols[o + nu] = cols[e];
      ccoef[o + nu++] = 0.0;
    }
    ccoef[o + u] += coef[e];
  }
  for (int u = nu; u < wc; ++u) {
    ccols[o + u] = nu > 0 ? ccols[o] : 0;
    ccoef[o + u] = 0.0;
  }
}

}  // namespace lsamd
