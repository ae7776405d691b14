// Multigrid-preconditioned flexible CG for several right-hand sides of one
// FD system at once (realisations that share the constraint locations and
// differ in observed values or normal directions: same matrix, other rhs).
// The R vectors of a level lie one after the other, v[r * n + i] (every
// load of a lane coalesced across the wavefront), and every launch of the
// V-cycle and of the outer iteration serves all R lanes: the coefficients,
// indices and Schwarz rows are read once for R lanes and the launch count
// per solve drops R-fold.  The hierarchy (coarse grids, Gram
// matrices, Schwarz blocks, coarsest inverse, Chebyshev polynomials) is the
// single-system one.  The outer flexible CG keeps x in FP64 and p, N p and
// the residual in FP32 (half the bytes of its vector passes), with the
// residual replaced by the FP64 rhs - N x every few iterations.  Lanes
// converge independently: a converged lane gets alpha = 0 and keeps its
// iterate; every kernel returns at once when all lanes are done.
#pragma once

namespace lsamd {

// levels with at least this many nodes apply their coarse operator one
// thread per node for all lanes (fewer, per-lane threads below)
constexpr int64_t kMrNodeThreads = 1 << 17;

// per right-hand side
struct MrCtl {
  double rho, rho_prev, beta, alpha, res2, thr, bnorm;
  int iter, done, converged, maxiter;
};

// ------------------------------------------------------------ packing ----
// out[r n + i] = in[r n + i] for r < nr, 0 for the padding lanes nr <= r < R
__global__ void k_mr_pack(const double* __restrict__ in, int nr, int R, int64_t n,
                          double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * R) return;
  out[e] = e < n * nr ? in[e] : 0.0;
}

// ----------------------------------------------------------- operator ----
// Every kernel of the V-cycle runs one thread per (lane, node), the lane in
// blockIdx.y: the loads of one lane are coalesced and the small levels still
// get R times the threads of a single solve.

// y = R x (class-table regulariser, ridge included)
template <typename T>
__global__ __launch_bounds__(256) void k_mr_stencil(const T* __restrict__ x, T* __restrict__ y,
                                                    GridDev g, const T* __restrict__ tab, int R,
                                                    const int* __restrict__ all_done) {
  if (*all_done) return;
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int64_t r = blockIdx.y;
  const T* xs = x + r * n;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int64_t nxy = (int64_t)nx * ny;
  const int k = (int)(id / nxy);
  const int64_t rr = id - (int64_t)k * nxy;
  const int j = (int)(rr / nx);
  const int i = (int)(rr - (int64_t)j * nx);
  const T* c = tab + ((cls7(k, nz) * 7 + cls7(j, ny)) * 7 + cls7(i, nx)) * kTaps;
  T acc = c[0] * xs[id];
#pragma unroll
  for (int t = 1; t < kTaps; ++t) {
    const int dx = kBTap[t][0], dy = kBTap[t][1], dz = kBTap[t][2];
    const bool ok = (unsigned)(i + dx) < (unsigned)nx && (unsigned)(j + dy) < (unsigned)ny &&
                    (unsigned)(k + dz) < (unsigned)nz;
    const T xv = xs[ok ? id + dx + (int64_t)nx * dy + nxy * dz : id];
    acc += c[t] * (ok ? xv : (T)0);
  }
  y[r * n + id] = acc;
}

// t[r nrows + row] = w_r^2 (a_r . x_r)
template <typename T>
__global__ __launch_bounds__(256) void k_mr_rows_t(const T* __restrict__ x, int64_t n, RowsDev Rw,
                                                   T* __restrict__ t, int R,
                                                   const int* __restrict__ all_done) {
  if (*all_done) return;
  const T* coef = (const T*)Rw.coef;
  const T* w2 = (const T*)Rw.w2;
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= Rw.nrows) return;
  const int64_t r = blockIdx.y;
  const T* xs = x + r * n;
  T acc = 0;
  for (int64_t q = Rw.row_ptr[row]; q < Rw.row_ptr[row + 1]; ++q) acc += coef[q] * xs[Rw.cols[q]];
  t[r * Rw.nrows + row] = w2[row] * acc;
}

// the same with the row part of x.Nx: t = w^2 (a.x) per lane and per-block
// partials (FP64) of sum_rows t (a.x), written upart[lane * gridDim.x + blk]
template <typename T>
__global__ __launch_bounds__(256) void k_mr_rows_tu(const T* __restrict__ x, int64_t n,
                                                    RowsDev Rw, T* __restrict__ t,
                                                    double* __restrict__ upart,
                                                    const int* __restrict__ all_done) {
  __shared__ double sh[4];
  if (*all_done) return;
  const T* coef = (const T*)Rw.coef;
  const T* w2 = (const T*)Rw.w2;
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t r = blockIdx.y;
  double u = 0.0;
  if (row < Rw.nrows) {
    const T* xs = x + r * n;
    T acc = 0;
    for (int64_t q = Rw.row_ptr[row]; q < Rw.row_ptr[row + 1]; ++q)
      acc += coef[q] * xs[Rw.cols[q]];
    const T tv = w2[row] * acc;
    t[r * Rw.nrows + row] = tv;
    u = (double)tv * (double)acc;
  }
  const double bs = block_sum(u, sh);
  if (threadIdx.x == 0) upart[r * gridDim.x + blockIdx.x] = bs;
}

// y[node] += sum over the node's entries of coef * t[row]
template <typename T>
__global__ __launch_bounds__(256) void k_mr_rows_scatter(RowsDev Rw, const T* __restrict__ t,
                                                         T* __restrict__ y, int64_t n, int R,
                                                         const int* __restrict__ all_done) {
  if (*all_done) return;
  const T* coef = (const T*)Rw.coef;
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= Rw.ndn) return;
  const int64_t r = blockIdx.y;
  const T* ts = t + r * Rw.nrows;
  T acc = 0;
  for (int64_t q = Rw.seg_ptr[d]; q < Rw.seg_ptr[d + 1]; ++q)
    acc += coef[Rw.seg_entry[q]] * ts[Rw.seg_row[q]];
  y[r * n + Rw.dnode[d]] += acc;
}

// coarse level: y = (class-table regulariser + Gram) x
__global__ __launch_bounds__(256) void k_mr_apply_coarse(MgDev L, const float* __restrict__ x,
                                                         float* __restrict__ y, int R,
                                                         const int* __restrict__ all_done) {
  if (*all_done) return;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= L.n) return;
  const int64_t r = blockIdx.y;
  y[r * L.n + id] = mg_apply_node(L, x + r * L.n, id);
}

// coarse level, one thread per node and all lanes (large coarse levels: the
// class row and the Gram row are read once for the R lanes)
template <int R>
__global__ __launch_bounds__(256) void k_mr_apply_coarse_all(MgDev L,
                                                             const float* __restrict__ x,
                                                             float* __restrict__ y,
                                                             const int* __restrict__ all_done) {
  if (*all_done) return;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= L.n) return;
  const int64_t n = L.n;
  int i, j, k;
  mg_ijk(id, L.g.nx, L.g.ny, i, j, k);
  const int nx = L.g.nx, ny = L.g.ny, nz = L.g.nz;
  const int64_t nxy = (int64_t)nx * ny;
  const float* c = L.tab + ((cls7(k, nz) * 7 + cls7(j, ny)) * 7 + cls7(i, nx)) * kTaps;
  float acc[R];
  const float c0 = c[0];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = c0 * x[r * n + id];
#pragma unroll
  for (int t = 1; t < kTaps; ++t) {
    const int dx = kBTap[t][0], dy = kBTap[t][1], dz = kBTap[t][2];
    const bool ok = (unsigned)(i + dx) < (unsigned)nx && (unsigned)(j + dy) < (unsigned)ny &&
                    (unsigned)(k + dz) < (unsigned)nz;
    if (!ok) continue;
    const int64_t o = id + dx + (int64_t)nx * dy + nxy * dz;
    const float ct = c[t];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] += ct * x[r * n + o];
  }
  if (L.gptr)
    for (int64_t e = L.gptr[id]; e < L.gptr[id + 1]; ++e) {
      const float a = L.gval[e];
      const int64_t o = L.gcol[e];
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] += a * x[r * n + o];
    }
#pragma unroll
  for (int r = 0; r < R; ++r) y[r * n + id] = acc[r];
}

// y += G x for the lanes of a coarse level (G its Gram rows), LG lanes per
// thread (lane group in blockIdx.y): the row is read once for the LG lanes;
// small levels take LG = 1 so that every lane has threads of its own.  The
// regulariser part of y is already there.
template <int LG>
__global__ __launch_bounds__(256) void k_mr_gram_add(const int64_t* __restrict__ gptr,
                                                     const int32_t* __restrict__ gcol,
                                                     const float* __restrict__ gval, int64_t n,
                                                     const float* __restrict__ x,
                                                     float* __restrict__ y,
                                                     const int* __restrict__ all_done) {
  if (*all_done) return;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int64_t e0 = gptr[id], e1 = gptr[id + 1];
  if (e0 == e1) return;
  const int64_t r0 = (int64_t)blockIdx.y * LG;
  float acc[LG];
#pragma unroll
  for (int l = 0; l < LG; ++l) acc[l] = 0.f;
  for (int64_t e = e0; e < e1; ++e) {
    const float a = gval[e];
    const int64_t o = gcol[e];
#pragma unroll
    for (int l = 0; l < LG; ++l) acc[l] += a * x[(r0 + l) * n + o];
  }
#pragma unroll
  for (int l = 0; l < LG; ++l) y[(r0 + l) * n + id] += acc[l];
}

// y += G x with G's rows sliced (SELL-64: the column and value loads of a
// wavefront are contiguous), LG lanes per thread (lane group in blockIdx.y).
// A slice without entries leaves its nodes' y untouched.
template <int LG>
__global__ __launch_bounds__(256) void k_mr_gram_sell(Sell G, int64_t n,
                                                      const float* __restrict__ x,
                                                      float* __restrict__ y,
                                                      const int* __restrict__ all_done,
                                                      int xcd) {
  if (*all_done) return;
  unsigned bx, by;
  xcd_block(xcd, bx, by);
  const int64_t r = (int64_t)bx * blockDim.x + threadIdx.x;
  if (r >= G.nrows) return;
  const int64_t s = r >> 6;
  const int w = G.width[s];
  if (w == 0) return;
  const int64_t o = G.off[s] + (r & 63);
  const int64_t r0 = (int64_t)by * LG;
  const float* xl = x + r0 * n;
  float acc[LG];
#pragma unroll
  for (int l = 0; l < LG; ++l) acc[l] = 0.f;
  for (int k = 0; k < w; k += 4) {
    const int64_t e = o + 64 * (int64_t)k;
    const int32_t j0 = G.col[e], j1 = G.col[e + 64], j2 = G.col[e + 128], j3 = G.col[e + 192];
    const float a0 = G.val[e], a1 = G.val[e + 64], a2 = G.val[e + 128], a3 = G.val[e + 192];
#pragma unroll
    for (int l = 0; l < LG; ++l) {
      const float* xx = xl + l * n;
      acc[l] += a0 * xx[j0] + a1 * xx[j1] + a2 * xx[j2] + a3 * xx[j3];
    }
  }
  const int64_t i = G.node ? G.node[r] : r;
#pragma unroll
  for (int l = 0; l < LG; ++l) y[(r0 + l) * n + i] += acc[l];
}

// ----------------------------------------------------------- smoother ----
// Chebyshev step with the Schwarz preconditioner M: mode 0 first step from
// zero (d = c0 M b, x = d), 1 first step (d = c0 M (b - q), x += d), 2 later
// step (d = c1 d + c2 M (b - q), x += d); mode | kMrLast: the last step of
// the smoothing, whose d is dead and not stored.  skip_cov (large levels):
// the covered nodes are left to k_mr_cheb_cov.
constexpr int kMrLast = 4;

__global__ __launch_bounds__(256) void k_mr_cheb(const float* __restrict__ b,
                                                 const float* __restrict__ q, MgSm S,
                                                 float* __restrict__ x, float* __restrict__ d,
                                                 int64_t n, int skip_cov, int mode, float c0,
                                                 float c1, float c2,
                                                 const int* __restrict__ all_done) {
  if (*all_done) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (skip_cov && S.sm[i] < 0.f) return;
  const int64_t r = blockIdx.y, e = r * n + i;
  const int m = mode & 3;
  const float z = mg_precond(i, S, b + r * n, q + r * n, m == 0);
  const float dv = m == 2 ? c1 * d[e] + c2 * z : c0 * z;
  if (!(mode & kMrLast)) d[e] = dv;
  x[e] = m == 0 ? dv : x[e] + dv;
}

// the same step for the covered nodes of a large level (compact list), one
// thread per (lane, listed node)
__global__ __launch_bounds__(256) void k_mr_cheb_cov(const int32_t* __restrict__ cov,
                                                     int64_t ncov, const float* __restrict__ b,
                                                     const float* __restrict__ q, MgSm S,
                                                     float* __restrict__ x, float* __restrict__ d,
                                                     int64_t n, int mode, float c0, float c1,
                                                     float c2, const int* __restrict__ all_done) {
  if (*all_done) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncov) return;
  const int64_t i = cov[t];
  const int64_t r = blockIdx.y, e = r * n + i;
  const int m = mode & 3;
  const float z = mg_precond(i, S, b + r * n, q + r * n, m == 0);
  const float dv = m == 2 ? c1 * d[e] + c2 * z : c0 * z;
  if (!(mode & kMrLast)) d[e] = dv;
  x[e] = m == 0 ? dv : x[e] + dv;
}

// the same step for LG lanes per thread: the covered node's row of summed
// block inverses is read once for the LG lanes (lane group in blockIdx.y);
// entries summed in the association of mg_precond
template <int LG>
__global__ __launch_bounds__(256) void k_mr_cheb_cov_lg(const int32_t* __restrict__ cov,
                                                        int64_t ncov, const float* __restrict__ b,
                                                        const float* __restrict__ q, MgSm S,
                                                        float* __restrict__ x,
                                                        float* __restrict__ d, int64_t n, int mode,
                                                        float c0, float c1, float c2,
                                                        const int* __restrict__ all_done, int xcd) {
  if (*all_done) return;
  unsigned bx, by;
  xcd_block(xcd, bx, by);
  const int64_t t = (int64_t)bx * blockDim.x + threadIdx.x;
  if (t >= ncov) return;
  const int64_t i = cov[t];
  const int64_t r0 = (int64_t)by * LG;
  const float* bl = b + r0 * n;
  const float* ql = q + r0 * n;
  const int m = mode & 3;
  const int fz = m == 0;
  float acc[LG];
#pragma unroll
  for (int l = 0; l < LG; ++l) acc[l] = 0.f;
  int64_t e = S.sptr[i];
  const int64_t e1 = S.sptr[i + 1];
  for (; e + 4 <= e1; e += 4) {
    const int64_t j0 = S.scol[e], j1 = S.scol[e + 1], j2 = S.scol[e + 2], j3 = S.scol[e + 3];
    const float a0 = S.sval[e], a1 = S.sval[e + 1], a2 = S.sval[e + 2], a3 = S.sval[e + 3];
#pragma unroll
    for (int l = 0; l < LG; ++l) {
      const float* bb = bl + l * n;
      const float* qq = ql + l * n;
      acc[l] += a0 * mg_r(bb, qq, j0, fz) + a1 * mg_r(bb, qq, j1, fz) +
                a2 * mg_r(bb, qq, j2, fz) + a3 * mg_r(bb, qq, j3, fz);
    }
  }
  for (; e < e1; ++e) {
    const int64_t j = S.scol[e];
    const float a = S.sval[e];
#pragma unroll
    for (int l = 0; l < LG; ++l) acc[l] += a * mg_r(bl + l * n, ql + l * n, j, fz);
  }
#pragma unroll
  for (int l = 0; l < LG; ++l) {
    const int64_t el = (r0 + l) * n + i;
    const float dv = m == 2 ? c1 * d[el] + c2 * acc[l] : c0 * acc[l];
    if (!(mode & kMrLast)) d[el] = dv;
    x[el] = m == 0 ? dv : x[el] + dv;
  }
}

// the same step over the rows of the sliced smoother matrix P (rows of the
// covered nodes, or of every node with the Jacobi scale as a one-entry row),
// LG lanes per thread (lane group in blockIdx.y)
template <int LG>
__global__ __launch_bounds__(256) void k_mr_cheb_sell(Sell P, const float* __restrict__ b,
                                                      const float* __restrict__ q,
                                                      float* __restrict__ x,
                                                      float* __restrict__ d, int64_t n, int mode,
                                                      float c0, float c1, float c2,
                                                      const int* __restrict__ all_done, int xcd) {
  if (*all_done) return;
  unsigned bx, by;
  xcd_block(xcd, bx, by);
  const int64_t r = (int64_t)bx * blockDim.x + threadIdx.x;
  if (r >= P.nrows) return;
  const int64_t s = r >> 6;
  const int w = P.width[s];
  const int64_t o = P.off[s] + (r & 63);
  const int64_t r0 = (int64_t)by * LG;
  const float* bl = b + r0 * n;
  const float* ql = q + r0 * n;
  const int m = mode & 3;
  const int fz = m == 0;
  float acc[LG];
#pragma unroll
  for (int l = 0; l < LG; ++l) acc[l] = 0.f;
  for (int k = 0; k < w; k += 4) {
    const int64_t e = o + 64 * (int64_t)k;
    const int32_t j0 = P.col[e], j1 = P.col[e + 64], j2 = P.col[e + 128], j3 = P.col[e + 192];
    const float a0 = P.val[e], a1 = P.val[e + 64], a2 = P.val[e + 128], a3 = P.val[e + 192];
#pragma unroll
    for (int l = 0; l < LG; ++l) {
      const float* bb = bl + l * n;
      const float* qq = ql + l * n;
      acc[l] += a0 * mg_r(bb, qq, j0, fz) + a1 * mg_r(bb, qq, j1, fz) +
                a2 * mg_r(bb, qq, j2, fz) + a3 * mg_r(bb, qq, j3, fz);
    }
  }
  const int64_t i = P.node ? P.node[r] : r;
#pragma unroll
  for (int l = 0; l < LG; ++l) {
    const int64_t el = (r0 + l) * n + i;
    const float dv = m == 2 ? c1 * d[el] + c2 * acc[l] : c0 * acc[l];
    if (!(mode & kMrLast)) d[el] = dv;
    x[el] = m == 0 ? dv : x[el] + dv;
  }
}

// Covered nodes of a large level, interleaved: ri[t R + l] = r of lane l at
// the t-th covered node (r = b, or b - q), so that a covered row's gather
// reads one contiguous run of R values per column instead of R lines
// n apart (lane-major gathers over R lanes overflow the L2 and re-fetch the
// same lines many times).  64 covered nodes per block, transposed in LDS so
// that both the lane-major reads and the interleaved writes are coalesced.
template <int R>
__global__ __launch_bounds__(256) void k_mr_cov_pack(const int32_t* __restrict__ cov,
                                                     int64_t ncov, const float* __restrict__ b,
                                                     const float* __restrict__ q, int64_t n,
                                                     int from_zero, float* __restrict__ ri,
                                                     const int* __restrict__ all_done) {
  __shared__ float tile[64][R + 1];
  if (*all_done) return;
  const int64_t t0 = (int64_t)blockIdx.x * 64;
  const int tl = threadIdx.x & 63, lg = threadIdx.x >> 6;
  const int64_t t = t0 + tl;
  if (t < ncov) {
    const int64_t i = cov[t];
    for (int l = lg; l < R; l += 4) {
      const int64_t e = (int64_t)l * n + i;
      tile[tl][l] = from_zero ? b[e] : b[e] - q[e];
    }
  }
  __syncthreads();
  const int64_t cnt = ncov - t0 < 64 ? ncov - t0 : 64;
  for (int k = threadIdx.x; k < cnt * R; k += 256) ri[t0 * R + k] = tile[k / R][k % R];
}

// M r at the covered nodes from the interleaved r: sliced rows over the
// covered list whose columns are covered-list positions; LG lanes per
// thread (lane group in blockIdx.y), read as LG / 4 float4 per column.
// Output lane-major over the list, zc[l ncov + t].
template <int R, int LG>
__global__ __launch_bounds__(256) void k_mr_cov_z(Sell P, const float* __restrict__ ri,
                                                  float* __restrict__ zc,
                                                  const int* __restrict__ all_done) {
  static_assert(LG % 4 == 0 && R % LG == 0, "lane groups of float4");
  if (*all_done) return;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= P.nrows) return;
  const int64_t s = r >> 6;
  const int w = P.width[s];
  const int64_t o = P.off[s] + (r & 63);
  const int l0 = blockIdx.y * LG;
  float acc[LG];
#pragma unroll
  for (int l = 0; l < LG; ++l) acc[l] = 0.f;
  for (int k = 0; k < w; k += 2) {
    const int64_t e = o + 64 * (int64_t)k;
    const int32_t h0 = P.col[e], h1 = P.col[e + 64];
    const float a0 = P.val[e], a1 = P.val[e + 64];
    const float4* v0 = (const float4*)(ri + (int64_t)h0 * R + l0);
    const float4* v1 = (const float4*)(ri + (int64_t)h1 * R + l0);
#pragma unroll
    for (int g = 0; g < LG / 4; ++g) {
      const float4 p = v0[g], u = v1[g];
      acc[4 * g] += a0 * p.x + a1 * u.x;
      acc[4 * g + 1] += a0 * p.y + a1 * u.y;
      acc[4 * g + 2] += a0 * p.z + a1 * u.z;
      acc[4 * g + 3] += a0 * p.w + a1 * u.w;
    }
  }
#pragma unroll
  for (int l = 0; l < LG; ++l) zc[(int64_t)(l0 + l) * P.nrows + r] = acc[l];
}

// Chebyshev step over every node of a level with a covered list, LG lanes
// per thread (lane group in blockIdx.y): M r is the Jacobi scale times r on
// uncovered nodes and zc (k_mr_cov_z) on covered ones, pos[i] the node's
// list position.  Writing whole lines (no skipped nodes) avoids the partial
// line writes that the covered runs left in a split update.
template <int LG>
__global__ __launch_bounds__(256) void k_mr_cheb_all(const float* __restrict__ b,
                                                     const float* __restrict__ q,
                                                     const float* __restrict__ sm,
                                                     const int32_t* __restrict__ pos,
                                                     const float* __restrict__ zc, int64_t ncov,
                                                     float* __restrict__ x, float* __restrict__ d,
                                                     int64_t n, int mode, float c0, float c1,
                                                     float c2, const int* __restrict__ all_done) {
  if (*all_done) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t r0 = (int64_t)blockIdx.y * LG;
  const int m = mode & 3;
  const float t = sm[i];
  float z[LG];
  if (t >= 0.f) {
#pragma unroll
    for (int l = 0; l < LG; ++l) {
      const int64_t e = (r0 + l) * n + i;
      z[l] = t * (m == 0 ? b[e] : b[e] - q[e]);
    }
  } else {
    const int64_t p = pos[i];
#pragma unroll
    for (int l = 0; l < LG; ++l) z[l] = zc[(r0 + l) * ncov + p];
  }
#pragma unroll
  for (int l = 0; l < LG; ++l) {
    const int64_t e = (r0 + l) * n + i;
    const float dv = m == 2 ? c1 * d[e] + c2 * z[l] : c0 * z[l];
    if (!(mode & kMrLast)) d[e] = dv;
    x[e] = m == 0 ? dv : x[e] + dv;
  }
}

// ------------------------------------------------- transfer, coarsest ----
// Restriction and prolongation with LG lanes per thread (lane group in
// blockIdx.y): the node's index arithmetic is done once for the LG lanes;
// sums in the order of mg_restrict_node / mg_prolong_node
template <int LG>
__global__ __launch_bounds__(256) void k_mr_restrict_lg(const float* __restrict__ bf,
                                                        const float* __restrict__ qf, GridDev gf,
                                                        float* __restrict__ bc, GridDev gc,
                                                        const int* __restrict__ all_done,
                                                        int xcd) {
  if (*all_done) return;
  unsigned bx, by;
  xcd_block(xcd, bx, by);
  const int64_t nc = (int64_t)gc.nx * gc.ny * gc.nz;
  const int64_t nf = (int64_t)gf.nx * gf.ny * gf.nz;
  const int64_t id = (int64_t)bx * blockDim.x + threadIdx.x;
  if (id >= nc) return;
  const int64_t r0 = (int64_t)by * LG;
  int I, J, K;
  mg_ijk(id, gc.nx, gc.ny, I, J, K);
  const int64_t fnxy = (int64_t)gf.nx * gf.ny;
  float acc[LG];
#pragma unroll
  for (int l = 0; l < LG; ++l) acc[l] = 0.f;
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
#pragma unroll
        for (int l = 0; l < LG; ++l) {
          const int64_t e = (r0 + l) * nf + f;
          acc[l] += w * (bf[e] - qf[e]);
        }
      }
    }
  }
#pragma unroll
  for (int l = 0; l < LG; ++l) bc[(r0 + l) * nc + id] = acc[l];
}

// The same restriction streamed along z: a thread per coarse column (I, J)
// of a band of 256 consecutive positions of the coarse plane walks KC coarse
// planes, summing each fine plane's 3 x 3 (I, J) neighbourhood once (the
// plane 2K + 1 serves K and K + 1).  Neighbouring columns share their fine
// rows inside the band, so the fine level is read about once, where the
// node-per-thread form re-reads the y and z neighbours from beyond the L2.
// blockIdx.x = band + nband * (z chunk); lane group in blockIdx.y.
template <int LG, int KC>
__global__ __launch_bounds__(256) void k_mr_restrict_z(const float* __restrict__ bf,
                                                       const float* __restrict__ qf, GridDev gf,
                                                       float* __restrict__ bc, GridDev gc,
                                                       int nband,
                                                       const int* __restrict__ all_done) {
  if (*all_done) return;
  unsigned bx, by;
  xcd_block(1, bx, by);
  const int band = (int)(bx % (unsigned)nband), kch = (int)(bx / (unsigned)nband);
  const int64_t pos = (int64_t)band * 256 + threadIdx.x;
  const int64_t ncp = (int64_t)gc.nx * gc.ny;
  if (pos >= ncp) return;
  const int I = (int)(pos % gc.nx), J = (int)(pos / gc.nx);
  const int K0 = kch * KC, K1 = K0 + KC < gc.nz ? K0 + KC : gc.nz;
  const int64_t nc = ncp * gc.nz;
  const int64_t fnxy = (int64_t)gf.nx * gf.ny;
  const int64_t nf = fnxy * gf.nz;
  const int64_t r0 = (int64_t)by * LG;
  // S(k) = sum over the in-grid (i, j) of {2I-1, 2I, 2I+1} x {2J-1, 2J, 2J+1}
  // of w_i w_j (bf - qf)(i, j, k); 0 outside the fine grid in z
  auto plane = [&](int k, float* s) {
#pragma unroll
    for (int l = 0; l < LG; ++l) s[l] = 0.f;
    if (k < 0 || k >= gf.nz) return;
    for (int dj = -1; dj <= 1; ++dj) {
      const int j = 2 * J + dj;
      if (j < 0 || j >= gf.ny) continue;
      const float wj = dj == 0 ? 1.f : 0.5f;
      for (int di = -1; di <= 1; ++di) {
        const int i = 2 * I + di;
        if (i < 0 || i >= gf.nx) continue;
        const float w = wj * (di == 0 ? 1.f : 0.5f);
        const int64_t f = (int64_t)i + (int64_t)j * gf.nx + (int64_t)k * fnxy;
#pragma unroll
        for (int l = 0; l < LG; ++l) {
          const int64_t e = (r0 + l) * nf + f;
          s[l] += w * (bf[e] - qf[e]);
        }
      }
    }
  };
  float sp[LG], s0[LG], s1[LG];
  plane(2 * K0 - 1, sp);
  for (int K = K0; K < K1; ++K) {
    plane(2 * K, s0);
    plane(2 * K + 1, s1);
    const int64_t c = pos + (int64_t)K * ncp;
#pragma unroll
    for (int l = 0; l < LG; ++l) {
      bc[(r0 + l) * nc + c] = 0.5f * sp[l] + s0[l] + 0.5f * s1[l];
      sp[l] = s1[l];
    }
  }
}

// xf += P xc streamed along z: a thread per fine column (i, j) of a band of
// 256 consecutive positions of the fine plane walks KC fine planes; the
// (i, j) interpolation of a coarse plane is formed once and serves the fine
// planes 2K - 1, 2K and 2K + 1.  blockIdx.x = band + nband * (z chunk); lane
// group in blockIdx.y.
template <int LG, int KC>
__global__ __launch_bounds__(256) void k_mr_prolong_z(const float* __restrict__ xc, GridDev gc,
                                                      float* __restrict__ xf, GridDev gf,
                                                      int nband,
                                                      const int* __restrict__ all_done) {
  static_assert(KC % 2 == 0, "chunks start on even fine planes");
  if (*all_done) return;
  unsigned bx, by;
  xcd_block(1, bx, by);
  const int band = (int)(bx % (unsigned)nband), kch = (int)(bx / (unsigned)nband);
  const int64_t pos = (int64_t)band * 256 + threadIdx.x;
  const int64_t fnxy = (int64_t)gf.nx * gf.ny;
  if (pos >= fnxy) return;
  const int i = (int)(pos % gf.nx), j = (int)(pos / gf.nx);
  const int k0 = kch * KC, k1 = k0 + KC < gf.nz ? k0 + KC : gf.nz;
  const int64_t cnxy = (int64_t)gc.nx * gc.ny;
  const int64_t nc = cnxy * gc.nz, nf = fnxy * gf.nz;
  const int64_t r0 = (int64_t)by * LG;
  const int i0 = i >> 1, j0 = j >> 1;
  const int ni = (i & 1) ? 2 : 1, nj = (j & 1) ? 2 : 1;
  const float wij = (ni == 2 ? 0.5f : 1.f) * (nj == 2 ? 0.5f : 1.f);
  // (i, j) interpolation of coarse plane K
  auto plane = [&](int K, float* s) {
#pragma unroll
    for (int l = 0; l < LG; ++l) s[l] = 0.f;
    for (int b = 0; b < nj; ++b)
      for (int a = 0; a < ni; ++a) {
        const int64_t o = (int64_t)(i0 + a) + (int64_t)(j0 + b) * gc.nx + (int64_t)K * cnxy;
#pragma unroll
        for (int l = 0; l < LG; ++l) s[l] += xc[(r0 + l) * nc + o];
      }
#pragma unroll
    for (int l = 0; l < LG; ++l) s[l] *= wij;
  };
  float pc[LG], pn[LG];
  plane(k0 >> 1, pc);
  for (int k = k0; k < k1; k += 2) {
    const int64_t f = pos + (int64_t)k * fnxy;
#pragma unroll
    for (int l = 0; l < LG; ++l) xf[(r0 + l) * nf + f] += pc[l];
    if (k + 1 >= k1) break;
    plane((k >> 1) + 1, pn);
#pragma unroll
    for (int l = 0; l < LG; ++l) {
      xf[(r0 + l) * nf + f + fnxy] += 0.5f * (pc[l] + pn[l]);
      pc[l] = pn[l];
    }
  }
}

template <int LG>
__global__ __launch_bounds__(256) void k_mr_prolong_lg(const float* __restrict__ xc, GridDev gc,
                                                       float* __restrict__ xf, GridDev gf,
                                                       const int* __restrict__ all_done,
                                                       int xcd) {
  if (*all_done) return;
  unsigned bx, by;
  xcd_block(xcd, bx, by);
  const int64_t nc = (int64_t)gc.nx * gc.ny * gc.nz;
  const int64_t nf = (int64_t)gf.nx * gf.ny * gf.nz;
  const int64_t id = (int64_t)bx * blockDim.x + threadIdx.x;
  if (id >= nf) return;
  const int64_t r0 = (int64_t)by * LG;
  int i, j, k;
  mg_ijk(id, gf.nx, gf.ny, i, j, k);
  const int i0 = i >> 1, j0 = j >> 1, k0 = k >> 1;
  const int ni = (i & 1) ? 2 : 1, nj = (j & 1) ? 2 : 1, nk = (k & 1) ? 2 : 1;
  const float wi = ni == 2 ? 0.5f : 1.f, wj = nj == 2 ? 0.5f : 1.f, wk = nk == 2 ? 0.5f : 1.f;
  const int64_t cnxy = (int64_t)gc.nx * gc.ny;
  float acc[LG];
#pragma unroll
  for (int l = 0; l < LG; ++l) acc[l] = 0.f;
  for (int c = 0; c < nk; ++c)
    for (int b = 0; b < nj; ++b)
      for (int a = 0; a < ni; ++a) {
        const int64_t o = (int64_t)(i0 + a) + (int64_t)(j0 + b) * gc.nx + (int64_t)(k0 + c) * cnxy;
#pragma unroll
        for (int l = 0; l < LG; ++l) acc[l] += xc[(r0 + l) * nc + o];
      }
#pragma unroll
  for (int l = 0; l < LG; ++l) xf[(r0 + l) * nf + id] += wi * wj * wk * acc[l];
}

// x = Cinv b on the coarsest level, one thread per (lane, row)
__global__ void k_mr_dense(const float* __restrict__ cinv, const float* __restrict__ b,
                           float* __restrict__ x, int nc, int R,
                           const int* __restrict__ all_done) {
  if (*all_done) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nc * R) return;
  const int r = e / nc, row = e - (e / nc) * nc;
  float acc = 0.f;
  for (int c = 0; c < nc; ++c) acc += cinv[(int64_t)row * nc + c] * b[r * nc + c];
  x[e] = acc;
}

// --------------------------------------------------------- outer loop ----
// Per-block partial sums of NP products per lane (256-thread blocks),
// written part[(k * R + r) * gridDim.x + blk]: every value is reduced per
// wavefront first, then the four wavefronts' sums are added in index order
// after one barrier.
template <int R, int NP>
__device__ inline void mr_block_partials(const double* v, double* part) {
  __shared__ double sh[4][NP * R];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NP * R; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) sh[wid][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < NP * R) {
    const int k = threadIdx.x;
    part[(int64_t)k * gridDim.x + blockIdx.x] = ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
  }
}
// the same for the G lanes lane0 .. lane0 + G - 1 of an R-lane group:
// v[k G + j] is product k of lane lane0 + j, stored where mr_block_partials
// would store it
template <int G, int NP>
__device__ inline void mr_block_partials_g(const double* v, double* part, int R, int lane0) {
  __shared__ double sh[4][NP * G];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NP * G; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) sh[wid][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < NP * G) {
    const int k = threadIdx.x, slot = k / G, j = k - slot * G;
    part[(int64_t)(slot * R + lane0 + j) * gridDim.x + blockIdx.x] =
        ((sh[0][k] + sh[1][k]) + sh[2][k]) + sh[3][k];
  }
}

// lanes per thread of the outer-iteration vector kernels (lane group in
// blockIdx.y): few enough that the per-lane scalars and sums leave room for
// eight wavefronts per SIMD
constexpr int mr_outer_lanes(int R) { return R < 8 ? R : 8; }

// Sum of the nblk partials of product k of lane r, by one wavefront (fixed
// order: four strided lane sums, then the butterfly); result in every lane.
template <int R>
__device__ inline double mr_wave_reduce(const double* part, int nblk, int k, int r) {
  const int lane = threadIdx.x & 63;
  const double* p = part + (int64_t)(k * R + r) * nblk;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  int b = lane;
  for (; b + 192 < nblk; b += 256) {
    v0 += p[b];
    v1 += p[b + 64];
    v2 += p[b + 128];
    v3 += p[b + 192];
  }
  for (; b < nblk; b += 64) v0 += p[b];
  return wave_sum((v0 + v1) + (v2 + v3));
}

// The one-block kernels run min(R, 16) wavefronts; wavefront w
// serves lanes w, w + nw, ...
constexpr int mr_fin_threads(int R) { return 64 * (R < 16 ? R : 16); }

// x = 0 ; r = (float) rhs ; partials of r.(minv r)
template <int R>
__global__ __launch_bounds__(256) void k_mr_init(const double* __restrict__ rhs,
                                                 const double* __restrict__ minv,
                                                 double* __restrict__ x, float* __restrict__ rf,
                                                 int64_t n, double* __restrict__ part) {
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double m = minv[i];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double v = rhs[r * n + i];
      x[r * n + i] = 0.0;
      rf[r * n + i] = (float)v;
      acc[r] += v * (m * v);
    }
  }
  mr_block_partials<R, 1>(acc, part);
}

// Residual replacement: r = (float) (rhs - N x) with N x in FP64 (q, from the
// FP64 apply of x), partials of r.(minv r); every few iterations, so that
// the FP32 recurrence r -= alpha q never drifts from the true residual by
// more than a few FP32 roundings of the residual since the last one
template <int R>
__global__ __launch_bounds__(256) void k_mr_resid(const double* __restrict__ rhs,
                                                  const double* __restrict__ q,
                                                  const double* __restrict__ minv,
                                                  float* __restrict__ rf, int64_t n,
                                                  double* __restrict__ part,
                                                  const int* __restrict__ all_done) {
  if (*all_done) return;
  constexpr int G = mr_outer_lanes(R);
  const int r0 = blockIdx.y * G;
  double acc[G];
#pragma unroll
  for (int j = 0; j < G; ++j) acc[j] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double m = minv[i];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const int64_t e = (r0 + j) * n + i;
      const double v = rhs[e] - q[e];
      rf[e] = (float)v;
      acc[j] += v * (m * v);
    }
  }
  mr_block_partials_g<G, 1>(acc, part, R, r0);
}

// after a residual replacement: the stopping test on the true residual
// (no iteration counted)
template <int R>
__global__ __launch_bounds__(mr_fin_threads(R)) void k_mr_resid_fin(const double* __restrict__ part,
                                                                    int nblk,
                                                                    MrCtl* __restrict__ ctl,
                                                                    int* __restrict__ all_done) {
  if (*all_done) return;
  __shared__ int done[R];
  for (int r = threadIdx.x >> 6; r < R; r += blockDim.x >> 6) {
    const double s = mr_wave_reduce<R>(part, nblk, 0, r);
    if ((threadIdx.x & 63) == 0) {
      MrCtl& c = ctl[r];
      if (!c.done) {
        c.res2 = s;
        c.converged = sqrt(s) < c.thr;
        c.done = c.converged;
      }
      done[r] = c.done;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int all = 1;
    for (int k = 0; k < R; ++k) all = all && done[k];
    *all_done = all;
  }
}


template <int R>
__global__ __launch_bounds__(mr_fin_threads(R)) void k_mr_init_fin(const double* __restrict__ part, int nblk,
                                                        int nr, double rtol, double atol,
                                                        int maxiter, MrCtl* __restrict__ ctl,
                                                        int* __restrict__ all_done) {
  __shared__ int done[R];
  for (int r = threadIdx.x >> 6; r < R; r += blockDim.x >> 6) {
    const double s = mr_wave_reduce<R>(part, nblk, 0, r);
    if ((threadIdx.x & 63) == 0) {
      MrCtl c = {};
      c.bnorm = sqrt(s);
      c.thr = fmax(atol, rtol * c.bnorm);
      c.res2 = s;
      c.maxiter = maxiter;
      c.converged = r < nr ? (sqrt(s) < c.thr || s == 0.0) : 1;
      c.done = c.converged || maxiter <= 0;
      ctl[r] = c;
      done[r] = c.done;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int all = 1;
    for (int k = 0; k < R; ++k) all = all && done[k];
    *all_done = all;
  }
}

// partials of z.r and z.q per lane (z: the FP32 V-cycle output; q still
// holds the previous N p, so z.r_prev = z.r + alpha_prev z.q)
template <int R>
__global__ __launch_bounds__(256) void k_mr_dots(const float* __restrict__ z,
                                                 const float* __restrict__ rr,
                                                 const float* __restrict__ q, int64_t n,
                                                 double* __restrict__ part,
                                                 const int* __restrict__ all_done) {
  if (*all_done) return;
  constexpr int G = mr_outer_lanes(R);
  const int r0 = blockIdx.y * G;
  double acc[2 * G];
#pragma unroll
  for (int r = 0; r < 2 * G; ++r) acc[r] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const int64_t e = (r0 + j) * n + i;
      const double zv = (double)z[e];
      acc[j] += zv * (double)rr[e];
      acc[G + j] += zv * (double)q[e];
    }
  mr_block_partials_g<G, 2>(acc, part, R, r0);
}

// rho = z.r ; beta = (z.r - z.r_prev) / rho_prev = -alpha_prev z.q / rho_prev
// (0 on the first step)
template <int R>
__global__ __launch_bounds__(mr_fin_threads(R)) void k_mr_dots_fin(const double* __restrict__ part,
                                                                   int nblk,
                                                                   MrCtl* __restrict__ ctl,
                                                                   const int* __restrict__ all_done) {
  if (*all_done) return;
  for (int r = threadIdx.x >> 6; r < R; r += blockDim.x >> 6) {
    const double a = mr_wave_reduce<R>(part, nblk, 0, r);
    const double b = mr_wave_reduce<R>(part, nblk, 1, r);
    if ((threadIdx.x & 63) == 0 && !ctl[r].done) {
      ctl[r].beta = ctl[r].iter == 0 ? 0.0 : -ctl[r].alpha * b / ctl[r].rho_prev;
      ctl[r].rho = a;
    }
  }
}

// p = z + beta p
template <int R>
__global__ void k_mr_dir(const float* __restrict__ z, float* __restrict__ p, int64_t n,
                         const MrCtl* __restrict__ ctl, const int* __restrict__ all_done) {
  if (*all_done) return;
  constexpr int G = mr_outer_lanes(R);
  const int r0 = blockIdx.y * G;
  double beta[G];
  bool first[G];
#pragma unroll
  for (int j = 0; j < G; ++j) {
    beta[j] = ctl[r0 + j].beta;
    first[j] = ctl[r0 + j].iter == 0;
  }
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const int64_t e = (r0 + j) * n + i;
      p[e] = first[j] ? z[e] : (float)((double)z[e] + beta[j] * (double)p[e]);
    }
}

// alpha = rho / p.q (0 on converged lanes), p.q summed from the stencil's
// tpl work-item partials and the rows' nbu block partials of each lane
template <int R>
__global__ __launch_bounds__(mr_fin_threads(R)) void k_mr_alpha_st(
    const double* __restrict__ tpart, int tpl, const double* __restrict__ upart, int nbu,
    MrCtl* __restrict__ ctl, const int* __restrict__ all_done) {
  if (*all_done) return;
  const int lane = threadIdx.x & 63;
  for (int r = threadIdx.x >> 6; r < R; r += blockDim.x >> 6) {
    double v = 0.0;
    for (int b = lane; b < tpl; b += 64) v += tpart[(int64_t)r * tpl + b];
    for (int b = lane; b < nbu; b += 64) v += upart[(int64_t)r * nbu + b];
    const double pq = wave_sum(v);
    if (lane == 0) ctl[r].alpha = ctl[r].done ? 0.0 : ctl[r].rho / pq;
  }
}

// x += alpha p (FP64 x) ; r -= alpha q (FP32 r, the V-cycle's input) ;
// partials of r.(minv r)
template <int R>
__global__ __launch_bounds__(256) void k_mr_upd(double* __restrict__ x,
                                                const float* __restrict__ p,
                                                const float* __restrict__ q,
                                                const double* __restrict__ minv,
                                                float* __restrict__ rf, int64_t n,
                                                double* __restrict__ part,
                                                const MrCtl* __restrict__ ctl,
                                                const int* __restrict__ all_done) {
  if (*all_done) return;
  constexpr int G = mr_outer_lanes(R);
  const int r0 = blockIdx.y * G;
  double alpha[G], acc[G];
#pragma unroll
  for (int j = 0; j < G; ++j) {
    alpha[j] = ctl[r0 + j].alpha;
    acc[j] = 0.0;
  }
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double m = minv[i];
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const int64_t e = (r0 + j) * n + i;
      x[e] += alpha[j] * (double)p[e];
      const double rn = (double)rf[e] - alpha[j] * (double)q[e];
      rf[e] = (float)rn;
      acc[j] += rn * (m * rn);
    }
  }
  mr_block_partials_g<G, 1>(acc, part, R, r0);
}

// rho_prev = rho ; iter++ ; stopping test on |S r| per lane
template <int R>
__global__ __launch_bounds__(mr_fin_threads(R)) void k_mr_upd_fin(const double* __restrict__ part,
                                                                  int nblk, MrCtl* __restrict__ ctl,
                                                                  int* __restrict__ all_done) {
  if (*all_done) return;
  __shared__ int done[R];
  for (int r = threadIdx.x >> 6; r < R; r += blockDim.x >> 6) {
    const double s = mr_wave_reduce<R>(part, nblk, 0, r);
    if ((threadIdx.x & 63) == 0) {
      MrCtl& c = ctl[r];
      if (!c.done) {
        c.rho_prev = c.rho;
        c.iter += 1;
        c.res2 = s;
        c.converged = sqrt(s) < c.thr;
        c.done = c.converged || c.iter >= c.maxiter;
      }
      done[r] = c.done;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int all = 1;
    for (int k = 0; k < R; ++k) all = all && done[k];
    *all_done = all;
  }
}

}  // namespace lsamd
