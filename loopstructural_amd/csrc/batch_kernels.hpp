// Batched ensemble of FD systems of different grid sizes solved together.
//
// The systems' node vectors are concatenated (system s owns the nodes
// [off_s, off_s + n_s)); the data rows of all systems form one CSR block
// with global column ids, rows of system s in [r0_s, r1_s).  Every
// regulariser or vector kernel runs over a table of node blocks, each block
// inside one system, so one launch serves the whole ensemble and systems
// that have converged drop out block by block.  CG scalars are per system
// (CgCtl[s]); the per-system reductions go through a kernel boundary (one
// workgroup per system, fixed order: deterministic).
//
// Regulariser: the FD second-derivative operator does not depend on the grid
// step (LI/_finite_difference_interpolator.py:643-692 uses unscaled masks),
// so all systems with the same weights share one 7^3 x 25 class table.
#pragma once

namespace lsamd {

struct BSys {
  GridDev g;
  int64_t off, n;  // node range in the concatenated vectors
  int64_t r0, r1;  // data rows
  int b0, b1;      // node blocks
  int t0, t1;      // stencil tiles
  const void* xp;  // pointer form (k_b_tiles<..., true>): the system's own x and y
  void* yp;
};

// Stencil work item of the batched plane-streaming kernel: system, xy tile
// and z part.  kind >= 0: bulk z-chunk `kind` of an xy tile whose class
// coefficients vary per lane; kind >= kUniTile: same for an xy tile that is
// interior in x and y (one coefficient set); kZbLow / kZbHigh: output planes
// [0,3) / [nz-3,nz).
constexpr int kUniTile = 1 << 20;
constexpr int kZbLow = -1, kZbHigh = -2;
constexpr int kBZc = 16;  // z-chunk of the batched bulk tiles

// Work items of one system (index sys) for z-chunk zc and tile width txt
// (64 x 8 or 32 x 16 tiles): bulk chunks of every xy tile (interior tiles
// flagged kUniTile), then both z-boundary parts.
inline void stencil_tiles(int sys, const GridDev& g, int zc, std::vector<int4>& out,
                          int txt = TX) {
  const int tw = txt, th = NT / txt;
  const int nbx = (int)ceil_div(g.nx, tw), nby = (int)ceil_div(g.ny, th);
  const int nbz = (int)ceil_div(g.nz - 6, zc);
  for (int ty = 0; ty < nby; ++ty)
    for (int tx = 0; tx < nbx; ++tx) {
      const bool inter = tx * tw >= 3 && tx * tw + tw - 1 <= g.nx - 4 && ty * th >= 3 &&
                         ty * th + th - 1 <= g.ny - 4;
      for (int bz = 0; bz < nbz; ++bz)
        out.push_back(make_int4(sys, tx, ty, inter ? kUniTile + bz : bz));
    }
  for (int ty = 0; ty < nby; ++ty)
    for (int tx = 0; tx < nbx; ++tx) {
      out.push_back(make_int4(sys, tx, ty, kZbLow));
      out.push_back(make_int4(sys, tx, ty, kZbHigh));
    }
}

constexpr int kBNodes = 1024;  // nodes per block (256 threads x 4)

// tap offsets in the order of build_class_table
__constant__ signed char kBTap[kTaps][3] = {
    {0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},   {2, 0, 0},  {-2, 0, 0},
    {0, 2, 0},  {0, -2, 0}, {2, 2, 0},  {-2, 2, 0}, {2, -2, 0},   {-2, -2, 0}, {0, 0, -1},
    {0, 0, 1},  {0, 0, -2}, {2, 0, -2}, {-2, 0, -2}, {0, 2, -2},  {0, -2, -2}, {0, 0, 2},
    {2, 0, 2},  {-2, 0, 2}, {0, 2, 2},  {0, -2, 2}};

// Plane-streaming stencil over a table of work items (one launch for a
// whole system or a whole ensemble); per-block partials of x.y.  The
// single-system plan is the same table with one system.  DMA: bulk planes
// staged by direct global -> LDS copies (st_bulk_dma) instead of through
// registers (st_bulk_body); same outputs bit for bit.  TXT: tile width of
// the table (64 x 8 or 32 x 16 tiles).  PTR: each system's x and y are the
// descriptor's own pointers xp, yp (systems in separate allocations; the
// descriptors' off is 0) instead of offsets into x and y.
template <typename T, int ZC, bool DMA = false, int TXT = 64, bool PTR = false>
__global__ __launch_bounds__(NT) void k_b_tiles(const T* __restrict__ xg, T* __restrict__ yg,
                                                const BSys* __restrict__ sys,
                                                const int4* __restrict__ tiles,
                                                const T* __restrict__ tab, Coef25<T> ku,
                                                double* __restrict__ part,
                                                const CgCtl* __restrict__ ctl) {
  __shared__ double shm[NT / 64];
  const int4 t = tiles[blockIdx.x];
  if (ctl[t.x].done) return;
  const BSys S = sys[t.x];
  // pointer form: descriptors with off = 0
  const T* x = PTR ? (const T*)S.xp : xg;
  T* y = PTR ? (T*)S.yp : yg;
  double dot;
  static_assert(!DMA || TXT == 64, "LDS-DMA forms use 64 x 8 tiles");
  if (t.w >= kUniTile) {
    if constexpr (DMA)
      dot = st_bulk_dma<T, ZC, true>(x + S.off, y + S.off, S.g, tab, ku, make_int2(t.y, t.z),
                                     t.w - kUniTile);
    else
      dot = st_bulk_body<T, ZC, 4, true, TXT>(x + S.off, y + S.off, S.g, tab, ku,
                                              make_int2(t.y, t.z), t.w - kUniTile);
  } else if (t.w >= 0) {
    if constexpr (DMA)
      dot = st_bulk_dma<T, ZC, false>(x + S.off, y + S.off, S.g, tab, ku, make_int2(t.y, t.z),
                                      t.w);
    else
      dot = st_bulk_body<T, ZC, 4, false, TXT>(x + S.off, y + S.off, S.g, tab, ku,
                                               make_int2(t.y, t.z), t.w);
  }
  else
    dot = st_zb_body<T, DMA, TXT>(x + S.off, y + S.off, S.g, tab, t.y, t.z,
                                  t.w == kZbLow ? 0 : 1);
  const double bs = block_sum(dot, shm);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// y = R x (+ ridge) for every node of every active system, one node per
// thread (reference form of k_b_tiles); per-block partials of x.y.
template <typename T>
__global__ __launch_bounds__(256) void k_b_stencil(const T* __restrict__ x, T* __restrict__ y,
                                                   const BSys* __restrict__ sys,
                                                   const int* __restrict__ blk_sys,
                                                   const T* __restrict__ tab,
                                                   double* __restrict__ part,
                                                   const CgCtl* __restrict__ ctl) {
  __shared__ double sh[4];
  const int s = blk_sys[blockIdx.x];
  if (ctl[s].done) return;
  const BSys S = sys[s];
  const int64_t start = (int64_t)(blockIdx.x - S.b0) * kBNodes;
  const int64_t end = start + kBNodes < S.n ? start + kBNodes : S.n;
  const int nx = S.g.nx, ny = S.g.ny, nz = S.g.nz;
  const int64_t nxy = (int64_t)nx * ny;
  const T* xs = x + S.off;
  double dot = 0.0;
  for (int64_t id = start + threadIdx.x; id < end; id += 256) {
    const int k = (int)(id / nxy);
    const int64_t rr = id - (int64_t)k * nxy;
    const int j = (int)(rr / nx);
    const int i = (int)(rr - (int64_t)j * nx);
    const T* c = tab + ((cls7(k, nz) * 7 + cls7(j, ny)) * 7 + cls7(i, nx)) * kTaps;
    const T xc = xs[id];
    T acc = c[0] * xc;
#pragma unroll
    for (int q = 1; q < kTaps; ++q) {
      const int dx = kBTap[q][0], dy = kBTap[q][1], dz = kBTap[q][2];
      const bool ok = (unsigned)(i + dx) < (unsigned)nx && (unsigned)(j + dy) < (unsigned)ny &&
                      (unsigned)(k + dz) < (unsigned)nz;
      const int64_t o = ok ? id + dx + (int64_t)nx * dy + nxy * dz : id;
      const T xv = xs[o];
      acc += c[q] * (ok ? xv : (T)0);
    }
    y[S.off + id] = acc;
    dot += (double)xc * (double)acc;
  }
  const double bs = block_sum(dot, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

__global__ void k_b_diag(const BSys* __restrict__ sys, const int* __restrict__ blk_sys,
                         const double* __restrict__ tab, double* __restrict__ diag) {
  const int s = blk_sys[blockIdx.x];
  const BSys S = sys[s];
  const int64_t start = (int64_t)(blockIdx.x - S.b0) * kBNodes;
  const int64_t end = start + kBNodes < S.n ? start + kBNodes : S.n;
  const int64_t nxy = (int64_t)S.g.nx * S.g.ny;
  for (int64_t id = start + threadIdx.x; id < end; id += blockDim.x) {
    const int k = (int)(id / nxy);
    const int64_t rr = id - (int64_t)k * nxy;
    const int j = (int)(rr / S.g.nx);
    const int i = (int)(rr - (int64_t)j * S.g.nx);
    diag[S.off + id] =
        tab[((cls7(k, S.g.nz) * 7 + cls7(j, S.g.ny)) * 7 + cls7(i, S.g.nx)) * kTaps];
  }
}

// t_r = w_r^2 (a_r . x), u_r = t_r (a_r . x) (row part of p.Np)
template <typename T>
__global__ __launch_bounds__(256) void k_b_rows_t(const T* __restrict__ x, RowsDev R,
                                                  T* __restrict__ t, double* __restrict__ u) {
  const T* coef = (const T*)R.coef;
  const T* w2 = (const T*)R.w2;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R.nrows) return;
  T s = 0;
  for (int64_t e = R.row_ptr[r]; e < R.row_ptr[r + 1]; ++e) s += coef[e] * x[R.cols[e]];
  const T tv = w2[r] * s;
  t[r] = tv;
  u[r] = (double)tv * (double)s;
}

// x = 0, r = rhs (masked systems keep r = 0), partials of r.(minv r)
template <typename T>
__global__ __launch_bounds__(256) void k_b_init(const double* __restrict__ rhs,
                                                const T* __restrict__ minv, T* __restrict__ x,
                                                T* __restrict__ r, const BSys* __restrict__ sys,
                                                const int* __restrict__ blk_sys,
                                                const int* __restrict__ skip,
                                                double* __restrict__ part) {
  __shared__ double sh[4];
  const int s = blk_sys[blockIdx.x];
  const BSys S = sys[s];
  const int64_t start = S.off + (int64_t)(blockIdx.x - S.b0) * kBNodes;
  const int64_t end = start + kBNodes < S.off + S.n ? start + kBNodes : S.off + S.n;
  const bool sk = skip && skip[s];
  double acc = 0.0;
  for (int64_t i = start + threadIdx.x; i < end; i += 256) {
    const T rv = sk ? (T)0 : (T)rhs[i];
    x[i] = (T)0;
    r[i] = rv;
    acc += (double)rv * (double)(minv[i] * rv);
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// one workgroup per system
__global__ __launch_bounds__(256) void k_b_init_fin(const double* __restrict__ part,
                                                    const BSys* __restrict__ sys, double rtol,
                                                    double atol, int maxiter,
                                                    const int* __restrict__ skip,
                                                    CgCtl* __restrict__ ctl) {
  __shared__ double sh[4];
  const int s = blockIdx.x;
  const BSys S = sys[s];
  const double tot = reduce_partials(part + S.b0, S.b1 - S.b0, sh);
  if (threadIdx.x == 0) {
    CgCtl& c = ctl[s];
    const double bn = sqrt(tot);
    const double thr = fmax(atol, rtol * bn);
    c.rho = tot;
    c.rho_prev = 1.0;
    c.bnorm = bn;
    c.thr = thr;
    c.alpha = 0.0;
    c.iter = 0;
    c.maxiter = maxiter;
    c.converged = (sqrt(tot) < thr) || bn == 0.0;
    c.done = c.converged || (skip && skip[s]);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_b_dir(const T* __restrict__ r, const T* __restrict__ minv,
                                               T* __restrict__ p, const BSys* __restrict__ sys,
                                               const int* __restrict__ blk_sys,
                                               const CgCtl* __restrict__ ctl) {
  const int s = blk_sys[blockIdx.x];
  const CgCtl& c = ctl[s];
  if (c.done) return;
  const BSys S = sys[s];
  const int64_t start = S.off + (int64_t)(blockIdx.x - S.b0) * kBNodes;
  const int64_t end = start + kBNodes < S.off + S.n ? start + kBNodes : S.off + S.n;
  const int first = c.iter == 0;
  const T beta = first ? (T)0 : (T)(c.rho / c.rho_prev);
  for (int64_t i = start + threadIdx.x; i < end; i += 256) {
    const T z = minv[i] * r[i];
    p[i] = first ? z : z + beta * p[i];
  }
}

// one workgroup per system: pq = stencil partials + row terms; alpha
__global__ __launch_bounds__(256) void k_b_alpha(const double* __restrict__ part,
                                                 const double* __restrict__ u,
                                                 const BSys* __restrict__ sys,
                                                 CgCtl* __restrict__ ctl) {
  __shared__ double sh[4];
  const int s = blockIdx.x;
  if (ctl[s].done) return;
  const BSys S = sys[s];
  const double a = reduce_partials(part + S.t0, S.t1 - S.t0, sh);
  const double b = reduce_partials(u + S.r0, (int)(S.r1 - S.r0), sh);
  if (threadIdx.x == 0) {
    ctl[s].pq_st = a;
    ctl[s].pq_rows = b;
    ctl[s].alpha = ctl[s].rho / (a + b);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_b_upd(T* __restrict__ x, T* __restrict__ r,
                                               const T* __restrict__ p, const T* __restrict__ q,
                                               const T* __restrict__ minv,
                                               const BSys* __restrict__ sys,
                                               const int* __restrict__ blk_sys,
                                               double* __restrict__ part,
                                               const CgCtl* __restrict__ ctl) {
  __shared__ double sh[4];
  const int s = blk_sys[blockIdx.x];
  if (ctl[s].done) return;
  const BSys S = sys[s];
  const int64_t start = S.off + (int64_t)(blockIdx.x - S.b0) * kBNodes;
  const int64_t end = start + kBNodes < S.off + S.n ? start + kBNodes : S.off + S.n;
  const T alpha = (T)ctl[s].alpha;
  double acc = 0.0;
  for (int64_t i = start + threadIdx.x; i < end; i += 256) {
    x[i] += alpha * p[i];
    const T rv = r[i] - alpha * q[i];
    r[i] = rv;
    acc += (double)rv * (double)(minv[i] * rv);
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// one workgroup per system: rho update and stopping test
__global__ __launch_bounds__(256) void k_b_rho(const double* __restrict__ part,
                                               const BSys* __restrict__ sys,
                                               CgCtl* __restrict__ ctl) {
  __shared__ double sh[4];
  const int s = blockIdx.x;
  if (ctl[s].done) return;
  const BSys S = sys[s];
  const double tot = reduce_partials(part + S.b0, S.b1 - S.b0, sh);
  if (threadIdx.x == 0) {
    CgCtl& c = ctl[s];
    c.rho_prev = c.rho;
    c.rho = tot;
    const int it = c.iter + 1;
    c.iter = it;
    const int conv = sqrt(tot) < c.thr;
    c.converged = conv;
    c.done = conv || it >= c.maxiter;
  }
}

// r64 = rhs - q64 over active systems' nodes; partials of |r|_M^2
__global__ __launch_bounds__(256) void k_b_residual(const double* __restrict__ rhs,
                                                    const double* __restrict__ q,
                                                    const double* __restrict__ minv,
                                                    double* __restrict__ r,
                                                    const BSys* __restrict__ sys,
                                                    const int* __restrict__ blk_sys,
                                                    double* __restrict__ part) {
  __shared__ double sh[4];
  const int s = blk_sys[blockIdx.x];
  const BSys S = sys[s];
  const int64_t start = S.off + (int64_t)(blockIdx.x - S.b0) * kBNodes;
  const int64_t end = start + kBNodes < S.off + S.n ? start + kBNodes : S.off + S.n;
  double acc = 0.0;
  for (int64_t i = start + threadIdx.x; i < end; i += 256) {
    const double rv = rhs[i] - q[i];
    r[i] = rv;
    acc += rv * (minv[i] * rv);
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// one workgroup per system: outer residual, outer done flag
__global__ __launch_bounds__(256) void k_b_outer_fin(const double* __restrict__ part,
                                                     const BSys* __restrict__ sys,
                                                     const CgCtl* __restrict__ ctl64,
                                                     double* __restrict__ res,
                                                     int* __restrict__ odone) {
  __shared__ double sh[4];
  const int s = blockIdx.x;
  const BSys S = sys[s];
  const double tot = reduce_partials(part + S.b0, S.b1 - S.b0, sh);
  if (threadIdx.x == 0) {
    res[s] = sqrt(tot);
    odone[s] = odone[s] || (sqrt(tot) < ctl64[s].thr);
  }
}

// c += x32 over all nodes (systems excluded from the inner solve have x32 = 0)
__global__ void k_b_accum(const float* __restrict__ d, double* __restrict__ c, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] += (double)d[i];
}

}  // namespace lsamd
