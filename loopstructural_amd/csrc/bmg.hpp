// Batched multigrid: the flexible PCG with the geometric V-cycle of mg.hpp for
// many independent FD systems at once, one launch sequence per iteration.
//
// Each system keeps its own hierarchy, built by the single-system code (its
// grids, smoother, Chebyshev coefficients, coarse Gram matrices, dense
// coarsest inverse).  A batched kernel runs over a table of 256-node blocks
// of every system at one level (the blocks of a system contiguous) and reads
// the system's descriptor: pointers into its own vectors and operators.
// Blocks of systems that have converged return at once.  Coarse level
// operators are applied node by node (class-table stencil, or the exact
// per-node formula, plus the data term as a Gram row).  Level 0 of an
// ensemble whose systems share one class table takes the plane-streaming
// stencil over all systems (k_b_tiles in its pointer form, one table of
// every system's work items) and adds its data Gram rows after it; otherwise
// it is applied node by node too.  Systems of different depth: at level
// l the smoothing and transfer kernels cover the systems for which l lies
// above their coarsest level, the dense kernel those whose coarsest level is
// l, so each system sees the operations of its own V-cycle in order.
// Reductions: per-block partials, then one workgroup per system sums its own
// blocks in block order (deterministic).
#pragma once

namespace lsamd {

// one system's level: FP32 operator, smoother, vectors, Chebyshev
// coefficients; the dense inverse on the system's coarsest level
struct BmgLevel {
  MgDev A;
  MgSm S;
  float *b, *x, *q, *d;
  float c0, c1[kMgDegMax], c2[kMgDegMax];
  const float* cinv;
  int nc;
  const int32_t* g0list;  // level 0: the nodes with a nonempty Gram row
  int64_t g0n;
};

// one system's outer iteration (FP64 on level 0)
struct BmgSys {
  GridDev g;
  int64_t n;
  const double* tab;  // FP64 class table, nullptr: per-node formula with w
  RegW<double> w;
  const int64_t* gptr;  // data Gram matrix (nullptr: no data rows)
  const int32_t* gcol;
  const double* gval;
  double *x, *r, *rp, *p, *q;
  const double* minv;
  const double* rhs;
  double* out;
  int b0, b1;  // its blocks in the level-0 table
  int t0, t1;  // its work items in the level-0 stencil table (empty: node by node)
  const int32_t* g0list;  // the nodes with a nonempty Gram row
  int64_t g0n;
  int gb0, gb1;  // its partials of p.q in the node-block or Gram-list table
};

// y = N x at node id in FP64: the regulariser from the class table (or the
// per-node formula) and the data Gram row
__device__ inline double bmg_apply64(const BmgSys& S, const double* __restrict__ x, int64_t id) {
  int i, j, k;
  mg_ijk(id, S.g.nx, S.g.ny, i, j, k);
  double acc;
  if (S.tab) {
    const int nx = S.g.nx, ny = S.g.ny, nz = S.g.nz;
    const int64_t nxy = (int64_t)nx * ny;
    const double* c = S.tab + ((cls7(k, nz) * 7 + cls7(j, ny)) * 7 + cls7(i, nx)) * kTaps;
    acc = c[0] * x[id];
#pragma unroll
    for (int t = 1; t < kTaps; ++t) {
      const int dx = kBTap[t][0], dy = kBTap[t][1], dz = kBTap[t][2];
      const bool ok = (unsigned)(i + dx) < (unsigned)nx && (unsigned)(j + dy) < (unsigned)ny &&
                      (unsigned)(k + dz) < (unsigned)nz;
      const double xv = x[ok ? id + dx + (int64_t)nx * dy + nxy * dz : id];
      acc += c[t] * (ok ? xv : 0.0);
    }
  } else {
    acc = general_node<double>(x, S.g, S.w, nullptr, i, j, k);
  }
  if (S.gptr) {
    for (int64_t e = S.gptr[id]; e < S.gptr[id + 1]; ++e) acc += S.gval[e] * x[S.gcol[e]];
  }
  return acc;
}

// Chebyshev step of the smoother on the blocks of one level.  mode 0: first
// step from zero (d = c0 M b, x = d); 1: first step (d = c0 M (b - q),
// x += d); 2: step k (d = c1[k] d + c2[k] M (b - q), x += d; d not stored
// on the last step).
__global__ __launch_bounds__(256) void k_bmg_cheb(const BmgLevel* __restrict__ lv,
                                                  const int2* __restrict__ blk, int mode, int k,
                                                  int last, const CgCtl* __restrict__ ctl) {
  const int2 t = blk[blockIdx.x];
  if (ctl[t.x].done) return;
  const BmgLevel& L = lv[t.x];
  const int64_t i = (int64_t)t.y * 256 + threadIdx.x;
  if (i >= L.A.n) return;
  if (mode == 2) {
    const float dv = L.c1[k] * L.d[i] + L.c2[k] * mg_precond(i, L.S, L.b, L.q, 0);
    if (!last) L.d[i] = dv;
    L.x[i] += dv;
  } else {
    const float dv = L.c0 * mg_precond(i, L.S, L.b, L.q, mode == 0);
    L.d[i] = dv;
    L.x[i] = mode == 0 ? dv : L.x[i] + dv;
  }
}

// q = N x of one level (FP32)
__global__ __launch_bounds__(256) void k_bmg_apply(const BmgLevel* __restrict__ lv,
                                                   const int2* __restrict__ blk,
                                                   const CgCtl* __restrict__ ctl) {
  const int2 t = blk[blockIdx.x];
  if (ctl[t.x].done) return;
  const BmgLevel& L = lv[t.x];
  const int64_t i = (int64_t)t.y * 256 + threadIdx.x;
  if (i < L.A.n) L.q[i] = mg_apply_node(L.A, L.x, i);
}

// coarse b = P^T (b - q) of the fine level (blocks of the coarse level)
__global__ __launch_bounds__(256) void k_bmg_restrict(const BmgLevel* __restrict__ fine,
                                                      const BmgLevel* __restrict__ coarse,
                                                      const int2* __restrict__ blk,
                                                      const CgCtl* __restrict__ ctl) {
  const int2 t = blk[blockIdx.x];
  if (ctl[t.x].done) return;
  const BmgLevel& F = fine[t.x];
  const BmgLevel& C = coarse[t.x];
  const int64_t i = (int64_t)t.y * 256 + threadIdx.x;
  if (i < C.A.n) C.b[i] = mg_restrict_node(F.b, F.q, F.A.g, C.A.g, i);
}

// fine x += P (coarse x) (blocks of the fine level)
__global__ __launch_bounds__(256) void k_bmg_prolong(const BmgLevel* __restrict__ fine,
                                                     const BmgLevel* __restrict__ coarse,
                                                     const int2* __restrict__ blk,
                                                     const CgCtl* __restrict__ ctl) {
  const int2 t = blk[blockIdx.x];
  if (ctl[t.x].done) return;
  const BmgLevel& F = fine[t.x];
  const BmgLevel& C = coarse[t.x];
  const int64_t i = (int64_t)t.y * 256 + threadIdx.x;
  if (i < F.A.n) F.x[i] += mg_prolong_node(C.x, C.A.g, F.A.g, i);
}

// x = C^-1 b on the coarsest level of the listed systems: row blockIdx.x of
// system sys[blockIdx.y], one wavefront per row (the order of k_mg_dense)
__global__ __launch_bounds__(64) void k_bmg_dense(const BmgLevel* __restrict__ lv,
                                                  const int* __restrict__ sys,
                                                  const CgCtl* __restrict__ ctl) {
  const int s = sys[blockIdx.y];
  if (ctl[s].done) return;
  const BmgLevel& L = lv[s];
  const int r = blockIdx.x, n = L.nc;
  if (r >= n) return;
  float acc = 0.f;
  for (int c = threadIdx.x; c < n; c += 64) acc += L.cinv[(int64_t)r * n + c] * L.b[c];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (threadIdx.x == 0) L.x[r] = acc;
}

// ------------------------------------------------------ outer iteration ----
// x = 0, r = rhs, the V-cycle input (float) r; partials of rhs.(minv rhs)
__global__ __launch_bounds__(256) void k_bmg_init(const BmgSys* __restrict__ sys,
                                                  const BmgLevel* __restrict__ lv0,
                                                  const int2* __restrict__ blk,
                                                  double* __restrict__ part) {
  __shared__ double sh[4];
  const int2 t = blk[blockIdx.x];
  const BmgSys& S = sys[t.x];
  const int64_t i = (int64_t)t.y * 256 + threadIdx.x;
  double acc = 0.0;
  if (i < S.n) {
    const double rv = S.rhs[i];
    S.x[i] = 0.0;
    S.r[i] = rv;
    lv0[t.x].b[i] = (float)rv;
    acc = rv * (S.minv[i] * rv);
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// one workgroup per system: thresholds as k_cg_init_fin
__global__ __launch_bounds__(256) void k_bmg_init_fin(const BmgSys* __restrict__ sys,
                                                      const double* __restrict__ part,
                                                      double rtol, double atol, int maxiter,
                                                      CgCtl* __restrict__ ctl) {
  __shared__ double sh[4];
  const BmgSys& S = sys[blockIdx.x];
  const double tot = reduce_partials(part + S.b0, S.b1 - S.b0, sh);
  if (threadIdx.x == 0) {
    CgCtl& c = ctl[blockIdx.x];
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
    c.done = c.converged;
  }
}

// partials of z.r and z.r_prev (z: the V-cycle output)
__global__ __launch_bounds__(256) void k_bmg_dots(const BmgSys* __restrict__ sys,
                                                  const BmgLevel* __restrict__ lv0,
                                                  const int2* __restrict__ blk,
                                                  double* __restrict__ part,
                                                  const CgCtl* __restrict__ ctl) {
  __shared__ double sh[4];
  const int2 t = blk[blockIdx.x];
  if (ctl[t.x].done) return;
  const BmgSys& S = sys[t.x];
  const int64_t i = (int64_t)t.y * 256 + threadIdx.x;
  double a = 0.0, b = 0.0;
  if (i < S.n) {
    const double zi = (double)lv0[t.x].x[i];
    a = zi * S.r[i];
    b = zi * S.rp[i];
  }
  const double sa = block_sum(a, sh);
  const double sb = block_sum(b, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sa;
    part[2 * blockIdx.x + 1] = sb;
  }
}

// per system: rho = z.r, beta = (z.r - z.r_prev) / rho_prev (0 first)
__global__ __launch_bounds__(256) void k_bmg_dots_fin(const BmgSys* __restrict__ sys,
                                                      const double* __restrict__ part,
                                                      CgCtl* __restrict__ ctl) {
  __shared__ double sh[4];
  CgCtl& c = ctl[blockIdx.x];
  if (c.done) return;
  const BmgSys& S = sys[blockIdx.x];
  double a = 0.0, b = 0.0;
  for (int i = S.b0 + threadIdx.x; i < S.b1; i += 256) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  const double sa = block_sum(a, sh);
  const double sb = block_sum(b, sh);
  if (threadIdx.x == 0) {
    c.acc = c.iter == 0 ? 0.0 : (sa - sb) / c.rho_prev;
    c.rho = sa;
  }
}

// p = z + beta p
__global__ __launch_bounds__(256) void k_bmg_dir(const BmgSys* __restrict__ sys,
                                                 const BmgLevel* __restrict__ lv0,
                                                 const int2* __restrict__ blk,
                                                 const CgCtl* __restrict__ ctl) {
  const int2 t = blk[blockIdx.x];
  const CgCtl& c = ctl[t.x];
  if (c.done) return;
  const BmgSys& S = sys[t.x];
  const int64_t i = (int64_t)t.y * 256 + threadIdx.x;
  if (i >= S.n) return;
  const double z = (double)lv0[t.x].x[i];
  S.p[i] = c.iter == 0 ? z : z + c.acc * S.p[i];
}

// q = N p (FP64); partials of p.q
__global__ __launch_bounds__(256) void k_bmg_apply64(const BmgSys* __restrict__ sys,
                                                     const int2* __restrict__ blk,
                                                     double* __restrict__ part,
                                                     const CgCtl* __restrict__ ctl) {
  __shared__ double sh[4];
  const int2 t = blk[blockIdx.x];
  if (ctl[t.x].done) return;
  const BmgSys& S = sys[t.x];
  const int64_t i = (int64_t)t.y * 256 + threadIdx.x;
  double acc = 0.0;
  if (i < S.n) {
    const double qv = bmg_apply64(S, S.p, i);
    S.q[i] = qv;
    acc = S.p[i] * qv;
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// per system: alpha = rho / p.q
// level 0 after the plane-streaming stencil: y += G0 x (the data Gram rows,
// FP32) over the listed nodes of the smoothed systems (blocks of 256 list
// entries)
__global__ __launch_bounds__(256) void k_bmg_gram0(const BmgLevel* __restrict__ lv,
                                                   const int2* __restrict__ blk,
                                                   const CgCtl* __restrict__ ctl) {
  const int2 t = blk[blockIdx.x];
  if (ctl[t.x].done) return;
  const BmgLevel& L = lv[t.x];
  const int64_t k = (int64_t)t.y * 256 + threadIdx.x;
  if (k >= L.g0n) return;
  const int64_t i = L.g0list[k];
  const int64_t e0 = L.A.gptr[i], e1 = L.A.gptr[i + 1];
  float acc = 0.f;
  for (int64_t e = e0; e < e1; ++e) acc += L.A.gval[e] * L.x[L.A.gcol[e]];
  L.q[i] += acc;
}

// the same in FP64 for the outer iteration: q += G0 p, per-block partials of
// p.(G0 p) (the stencil's own partials hold p.(R p))
__global__ __launch_bounds__(256) void k_bmg_gram64(const BmgSys* __restrict__ sys,
                                                    const int2* __restrict__ blk,
                                                    double* __restrict__ part,
                                                    const CgCtl* __restrict__ ctl) {
  __shared__ double sh[4];
  const int2 t = blk[blockIdx.x];
  if (ctl[t.x].done) return;
  const BmgSys& S = sys[t.x];
  const int64_t k = (int64_t)t.y * 256 + threadIdx.x;
  double acc = 0.0;
  if (k < S.g0n) {
    const int64_t i = S.g0list[k];
    const int64_t e0 = S.gptr[i], e1 = S.gptr[i + 1];
    double g = 0.0;
    for (int64_t e = e0; e < e1; ++e) g += S.gval[e] * S.p[S.gcol[e]];
    S.q[i] += g;
    acc = S.p[i] * g;
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// per system: alpha = rho / p.q, p.q summed from its partials [gb0, gb1)
// (node blocks, or Gram-list blocks in the plane-streaming form) and, in that
// form, its work items' partials
__global__ __launch_bounds__(256) void k_bmg_alpha(const BmgSys* __restrict__ sys,
                                                   const double* __restrict__ part,
                                                   const double* __restrict__ tpart,
                                                   CgCtl* __restrict__ ctl) {
  __shared__ double sh[4];
  CgCtl& c = ctl[blockIdx.x];
  if (c.done) return;
  const BmgSys& S = sys[blockIdx.x];
  double pq = reduce_partials(part + S.gb0, S.gb1 - S.gb0, sh);
  __syncthreads();
  if (S.t1 > S.t0) pq += reduce_partials(tpart + S.t0, S.t1 - S.t0, sh);
  if (threadIdx.x == 0) c.alpha = c.rho / pq;
}

// x += alpha p ; r_prev = r ; r -= alpha q ; V-cycle input (float) r ;
// partials of r.(minv r)
__global__ __launch_bounds__(256) void k_bmg_upd(const BmgSys* __restrict__ sys,
                                                 const BmgLevel* __restrict__ lv0,
                                                 const int2* __restrict__ blk,
                                                 double* __restrict__ part,
                                                 const CgCtl* __restrict__ ctl) {
  __shared__ double sh[4];
  const int2 t = blk[blockIdx.x];
  const CgCtl& c = ctl[t.x];
  if (c.done) return;
  const BmgSys& S = sys[t.x];
  const int64_t i = (int64_t)t.y * 256 + threadIdx.x;
  double acc = 0.0;
  if (i < S.n) {
    const double alpha = c.alpha;
    S.x[i] += alpha * S.p[i];
    const double rv = S.r[i];
    S.rp[i] = rv;
    const double rn = rv - alpha * S.q[i];
    S.r[i] = rn;
    lv0[t.x].b[i] = (float)rn;
    acc = rn * (S.minv[i] * rn);
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// per system: rho_prev = rho ; iter++ ; stopping test on |S r|
__global__ __launch_bounds__(256) void k_bmg_upd_fin(const BmgSys* __restrict__ sys,
                                                     const double* __restrict__ part,
                                                     CgCtl* __restrict__ ctl) {
  __shared__ double sh[4];
  CgCtl& c = ctl[blockIdx.x];
  if (c.done) return;
  const BmgSys& S = sys[blockIdx.x];
  const double tot = reduce_partials(part + S.b0, S.b1 - S.b0, sh);
  if (threadIdx.x == 0) {
    c.rho_prev = c.rho;
    const int it = c.iter + 1;
    c.iter = it;
    c.res2 = tot;
    const int conv = sqrt(tot) < c.thr;
    c.converged = conv;
    c.done = conv || it >= c.maxiter;
  }
}

// flag = every system done
__global__ __launch_bounds__(256) void k_bmg_all_done(const CgCtl* __restrict__ ctl, int nsys,
                                                      int* __restrict__ flag) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  for (int s = threadIdx.x; s < nsys; s += 256)
    if (!ctl[s].done) any = 1;
  __syncthreads();
  if (threadIdx.x == 0) *flag = any ? 0 : 1;
}

// solution out (every system)
__global__ __launch_bounds__(256) void k_bmg_out(const BmgSys* __restrict__ sys,
                                                 const int2* __restrict__ blk) {
  const int2 t = blk[blockIdx.x];
  const BmgSys& S = sys[t.x];
  const int64_t i = (int64_t)t.y * 256 + threadIdx.x;
  if (i < S.n) S.out[i] = S.x[i];
}

}  // namespace lsamd
