// LSMR (Fong & Saunders, SIAM J. Sci. Comput. 33 (2011), the algorithm of
// scipy.sparse.linalg.lsmr) on the rectangular FD least-squares operator,
// matrix-free.  The operator is A S with the reference's row blocks: data
// rows w_r a_r, regulariser rows sqrt(W_f(i)) d_f(i) (an interior row and the
// border rows of the same mask at one node are merged into one row of the
// summed squared weight, which leaves A^T A, A^T b and |b| unchanged and with
// them the LSMR iterates), ridge rows ridge * e_i; S = diag(A^T A)^-1/2 is
// the column scaling.  The scalar recurrences run in one-thread kernels on a
// device control block so that chunks of iterations replay from a graph.
#pragma once

namespace lsamd {

struct LsCtl {
  // Golub-Kahan and rotation state
  double alpha, beta, inv_beta, fwd_coef;
  double zetabar, alphabar, rho, rhobar, cbar, sbar, zeta;
  double betadd, betad, rhodold, tautildeold, thetatilde, d;
  double normA2, maxrbar, minrbar, normA, condA, normx, normr, normar, normb;
  // coefficients of the vector updates of this iteration
  double inv_alpha, c_hbar, c_x, c_h;
  double atol, btol, ctol;
  int itn, maxiter, istop, done;
  int skipv;  // beta == 0 this step: v and alpha kept (as scipy's lsmr)
};

// (c, s, r) with c a + s b = r, -s a + c b = 0 (stable)
__device__ inline void sym_ortho(double a, double b, double& c, double& s, double& r) {
  if (b == 0.0) {
    c = a == 0.0 ? 1.0 : (a > 0.0 ? 1.0 : -1.0);
    s = 0.0;
    r = fabs(a);
  } else if (a == 0.0) {
    c = 0.0;
    s = b > 0.0 ? 1.0 : -1.0;
    r = fabs(b);
  } else if (fabs(b) > fabs(a)) {
    const double tau = a / b;
    s = (b > 0.0 ? 1.0 : -1.0) / sqrt(1.0 + tau * tau);
    c = s * tau;
    r = b / s;
  } else {
    const double tau = b / a;
    c = (a > 0.0 ? 1.0 : -1.0) / sqrt(1.0 + tau * tau);
    s = c * tau;
    r = a / c;
  }
}

__device__ inline void ls_ijk(int64_t id, const GridDev& g, int& i, int& j, int& k) {
  const int64_t nxy = (int64_t)g.nx * g.ny;
  k = (int)(id / nxy);
  const int64_t r = id - (int64_t)k * nxy;
  j = (int)(r / g.nx);
  i = (int)(r - (int64_t)j * g.nx);
}

// value of the normalised row of family f centred at (i,j,k)
__device__ inline double ls_row(const double* x, const GridDev& g, int f, int i, int j, int k) {
  const double is6 = 0.40824829046386301637;
  switch (f) {
    case F_XX:
      return is6 * (ldx(x, g, i - 1, j, k) - 2.0 * ldx(x, g, i, j, k) + ldx(x, g, i + 1, j, k));
    case F_YY:
      return is6 * (ldx(x, g, i, j - 1, k) - 2.0 * ldx(x, g, i, j, k) + ldx(x, g, i, j + 1, k));
    case F_ZZ:
      return is6 * (ldx(x, g, i, j, k - 1) - 2.0 * ldx(x, g, i, j, k) + ldx(x, g, i, j, k + 1));
    case F_XY:
      return 0.5 * (-ldx(x, g, i - 1, j - 1, k) + ldx(x, g, i + 1, j - 1, k) +
                    ldx(x, g, i - 1, j + 1, k) - ldx(x, g, i + 1, j + 1, k));
    case F_XZ:
      return 0.5 * (-ldx(x, g, i - 1, j, k - 1) + ldx(x, g, i + 1, j, k - 1) +
                    ldx(x, g, i - 1, j, k + 1) - ldx(x, g, i + 1, j, k + 1));
    default:
      return 0.5 * (-ldx(x, g, i, j - 1, k - 1) + ldx(x, g, i, j + 1, k - 1) +
                    ldx(x, g, i, j - 1, k + 1) - ldx(x, g, i, j + 1, k + 1));
  }
}

// xs = S v
__global__ void k_ls_scale(const double* __restrict__ v, const double* __restrict__ minv,
                           double* __restrict__ xs, int64_t n, const LsCtl* ctl) {
  if (ctl->done) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xs[i] = sqrt(minv[i]) * v[i];
}

// regulariser and ridge blocks: u = A_reg xs - fwd_coef u (init: u = 0 - 0)
__global__ __launch_bounds__(256) void k_ls_reg_fwd(const double* __restrict__ xs, GridDev g,
                                                    RegW<double> w, const double* scale2,
                                                    double* __restrict__ ureg,
                                                    double* __restrict__ urid, int64_t n,
                                                    int init, double* __restrict__ part,
                                                    const LsCtl* ctl) {
  if (ctl->done) return;
  __shared__ double sh[4];
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (id < n) {
    int i, j, k;
    ls_ijk(id, g, i, j, k);
    const double a = init ? 0.0 : ctl->fwd_coef;
    for (int f = 0; f < 6; ++f) {
      const double W = row_weight(g, w, scale2, f, i, j, k);
      const double un = init ? 0.0
                             : (W > 0.0 ? sqrt(W) * ls_row(xs, g, f, i, j, k) : 0.0) -
                                   a * ureg[f * n + id];
      ureg[f * n + id] = un;
      acc += un * un;
    }
    const double un = init ? 0.0 : sqrt(w.ridge) * xs[id] - a * urid[id];
    urid[id] = un;
    acc += un * un;
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// data block: u_r = w_r (a_r . xs) - fwd_coef u_r, or u_r = w_r b_r (init)
__global__ __launch_bounds__(256) void k_ls_data_fwd(const double* __restrict__ xs, RowsDev R,
                                                     const double* __restrict__ w,
                                                     const double* __restrict__ b,
                                                     double* __restrict__ ud, int init,
                                                     double* __restrict__ part,
                                                     const LsCtl* ctl) {
  if (ctl->done) return;
  __shared__ double sh[4];
  const double* coef = (const double*)R.coef;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  if (r < R.nrows) {
    double un;
    if (init) {
      un = w[r] * b[r];
    } else {
      double s = 0.0;
      for (int64_t e = R.row_ptr[r]; e < R.row_ptr[r + 1]; ++e) s += coef[e] * xs[R.cols[e]];
      un = w[r] * s - ctl->fwd_coef * ud[r];
    }
    ud[r] = un;
    acc = un * un;
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// beta = |u| from the partials of both blocks (one block of 256)
__global__ __launch_bounds__(256) void k_ls_beta(const double* __restrict__ p1, int n1,
                                                 const double* __restrict__ p2, int n2,
                                                 LsCtl* ctl, int init) {
  if (ctl->done) return;
  __shared__ double sh[4];
  const double a = reduce_partials(p1, n1, sh);
  const double b = reduce_partials(p2, n2, sh);
  if (threadIdx.x != 0) return;
  const double beta = sqrt(a + b);
  if (init) {
    ctl->beta = beta;
    ctl->inv_beta = beta > 0.0 ? 1.0 / beta : 0.0;
    ctl->normb = beta;
    ctl->normr = beta;
    ctl->skipv = 0;
    if (!(beta > 0.0)) {
      ctl->istop = 0;
      ctl->done = 1;  // b = 0: x = 0
    }
  } else {
    ctl->beta = beta;
    ctl->inv_beta = beta > 0.0 ? 1.0 / beta : 0.0;
    ctl->skipv = beta > 0.0 ? 0 : 1;
  }
}

// t_r = w_r u_r / beta (data block of A^T u)
__global__ void k_ls_data_t(const double* __restrict__ ud, const double* __restrict__ w,
                            double* __restrict__ t, int64_t nrows, const LsCtl* ctl) {
  if (ctl->done) return;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nrows) t[r] = w[r] * ud[r] * ctl->inv_beta;
}

// y = (A_reg^T u_reg + ridge u_rid) / beta
__global__ void k_ls_reg_t(const double* __restrict__ ureg, const double* __restrict__ urid,
                           GridDev g, RegW<double> w, const double* scale2,
                           double* __restrict__ y, int64_t n, const LsCtl* ctl) {
  if (ctl->done) return;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  int i, j, k;
  ls_ijk(id, g, i, j, k);
  const double is6 = 0.40824829046386301637;
  auto U = [&](int f, int a, int b, int c) -> double {
    if (!in_grid(g, a, b, c)) return 0.0;
    const double W = row_weight(g, w, scale2, f, a, b, c);
    return W > 0.0 ? sqrt(W) * ureg[f * n + nid(g, a, b, c)] : 0.0;
  };
  double y0 = sqrt(w.ridge) * urid[id];
  for (int d = -1; d <= 1; ++d) {
    const double c = (d == 0 ? -2.0 : 1.0) * is6;
    y0 += c * (U(F_XX, i - d, j, k) + U(F_YY, i, j - d, k) + U(F_ZZ, i, j, k - d));
  }
  for (int du = -1; du <= 1; du += 2)
    for (int dv = -1; dv <= 1; dv += 2) {
      const double c = (du * dv > 0) ? -0.5 : 0.5;
      y0 += c * (U(F_XY, i - du, j - dv, k) + U(F_XZ, i - du, j, k - dv) +
                 U(F_YZ, i, j - du, k - dv));
    }
  y[id] = y0 * ctl->inv_beta;
}

// v = S y - beta v (init: v = S y, the old v never read); partials of |v|^2
__global__ __launch_bounds__(256) void k_ls_vupd(const double* __restrict__ y,
                                                 const double* __restrict__ minv,
                                                 double* __restrict__ v, int64_t n, int init,
                                                 double* __restrict__ part, const LsCtl* ctl) {
  if (ctl->done || ctl->skipv) return;
  __shared__ double sh[4];
  const double beta = ctl->beta;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double sy = sqrt(minv[i]) * y[i];
    const double vn = init ? sy : sy - beta * v[i];  // v is not read on init
    v[i] = vn;
    acc += vn * vn;
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// alpha = |v| and the rotations of one LSMR step (one block of 256, the
// scalar work in thread 0)
__global__ __launch_bounds__(256) void k_ls_step(const double* __restrict__ part, int np,
                                                 LsCtl* c, int init) {
  if (c->done) return;
  __shared__ double sh[4];
  const double s2 = reduce_partials(part, np, sh);
  if (threadIdx.x != 0) return;
  double alpha;
  if (!init && c->skipv) {
    alpha = c->alpha;  // v already normalised
    c->inv_alpha = 1.0;
  } else {
    alpha = sqrt(s2);
    c->inv_alpha = alpha > 0.0 ? 1.0 / alpha : 0.0;
  }
  if (init) {
    c->alpha = alpha;
    c->zetabar = alpha * c->beta;
    c->alphabar = alpha;
    c->rho = 1.0;
    c->rhobar = 1.0;
    c->cbar = 1.0;
    c->sbar = 0.0;
    c->zeta = 0.0;
    c->betadd = c->beta;
    c->betad = 0.0;
    c->rhodold = 1.0;
    c->tautildeold = 0.0;
    c->thetatilde = 0.0;
    c->d = 0.0;
    c->normA2 = alpha * alpha;
    c->maxrbar = 0.0;
    c->minrbar = 1e100;
    c->normA = sqrt(c->normA2);
    c->condA = 1.0;
    c->normx = 0.0;
    c->normar = alpha * c->beta;
    c->itn = 0;
    c->istop = 0;
    // h = v (scaled), hbar = 0, x = 0
    c->c_hbar = 0.0;
    c->c_x = 0.0;
    c->c_h = 0.0;
    c->fwd_coef = alpha * c->inv_beta;  // next forward subtracts alpha * (u / beta)
    if (c->normar == 0.0) c->done = 1;
    return;
  }
  const double beta = c->beta;
  // Qhat (damp = 0): chat = 1, shat = 0, alphahat = alphabar
  double chat, shat, alphahat;
  sym_ortho(c->alphabar, 0.0, chat, shat, alphahat);
  const double rhoold = c->rho;
  double cc, ss, rho;
  sym_ortho(alphahat, beta, cc, ss, rho);
  const double thetanew = ss * alpha;
  c->alphabar = cc * alpha;
  const double rhobarold = c->rhobar;
  const double zetaold = c->zeta;
  const double thetabar = c->sbar * rho;
  const double rhotemp = c->cbar * rho;
  double cbar, sbar, rhobar;
  sym_ortho(c->cbar * rho, thetanew, cbar, sbar, rhobar);
  c->cbar = cbar;
  c->sbar = sbar;
  c->rhobar = rhobar;
  c->rho = rho;
  c->zeta = cbar * c->zetabar;
  c->zetabar = -sbar * c->zetabar;
  // vector updates: hbar = h - c_hbar hbar ; x += c_x hbar ; h = v - c_h h
  c->c_hbar = thetabar * rho / (rhoold * rhobarold);
  c->c_x = c->zeta / (rho * rhobar);
  c->c_h = thetanew / rho;
  // |r| estimate
  const double betaacute = chat * c->betadd;
  const double betacheck = -shat * c->betadd;
  const double betahat = cc * betaacute;
  c->betadd = -ss * betaacute;
  const double thetatildeold = c->thetatilde;
  double ctildeold, stildeold, rhotildeold;
  sym_ortho(c->rhodold, thetabar, ctildeold, stildeold, rhotildeold);
  c->thetatilde = stildeold * rhobar;
  c->rhodold = ctildeold * rhobar;
  c->betad = -stildeold * c->betad + ctildeold * betahat;
  c->tautildeold = (zetaold - thetatildeold * c->tautildeold) / rhotildeold;
  const double taud = (c->zeta - c->thetatilde * c->tautildeold) / c->rhodold;
  c->d = c->d + betacheck * betacheck;
  c->normr = sqrt(c->d + (c->betad - taud) * (c->betad - taud) + c->betadd * c->betadd);
  // |A| and cond(A) estimates
  c->normA2 = c->normA2 + beta * beta;
  c->normA = sqrt(c->normA2);
  c->normA2 = c->normA2 + alpha * alpha;
  c->maxrbar = fmax(c->maxrbar, rhobarold);
  if (c->itn >= 1) c->minrbar = fmin(c->minrbar, rhobarold);  // itn counts finished steps
  c->condA = fmax(c->maxrbar, rhotemp) / fmin(c->minrbar, rhotemp);
  c->normar = fabs(c->zetabar);
  c->alpha = alpha;
}

// v /= alpha ; hbar = h - c_hbar hbar ; x += c_x hbar ; h = v - c_h h ;
// partials of |x|^2   (init: v /= alpha ; h = v ; hbar = 0 ; x = 0)
__global__ __launch_bounds__(256) void k_ls_xupd(double* __restrict__ v, double* __restrict__ h,
                                                 double* __restrict__ hbar,
                                                 double* __restrict__ x, int64_t n, int init,
                                                 double* __restrict__ part, const LsCtl* c) {
  if (c->done) return;
  __shared__ double sh[4];
  const double ia = c->inv_alpha, ch = c->c_hbar, cx = c->c_x, chh = c->c_h;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double vn = v[i] * ia;
    v[i] = vn;
    if (init) {
      h[i] = vn;
      hbar[i] = 0.0;
      x[i] = 0.0;
    } else {
      const double hb = h[i] - ch * hbar[i];
      hbar[i] = hb;
      const double xn = x[i] + cx * hb;
      x[i] = xn;
      acc += xn * xn;
      h[i] = vn - chh * h[i];
    }
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// |x| and the stopping tests, istop numbered as scipy's lsmr (one block)
__global__ __launch_bounds__(256) void k_ls_test(const double* __restrict__ part, int np,
                                                 LsCtl* c) {
  if (c->done) return;
  __shared__ double sh[4];
  const double s2 = reduce_partials(part, np, sh);
  if (threadIdx.x != 0) return;
  c->normx = sqrt(s2);
  c->itn += 1;
  c->fwd_coef = c->alpha * c->inv_beta;
  const double test1 = c->normr / c->normb;
  const double test2 = (c->normA * c->normr) != 0.0 ? c->normar / (c->normA * c->normr)
                                                   : __builtin_inf();
  const double test3 = 1.0 / c->condA;
  const double t1 = test1 / (1.0 + c->normA * c->normx / c->normb);
  const double rtol = c->btol + c->atol * c->normA * c->normx / c->normb;
  int istop = 0;
  if (c->itn >= c->maxiter) istop = 7;
  if (1.0 + test3 <= 1.0) istop = 6;
  if (1.0 + test2 <= 1.0) istop = 5;
  if (1.0 + t1 <= 1.0) istop = 4;
  if (test3 <= c->ctol) istop = 3;
  if (test2 <= c->atol) istop = 2;
  if (test1 <= rtol) istop = 1;
  c->istop = istop;
  if (istop > 0) c->done = 1;
}

// c = S x
__global__ void k_ls_out(const double* __restrict__ x, const double* __restrict__ minv,
                         double* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = sqrt(minv[i]) * x[i];
}

}  // namespace lsamd
