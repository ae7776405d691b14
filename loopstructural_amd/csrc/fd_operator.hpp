// Normal-equations operator N = A_d^T W^2 A_d + sum_f D_f^T diag(W_f) D_f + ridge2 I
// of the finite-difference interpolator on a structured grid (device side).
//
// Regulariser semantics (LI/_finite_difference_interpolator.py:643-692,
// 1052-1164; LC/supports/_3d_structured_grid.py:164-293,489-510;
// LC/math/finite_difference_stencil.py:12-38):
//  * six interior families xx,yy,zz (mask [1,-2,1]/sqrt6 after row
//    normalisation) and xy,xz,yz (+-1/2 at the four (+-1,+-1) corners), one
//    row per node strictly interior on all three axes, weight w_f
//    (w_xx = dxx, w_xy = dxy/4);
//  * border families dx/dy/dz lower/upper: the same axial mask at EVERY node
//    whose index on that axis is 1 or n-2 (both when n == 3), weight dx/dy/dz;
//  * optional per-node scale s (use_regularisation_weight_scale): every row
//    centred at node i has weight s_i * w.
// Row i of family f contributes W_f(i) d_f(i) d_f(i)^T with W_f(i) the sum
// of squared weights of the rows of f centred at i.
#pragma once

#include "lsamd_common.hpp"

namespace lsamd {

template <typename T>
struct RegW {
  T xx, yy, zz, xy, xz, yz, bx, by, bz, ridge;
};

// 25-tap composite kernel K = sum_f W_f autocorr(d_f) + ridge2 delta, valid
// at nodes whose distance to every face is >= 3 (all rows around them are
// interior rows of uniform weight).  Taps: centre, +-1 and +-2 on each axis,
// (+-2,+-2) in each coordinate plane.
template <typename T>
struct Composite {
  T c0, cx1, cy1, cz1, cx2, cy2, cz2, cxy, cxz, cyz;
};

template <typename T>
inline Composite<T> make_composite(const lsamd_fd_reg& r) {
  Composite<T> k;
  k.c0 = (T)(r.w2_xx + r.w2_yy + r.w2_zz + r.w2_xy + r.w2_xz + r.w2_yz + r.ridge2);
  k.cx1 = (T)(-2.0 / 3.0 * r.w2_xx);
  k.cy1 = (T)(-2.0 / 3.0 * r.w2_yy);
  k.cz1 = (T)(-2.0 / 3.0 * r.w2_zz);
  k.cx2 = (T)(r.w2_xx / 6.0 - 0.5 * r.w2_xy - 0.5 * r.w2_xz);
  k.cy2 = (T)(r.w2_yy / 6.0 - 0.5 * r.w2_xy - 0.5 * r.w2_yz);
  k.cz2 = (T)(r.w2_zz / 6.0 - 0.5 * r.w2_xz - 0.5 * r.w2_yz);
  k.cxy = (T)(0.25 * r.w2_xy);
  k.cxz = (T)(0.25 * r.w2_xz);
  k.cyz = (T)(0.25 * r.w2_yz);
  return k;
}

template <typename T>
inline RegW<T> make_regw(const lsamd_fd_reg& r) {
  RegW<T> w;
  w.xx = (T)r.w2_xx;
  w.yy = (T)r.w2_yy;
  w.zz = (T)r.w2_zz;
  w.xy = (T)r.w2_xy;
  w.xz = (T)r.w2_xz;
  w.yz = (T)r.w2_yz;
  w.bx = (T)r.w2_bx;
  w.by = (T)r.w2_by;
  w.bz = (T)r.w2_bz;
  w.ridge = (T)r.ridge2;
  return w;
}

__device__ inline bool in_grid(const GridDev& g, int i, int j, int k) {
  return i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
}
__device__ inline int64_t nid(const GridDev& g, int i, int j, int k) {
  return (int64_t)i + (int64_t)g.nx * ((int64_t)j + (int64_t)g.ny * k);
}

enum Fam { F_XX = 0, F_YY, F_ZZ, F_XY, F_XZ, F_YZ };

// Squared weight of the rows of family f centred at node (i,j,k); 0 when
// no such row exists (including nodes outside the grid).
template <typename T>
__device__ inline T row_weight(const GridDev& g, const RegW<T>& w, const double* scale2, int f,
                               int i, int j, int k) {
  if (!in_grid(g, i, j, k)) return (T)0;
  const bool I = i >= 1 && i <= g.nx - 2 && j >= 1 && j <= g.ny - 2 && k >= 1 && k <= g.nz - 2;
  T r = (T)0;
  switch (f) {
    case F_XX:
      r = (I ? w.xx : (T)0);
      if (g.nx >= 3) r += (i == 1 ? w.bx : (T)0) + (i == g.nx - 2 ? w.bx : (T)0);
      break;
    case F_YY:
      r = (I ? w.yy : (T)0);
      if (g.ny >= 3) r += (j == 1 ? w.by : (T)0) + (j == g.ny - 2 ? w.by : (T)0);
      break;
    case F_ZZ:
      r = (I ? w.zz : (T)0);
      if (g.nz >= 3) r += (k == 1 ? w.bz : (T)0) + (k == g.nz - 2 ? w.bz : (T)0);
      break;
    case F_XY:
      r = I ? w.xy : (T)0;
      break;
    case F_XZ:
      r = I ? w.xz : (T)0;
      break;
    default:
      r = I ? w.yz : (T)0;
      break;
  }
  if (scale2 && r != (T)0) r *= (T)scale2[nid(g, i, j, k)];
  return r;
}

template <typename T>
__device__ inline T ldx(const T* x, const GridDev& g, int i, int j, int k) {
  return in_grid(g, i, j, k) ? x[nid(g, i, j, k)] : (T)0;
}

// Exact y = (R^T W R + ridge2 I) x at one node, by the two-pass formula
// y_j = sum_f sum_o c_f(o) W_f(j-o) (d_f . x)(j-o).
template <typename T>
__device__ inline T general_node(const T* __restrict__ x, const GridDev& g, const RegW<T>& w,
                                 const double* scale2, int i, int j, int k) {
  const T is6 = (T)0.40824829046386301637;  // 1/sqrt(6)
  const T h = (T)0.5;
  T y = w.ridge * ldx(x, g, i, j, k);
  // axial families: column j sits at offset d in the row centred at j-d
  for (int d = -1; d <= 1; ++d) {
    const T c = (d == 0 ? (T)-2 : (T)1) * is6;
    T W = row_weight(g, w, scale2, F_XX, i - d, j, k);
    if (W != (T)0) {
      const int a = i - d;
      y += c * W * is6 *
           (ldx(x, g, a - 1, j, k) - (T)2 * ldx(x, g, a, j, k) + ldx(x, g, a + 1, j, k));
    }
    W = row_weight(g, w, scale2, F_YY, i, j - d, k);
    if (W != (T)0) {
      const int a = j - d;
      y += c * W * is6 *
           (ldx(x, g, i, a - 1, k) - (T)2 * ldx(x, g, i, a, k) + ldx(x, g, i, a + 1, k));
    }
    W = row_weight(g, w, scale2, F_ZZ, i, j, k - d);
    if (W != (T)0) {
      const int a = k - d;
      y += c * W * is6 *
           (ldx(x, g, i, j, a - 1) - (T)2 * ldx(x, g, i, j, a) + ldx(x, g, i, j, a + 1));
    }
  }
  // cross families: d(u,v) = 1/2 (-x[-1,-1] + x[+1,-1] + x[-1,+1] - x[+1,+1])
  for (int du = -1; du <= 1; du += 2) {
    for (int dv = -1; dv <= 1; dv += 2) {
      const T c = (du * dv > 0) ? -h : h;  // coefficient at offset (du, dv)
      T W = row_weight(g, w, scale2, F_XY, i - du, j - dv, k);
      if (W != (T)0) {
        const int a = i - du, b = j - dv;
        const T r = h * (-ldx(x, g, a - 1, b - 1, k) + ldx(x, g, a + 1, b - 1, k) +
                         ldx(x, g, a - 1, b + 1, k) - ldx(x, g, a + 1, b + 1, k));
        y += c * W * r;
      }
      W = row_weight(g, w, scale2, F_XZ, i - du, j, k - dv);
      if (W != (T)0) {
        const int a = i - du, b = k - dv;
        const T r = h * (-ldx(x, g, a - 1, j, b - 1) + ldx(x, g, a + 1, j, b - 1) +
                         ldx(x, g, a - 1, j, b + 1) - ldx(x, g, a + 1, j, b + 1));
        y += c * W * r;
      }
      W = row_weight(g, w, scale2, F_YZ, i, j - du, k - dv);
      if (W != (T)0) {
        const int a = j - du, b = k - dv;
        const T r = h * (-ldx(x, g, i, a - 1, b - 1) + ldx(x, g, i, a + 1, b - 1) +
                         ldx(x, g, i, a - 1, b + 1) - ldx(x, g, i, a + 1, b + 1));
        y += c * W * r;
      }
    }
  }
  return y;
}

// diag(R^T W R + ridge2 I) at one node.
__device__ inline double general_diag(const GridDev& g, const RegW<double>& w,
                                      const double* scale2, int i, int j, int k) {
  double y = w.ridge;
  for (int d = -1; d <= 1; ++d) {
    const double c2 = (d == 0 ? 4.0 : 1.0) / 6.0;
    y += c2 * row_weight(g, w, scale2, F_XX, i - d, j, k);
    y += c2 * row_weight(g, w, scale2, F_YY, i, j - d, k);
    y += c2 * row_weight(g, w, scale2, F_ZZ, i, j, k - d);
  }
  for (int du = -1; du <= 1; du += 2)
    for (int dv = -1; dv <= 1; dv += 2) {
      y += 0.25 * row_weight(g, w, scale2, F_XY, i - du, j - dv, k);
      y += 0.25 * row_weight(g, w, scale2, F_XZ, i - du, j, k - dv);
      y += 0.25 * row_weight(g, w, scale2, F_YZ, i, j - du, k - dv);
    }
  return y;
}

// Boundary shell = nodes at distance <= 2 from a face; the bulk (composite
// kernel) covers i in [3, n-4] on every axis.
struct Shell {
  int64_t count;
  int mx, my, mz;  // bulk extent per axis (n - 6), <= 0 means no bulk
  bool all;        // every node handled by the general formula
};

inline Shell make_shell(const GridDev& g, bool force_all) {
  Shell s;
  s.mx = g.nx - 6;
  s.my = g.ny - 6;
  s.mz = g.nz - 6;
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  s.all = force_all || s.mx <= 0 || s.my <= 0 || s.mz <= 0;
  s.count = s.all ? n : n - (int64_t)s.mx * s.my * s.mz;
  return s;
}

// shell index -> node (i,j,k)
__device__ inline void shell_decode(const GridDev& g, const Shell& s, int64_t t, int& i, int& j,
                                    int& k) {
  const int64_t nxy = (int64_t)g.nx * g.ny;
  if (s.all) {
    k = (int)(t / nxy);
    const int64_t r = t - (int64_t)k * nxy;
    j = (int)(r / g.nx);
    i = (int)(r - (int64_t)j * g.nx);
    return;
  }
  if (t < 3 * nxy) {
    k = (int)(t / nxy);
  } else if (t < 6 * nxy) {
    t -= 3 * nxy;
    k = g.nz - 3 + (int)(t / nxy);
  } else {
    t -= 6 * nxy;
    const int64_t ring = nxy - (int64_t)s.mx * s.my;
    k = 3 + (int)(t / ring);
    int64_t r = t - (int64_t)(k - 3) * ring;
    if (r < 3 * (int64_t)g.nx) {
      j = (int)(r / g.nx);
      i = (int)(r - (int64_t)j * g.nx);
    } else if (r < 6 * (int64_t)g.nx) {
      r -= 3 * (int64_t)g.nx;
      j = g.ny - 3 + (int)(r / g.nx);
      i = (int)(r - (int64_t)(j - (g.ny - 3)) * g.nx);
    } else {
      r -= 6 * (int64_t)g.nx;
      j = 3 + (int)(r / 6);
      const int c = (int)(r - (int64_t)(j - 3) * 6);
      i = c < 3 ? c : g.nx - 6 + c;
    }
    return;
  }
  const int64_t r = t - (t / nxy) * nxy;
  j = (int)(r / g.nx);
  i = (int)(r - (int64_t)j * g.nx);
}

}  // namespace lsamd
