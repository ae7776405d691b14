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

__host__ __device__ inline bool in_grid(const GridDev& g, int i, int j, int k) {
  return i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
}
__host__ __device__ inline int64_t nid(const GridDev& g, int i, int j, int k) {
  return (int64_t)i + (int64_t)g.nx * ((int64_t)j + (int64_t)g.ny * k);
}

enum Fam { F_XX = 0, F_YY, F_ZZ, F_XY, F_XZ, F_YZ };

// Squared weight of the rows of family f centred at node (i,j,k); 0 when
// no such row exists (including nodes outside the grid).
template <typename T>
__host__ __device__ inline T row_weight(const GridDev& g, const RegW<T>& w, const double* scale2, int f,
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

// ------------------------------------------------------- class table ----
// On a grid with >= 7 nodes per axis, the operator row of node j depends on
// j only through the "class" of each coordinate: 0,1,2 (distance to the low
// face), 3 (distance >= 3 from both faces), 4,5,6 (n-3, n-2, n-1).  A node's
// 25 coefficients are therefore one of 7^3 sets, computed exactly on a
// virtual 7x7x7 grid whose node (cx,cy,cz) has the same row layout around it.
// Tap order: in-plane 0:(0,0) 1:(+1,0) 2:(-1,0) 3:(0,+1) 4:(0,-1) 5:(+2,0)
// 6:(-2,0) 7:(0,+2) 8:(0,-2) 9:(+2,+2) 10:(-2,+2) 11:(+2,-2) 12:(-2,-2);
// 13:(0,0,-1) 14:(0,0,+1); 15..19 plane z-2: (0,0),(+2,0),(-2,0),(0,+2),
// (0,-2); 20..24 plane z+2: same.  Taps reaching outside the grid have
// coefficient 0 (every row touching them is absent).
constexpr int kTaps = 25;
constexpr int kClasses = 343;

__host__ __device__ inline int cls7(int i, int n) { return i < 3 ? i : (i > n - 4 ? i - (n - 7) : 3); }

inline void build_class_table(const lsamd_fd_reg& r, double* tab) {
  static const int taps[kTaps][3] = {
      {0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},  {2, 0, 0},  {-2, 0, 0},
      {0, 2, 0},  {0, -2, 0}, {2, 2, 0},  {-2, 2, 0}, {2, -2, 0},  {-2, -2, 0}, {0, 0, -1},
      {0, 0, 1},  {0, 0, -2}, {2, 0, -2}, {-2, 0, -2}, {0, 2, -2}, {0, -2, -2}, {0, 0, 2},
      {2, 0, 2},  {-2, 0, 2}, {0, 2, 2},  {0, -2, 2}};
  struct Off {
    int d[3];
    double c;
  };
  Off fam[6][4];
  int nfam[6] = {3, 3, 3, 4, 4, 4};
  const double s6 = 0.40824829046386301637;  // 1/sqrt(6)
  for (int a = 0; a < 3; ++a)
    for (int t = -1; t <= 1; ++t) {
      Off& o = fam[a][t + 1];
      o.d[0] = o.d[1] = o.d[2] = 0;
      o.d[a] = t;
      o.c = (t == 0 ? -2.0 : 1.0) * s6;
    }
  const int pl[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int f = 0; f < 3; ++f) {
    int q = 0;
    for (int u = -1; u <= 1; u += 2)
      for (int v = -1; v <= 1; v += 2) {
        Off& o = fam[3 + f][q++];
        o.d[0] = o.d[1] = o.d[2] = 0;
        o.d[pl[f][0]] = u;
        o.d[pl[f][1]] = v;
        o.c = (u * v > 0) ? -0.5 : 0.5;
      }
  }
  // squared row weight of family f centred at (i,j,k) of the 7-grid
  auto W = [&](int f, int i, int j, int k) {
    if (i < 0 || i > 6 || j < 0 || j > 6 || k < 0 || k > 6) return 0.0;
    const bool I = i >= 1 && i <= 5 && j >= 1 && j <= 5 && k >= 1 && k <= 5;
    switch (f) {
      case 0: return (I ? r.w2_xx : 0.0) + (i == 1 ? r.w2_bx : 0.0) + (i == 5 ? r.w2_bx : 0.0);
      case 1: return (I ? r.w2_yy : 0.0) + (j == 1 ? r.w2_by : 0.0) + (j == 5 ? r.w2_by : 0.0);
      case 2: return (I ? r.w2_zz : 0.0) + (k == 1 ? r.w2_bz : 0.0) + (k == 5 ? r.w2_bz : 0.0);
      case 3: return I ? r.w2_xy : 0.0;
      case 4: return I ? r.w2_xz : 0.0;
      default: return I ? r.w2_yz : 0.0;
    }
  };
  for (int cz = 0; cz < 7; ++cz)
    for (int cy = 0; cy < 7; ++cy)
      for (int cx = 0; cx < 7; ++cx) {
        double* out = tab + ((cz * 7 + cy) * 7 + cx) * kTaps;
        for (int t = 0; t < kTaps; ++t) {
          const int* dd = taps[t];
          double acc = (t == 0) ? r.ridge2 : 0.0;
          for (int f = 0; f < 6; ++f)
            for (int a = 0; a < nfam[f]; ++a) {
              // row centred at j - o1 holds column j at offset o1 and column
              // j + dd at offset o1 + dd
              const Off& o1 = fam[f][a];
              double c2 = 0.0;
              for (int b = 0; b < nfam[f]; ++b)
                if (fam[f][b].d[0] == o1.d[0] + dd[0] && fam[f][b].d[1] == o1.d[1] + dd[1] &&
                    fam[f][b].d[2] == o1.d[2] + dd[2])
                  c2 = fam[f][b].c;
              if (c2 == 0.0) continue;
              acc += o1.c * W(f, cx - o1.d[0], cy - o1.d[1], cz - o1.d[2]) * c2;
            }
          out[t] = acc;
        }
      }
}

}  // namespace lsamd
