// P1 (piecewise-linear, 5 tets per cell) point location, rows, evaluation
// and edge-jump rows.  Compiled with -ffp-contract=off (and the pragma) so
// that the barycentric coordinates and evaluated values round like the
// reference's numpy expressions: tet selection, coordinates and values are
// bit-identical.  Element gradients use the adjugate inverse, which agrees
// with numpy's LU-based inverse to a few ulp.
//
// Reference: LC/supports/_3d_structured_tetra.py:208-262 (evaluation),
// :264-365 (inside, location), :376-545 (elements, gradients), :626-740
// (neighbours), :120-194 (face table and normals); LI/_p1interpolator.py:
// 73-113 (normal and value rows), :115-166 (edge jumps), :248-296 (gradient
// orthogonal rows); LI/_discrete_interpolator.py:361-366 (row normalisation).
#pragma clang fp contract(off)

#include <math.h>

#include <map>
#include <vector>

#include "p1_mesh.hpp"

namespace lsamd {
namespace {

// np.einsum('ikj,ikj->ik') over 3 components: (p0 + p2) + p1
__device__ inline double dot3(const double* a, const double* b) {
  return (a[0] * b[0] + a[2] * b[2]) + a[1] * b[1];
}
__device__ inline void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

struct TetLoc {
  bool inside;
  int cell[3];
  int loc;
  double bc[4];
};

// get_element_for_location for one point: strict box test, cell lookup,
// barycentric coordinates by scalar triple products in all 5 tets, the last
// tet with all coordinates >= 0 wins.
__device__ TetLoc p1_locate(const P1Mesh& m, const double* p) {
  TetLoc L;
  bool ins = true;
  for (int a = 0; a < 3; ++a) {
    ins = ins && (p[a] > m.o[a]);
    ins = ins && (p[a] < m.o[a] + m.h[a] * (double)m.c[a]);
  }
  for (int a = 0; a < 3; ++a) {
    L.cell[a] = 0;
    if (ins) {
      long long c = (long long)floor((p[a] - m.o[a]) / m.h[a]);
      if (c < 0) c = 0;
      if (c > m.c[a] - 1) c = m.c[a] - 1;
      L.cell[a] = (int)c;
    }
  }
  L.loc = -1;
  for (int q = 0; q < 4; ++q) L.bc[q] = __builtin_nan("");
  for (int t = 0; t < 5; ++t) {
    double V[4][3];
    tet_vertices(m, L.cell, t, V);
    double vap[3], vbp[3], vab[3], vac[3], vad[3], vbc[3], vbd[3], cr[3];
    for (int a = 0; a < 3; ++a) {
      vap[a] = p[a] - V[0][a];
      vbp[a] = p[a] - V[1][a];
      vab[a] = V[1][a] - V[0][a];
      vac[a] = V[2][a] - V[0][a];
      vad[a] = V[3][a] - V[0][a];
      vbc[a] = V[2][a] - V[1][a];
      vbd[a] = V[3][a] - V[1][a];
    }
    cross3(vbd, vbc, cr);
    const double va = dot3(vbp, cr) / 6.0;
    cross3(vac, vad, cr);
    const double vb = dot3(vap, cr) / 6.0;
    const double v = dot3(vab, cr) / 6.0;
    cross3(vad, vab, cr);
    const double vc = dot3(vap, cr) / 6.0;
    cross3(vab, vac, cr);
    const double vd = dot3(vap, cr) / 6.0;
    const double c0 = va / v, c1 = vb / v, c2 = vc / v, c3 = vd / v;
    if (c0 >= 0 && c1 >= 0 && c2 >= 0 && c3 >= 0) {
      L.loc = t;
      L.bc[0] = c0;
      L.bc[1] = c1;
      L.bc[2] = c2;
      L.bc[3] = c3;
    }
  }
  L.inside = ins && L.loc >= 0;
  return L;
}

__device__ inline int64_t loc_tet(const P1Mesh& m, const TetLoc& L) {
  return L.loc + 5 * p1_cell(m, L.cell[0], L.cell[1], L.cell[2]);
}

__global__ void k_p1_locate(P1Mesh m, const double* __restrict__ pts, int64_t n,
                            int64_t* __restrict__ tet, int64_t* __restrict__ nodes,
                            double* __restrict__ bc, double* __restrict__ grad,
                            uint8_t* __restrict__ inside) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  const TetLoc L = p1_locate(m, p);
  const double nan = __builtin_nan("");
  if (inside) inside[i] = L.inside ? 1 : 0;
  if (tet) tet[i] = L.inside ? loc_tet(m, L) : -1;
  int64_t nd[4] = {-1, -1, -1, -1};
  if (L.inside) tet_nodes(m, loc_tet(m, L), nd);
  if (nodes)
    for (int q = 0; q < 4; ++q) nodes[4 * i + q] = nd[q];
  if (bc)
    for (int q = 0; q < 4; ++q) bc[4 * i + q] = L.inside ? L.bc[q] : nan;
  if (grad) {
    double G[3][4];
    if (L.inside) {
      double V[4][3];
      tet_vertices(m, L.cell, L.loc, V);
      grad_from_vertices(V, G);
    }
    for (int d = 0; d < 3; ++d)
      for (int q = 0; q < 4; ++q) grad[12 * i + 4 * d + q] = L.inside ? G[d][q] : nan;
  }
}

__global__ void k_p1_elements(P1Mesh m, int64_t ntet, int64_t* __restrict__ elements,
                              int64_t* __restrict__ neighbours) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntet) return;
  if (elements) {
    int64_t nd[4];
    tet_nodes(m, t, nd);
    for (int q = 0; q < 4; ++q) elements[4 * t + q] = nd[q];
  }
  if (neighbours) {
    const int loc = (int)(t % 5);
    const int64_t cid = t / 5;
    if (loc == 0) {
      for (int q = 0; q < 4; ++q) neighbours[4 * t + q] = t + 1 + q;
      return;
    }
    int cell[3];
    p1_cell_ijk(m, cid, cell);
    const int odd = (cell[0] + cell[1] + cell[2]) & 1;
    neighbours[4 * t] = 5 * cid;
    for (int a = 0; a < 3; ++a) {
      int nc[3] = {cell[0], cell[1], cell[2]};
      nc[a] += nb_step(odd, loc, a);
      const bool ok = nc[a] >= 0 && nc[a] < m.c[a];
      neighbours[4 * t + 1 + a] =
          ok ? 5 * p1_cell(m, nc[0], nc[1], nc[2]) + nb_local(odd, loc, a) : -1;
    }
  }
}

// Gradients (and optionally volumes) of given tets; a negative id counts from
// the end like numpy indexing.
__global__ void k_p1_tet_grads(P1Mesh m, const int64_t* __restrict__ tets, int64_t n, int64_t ntet,
                               double* __restrict__ grads, double* __restrict__ size) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t t = tets[i];
  if (t < 0) t += ntet;
  if (t < 0 || t >= ntet) t = 0;  // rejected on the host for the known-size case
  if (grads) {
    double G[3][4];
    tet_grad(m, t, G);
    for (int d = 0; d < 3; ++d)
      for (int q = 0; q < 4; ++q) grads[12 * i + 4 * d + q] = G[d][q];
  }
  if (size) size[i] = tet_volume(m, t);
}

// One thread per cell: the pairs the cell owns, in reference order.
__global__ void k_p1_pairs(P1Mesh m, int64_t* __restrict__ pairs, int64_t* __restrict__ faces,
                           int32_t* __restrict__ cols, double* __restrict__ coef,
                           double* __restrict__ w, double* __restrict__ norm,
                           const double* __restrict__ scale, double cgw) {
  const int64_t ncell = (int64_t)m.c[0] * m.c[1] * m.c[2];
  const int64_t cid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cid >= ncell) return;
  int cell[3];
  p1_cell_ijk(m, cid, cell);
  int64_t r = p1_pair_base(m, cell[0], cell[1], cell[2]);
  p1_cell_pairs(m, cell, [&](int64_t t1, int64_t t2) {
    if (pairs) {
      pairs[2 * r] = t1;
      pairs[2 * r + 1] = t2;
    }
    if (faces || cols || coef || w || norm) {
      int64_t col[8], f[3];
      double a[8];
      p1_edge_row(m, t1, t2, col, a, f);
      if (faces)
        for (int q = 0; q < 3; ++q) faces[3 * r + q] = f[q];
      if (cols)
        for (int q = 0; q < 8; ++q) cols[8 * r + q] = (int32_t)col[q];
      const double len = p1_norm8(a);
      if (norm) norm[r] = len;
      if (coef)
        for (int q = 0; q < 8; ++q) coef[8 * r + q] = len > 0 ? a[q] / len : a[q];
      if (w) {
        double rw = cgw;
        if (scale) {
          const double st = (((scale[col[0]] + scale[col[1]]) + scale[col[2]]) + scale[col[3]]) / 4;
          const double sn = (((scale[col[4]] + scale[col[5]]) + scale[col[6]]) + scale[col[7]]) / 4;
          rw = 0.5 * (st + sn) * cgw;
        }
        w[r] = rw;
      }
    }
    ++r;
  });
}

// Rows for points inside the mesh (1 value row, or 3 normal rows point-major,
// or 1 gradient-orthogonal row per point).  Per-point weights are ignored,
// as in the reference.
__global__ void k_p1_build_rows(P1Mesh m, int kind, const double* __restrict__ pts, int64_t n,
                                int64_t ld, double wscale, double bval, int32_t* __restrict__ cols,
                                double* __restrict__ coef, double* __restrict__ b,
                                double* __restrict__ w, double* __restrict__ norm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* p = pts + i * ld;
  const TetLoc L = p1_locate(m, p);
  int64_t nd[4] = {0, 0, 0, 0};
  double G[3][4];
  const int64_t t = L.inside ? loc_tet(m, L) : 0;
  if (L.inside) tet_nodes(m, t, nd);
  if (kind == LSAMD_ROWS_VALUE) {
    double a[4];
    for (int q = 0; q < 4; ++q) a[q] = L.inside ? L.bc[q] : 0.0;
    const double len = sqrt(((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]) + a[3] * a[3]);
    for (int q = 0; q < 4; ++q) {
      cols[4 * i + q] = (int32_t)nd[q];
      coef[4 * i + q] = len > 0 ? a[q] / len : a[q];
    }
    b[i] = len > 0 ? p[3] / len : p[3];
    w[i] = wscale;
    if (norm) norm[i] = len;
    return;
  }
  if (L.inside) {
    double V[4][3];
    tet_vertices(m, L.cell, L.loc, V);
    grad_from_vertices(V, G);
  } else {
    for (int d = 0; d < 3; ++d)
      for (int q = 0; q < 4; ++q) G[d][q] = 0.0;
  }
  if (kind == LSAMD_ROWS_NORMAL) {
    for (int d = 0; d < 3; ++d) {
      const int64_t r = 3 * i + d;
      const double* a = G[d];
      const double len = sqrt(((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]) + a[3] * a[3]);
      for (int q = 0; q < 4; ++q) {
        cols[4 * r + q] = (int32_t)nd[q];
        coef[4 * r + q] = len > 0 ? a[q] / len : a[q];
      }
      b[r] = len > 0 ? p[3 + d] / len : p[3 + d];
      w[r] = wscale;
      if (norm) norm[r] = len;
    }
    return;
  }
  // gradient orthogonal: unit vector . grad, weight scaled by the tet volume
  double v[3] = {p[3], p[4], p[5]};
  const double vn = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
  if (vn > 0)
    for (int d = 0; d < 3; ++d) v[d] = v[d] / vn;
  double a[4];
  for (int q = 0; q < 4; ++q) a[q] = (v[0] * G[0][q] + v[1] * G[1][q]) + v[2] * G[2][q];
  const double len = sqrt(((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]) + a[3] * a[3]);
  for (int q = 0; q < 4; ++q) {
    cols[4 * i + q] = (int32_t)nd[q];
    coef[4 * i + q] = len > 0 ? a[q] / len : a[q];
  }
  b[i] = len > 0 ? bval / len : bval;
  w[i] = wscale * (L.inside ? tet_volume(m, t) : 0.0);
  if (norm) norm[i] = len;
}

__global__ void k_p1_evaluate(P1Mesh m, const double* __restrict__ c,
                              const double* __restrict__ pts, int64_t n,
                              double* __restrict__ values, double* __restrict__ grads) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
  const TetLoc L = p1_locate(m, p);
  const double nan = __builtin_nan("");
  if (!L.inside) {
    if (values) values[i] = nan;
    if (grads)
      for (int d = 0; d < 3; ++d) grads[3 * i + d] = nan;
    return;
  }
  int64_t nd[4];
  tet_nodes(m, loc_tet(m, L), nd);
  const double cv[4] = {c[nd[0]], c[nd[1]], c[nd[2]], c[nd[3]]};
  if (values)
    values[i] = ((L.bc[0] * cv[0] + L.bc[1] * cv[1]) + L.bc[2] * cv[2]) + L.bc[3] * cv[3];
  if (grads) {
    double V[4][3], G[3][4];
    tet_vertices(m, L.cell, L.loc, V);
    grad_from_vertices(V, G);
    for (int d = 0; d < 3; ++d)
      grads[3 * i + d] = ((G[d][0] * cv[0] + G[d][1] * cv[1]) + G[d][2] * cv[2]) + G[d][3] * cv[3];
  }
}

int check_mesh(const lsamd_grid* g, const char* fn) {
  if (int rc = check_uniform_grid(g, fn)) return rc;
  const int64_t ntet = 5 * (g->nnodes[0] - 1) * (g->nnodes[1] - 1) * (g->nnodes[2] - 1);
  if (ntet > ((int64_t)1 << 40)) return fail(LSAMD_ERR_ARG, std::string(fn) + ": mesh too large");
  return LSAMD_OK;
}

}  // namespace

void build_p1_table(const lsamd_grid& g, double w2, double* tab, int* nv) {
  lsamd_grid vg = g;
  for (int a = 0; a < 3; ++a) {
    nv[a] = p1_virtual_n((int)g.nnodes[a]);
    vg.nnodes[a] = nv[a];
    vg.origin[a] = 0.0;
  }
  const P1Mesh m = to_p1(vg);
  const int64_t nn = (int64_t)nv[0] * nv[1] * nv[2];
  for (int64_t q = 0; q < nn * kP1Taps; ++q) tab[q] = 0.0;
  int tap_of[5][5][5];
  for (int x = 0; x < 5; ++x)
    for (int y = 0; y < 5; ++y)
      for (int z = 0; z < 5; ++z) tap_of[x][y][z] = -1;
  for (int q = 0; q < kP1Taps; ++q) {
    int d[3];
    p1_tap_offset(q, d);
    tap_of[d[0] + 2][d[1] + 2][d[2] + 2] = q;
  }
  const int64_t ncell = (int64_t)m.c[0] * m.c[1] * m.c[2];
  for (int64_t cid = 0; cid < ncell; ++cid) {
    int cell[3];
    p1_cell_ijk(m, cid, cell);
    p1_cell_pairs(m, cell, [&](int64_t t1, int64_t t2) {
      int64_t col[8];
      double a[8];
      p1_edge_row(m, t1, t2, col, a, nullptr);
      const double len = p1_norm8(a);
      if (len > 0)
        for (int q = 0; q < 8; ++q) a[q] /= len;
      // merge repeated nodes
      int64_t u[8];
      double r[8];
      int nu = 0;
      for (int q = 0; q < 8; ++q) {
        int k = 0;
        while (k < nu && u[k] != col[q]) ++k;
        if (k == nu) {
          u[nu] = col[q];
          r[nu++] = a[q];
        } else {
          r[k] += a[q];
        }
      }
      for (int p = 0; p < nu; ++p) {
        const int64_t np_ = u[p];
        const int pi = (int)(np_ % nv[0]), pj = (int)(np_ / nv[0] % nv[1]),
                  pk = (int)(np_ / nv[0] / nv[1]);
        for (int q = 0; q < nu; ++q) {
          const int64_t nq = u[q];
          const int dx = (int)(nq % nv[0]) - pi, dy = (int)(nq / nv[0] % nv[1]) - pj,
                    dz = (int)(nq / nv[0] / nv[1]) - pk;
          const int tp = (abs(dx) <= 2 && abs(dy) <= 2 && abs(dz) <= 2)
                             ? tap_of[dx + 2][dy + 2][dz + 2]
                             : -1;
          if (tp >= 0) tab[np_ * kP1Taps + tp] += w2 * r[p] * r[q];
        }
      }
    });
  }
}

}  // namespace lsamd

using namespace lsamd;

extern "C" int lsamd_p1_locate(const lsamd_grid* g, const double* pts, int64_t n, int64_t* tet,
                               int64_t* nodes, double* bc, double* grad, uint8_t* inside,
                               void* stream) {
  if (int rc = injected("p1_locate")) return rc;
  if (int rc = check_mesh(g, "lsamd_p1_locate")) return rc;
  if (n < 0) return fail(LSAMD_ERR_ARG, "lsamd_p1_locate: n < 0");
  if (n == 0) return LSAMD_OK;
  if (!pts) return fail(LSAMD_ERR_ARG, "lsamd_p1_locate: pts is NULL");
  hipLaunchKernelGGL(k_p1_locate, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, to_p1(*g), pts, n, tet, nodes, bc, grad, inside);
  LSAMD_CHECK_LAUNCH("k_p1_locate");
  return LSAMD_OK;
}

extern "C" int lsamd_p1_elements(const lsamd_grid* g, int64_t* elements, int64_t* neighbours,
                                 void* stream) {
  if (int rc = injected("p1_elements")) return rc;
  if (int rc = check_mesh(g, "lsamd_p1_elements")) return rc;
  const P1Mesh m = to_p1(*g);
  const int64_t ntet = 5 * (int64_t)m.c[0] * m.c[1] * m.c[2];
  hipLaunchKernelGGL(k_p1_elements, dim3((unsigned)ceil_div(ntet, 256)), dim3(256), 0,
                     (hipStream_t)stream, m, ntet, elements, neighbours);
  LSAMD_CHECK_LAUNCH("k_p1_elements");
  return LSAMD_OK;
}

extern "C" int lsamd_p1_tet_gradients(const lsamd_grid* g, const int64_t* tets, int64_t n,
                                      double* grads, double* size, void* stream) {
  if (int rc = injected("p1_tet_gradients")) return rc;
  if (int rc = check_mesh(g, "lsamd_p1_tet_gradients")) return rc;
  if (n < 0) return fail(LSAMD_ERR_ARG, "lsamd_p1_tet_gradients: n < 0");
  if (n == 0) return LSAMD_OK;
  if (!tets) return fail(LSAMD_ERR_ARG, "lsamd_p1_tet_gradients: tets is NULL");
  const P1Mesh m = to_p1(*g);
  const int64_t ntet = 5 * (int64_t)m.c[0] * m.c[1] * m.c[2];
  hipLaunchKernelGGL(k_p1_tet_grads, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, m, tets, n, ntet, grads, size);
  LSAMD_CHECK_LAUNCH("k_p1_tet_grads");
  return LSAMD_OK;
}

extern "C" int64_t lsamd_p1_n_pairs(const lsamd_grid* g) {
  if (check_mesh(g, "lsamd_p1_n_pairs")) return -1;
  return p1_n_pairs(to_p1(*g));
}

extern "C" int lsamd_p1_pairs(const lsamd_grid* g, int64_t* pairs, int64_t* faces, int32_t* cols,
                              double* coef, double* w, double* norm, const double* scale,
                              double cgw, void* stream) {
  if (int rc = injected("p1_pairs")) return rc;
  if (int rc = check_mesh(g, "lsamd_p1_pairs")) return rc;
  if (cols && g->nnodes[0] * g->nnodes[1] * g->nnodes[2] > 2147483647LL)
    return fail(LSAMD_ERR_ARG, "lsamd_p1_pairs: mesh exceeds int32 node ids");
  const P1Mesh m = to_p1(*g);
  const int64_t ncell = (int64_t)m.c[0] * m.c[1] * m.c[2];
  hipLaunchKernelGGL(k_p1_pairs, dim3((unsigned)ceil_div(ncell, 128)), dim3(128), 0,
                     (hipStream_t)stream, m, pairs, faces, cols, coef, w, norm, scale, cgw);
  LSAMD_CHECK_LAUNCH("k_p1_pairs");
  return LSAMD_OK;
}

extern "C" int lsamd_p1_build_rows(const lsamd_grid* g, int32_t kind, const double* pts, int64_t n,
                                   int64_t ld, double wscale, double bval, int32_t* cols,
                                   double* coef, double* b, double* w, double* norm, void* stream) {
  if (int rc = injected("p1_build_rows")) return rc;
  if (int rc = check_mesh(g, "lsamd_p1_build_rows")) return rc;
  if (kind != LSAMD_ROWS_VALUE && kind != LSAMD_ROWS_NORMAL && kind != LSAMD_ROWS_TANGENT)
    return fail(LSAMD_ERR_ARG, "lsamd_p1_build_rows: kind must be VALUE, NORMAL or TANGENT");
  const int need = kind == LSAMD_ROWS_VALUE ? 4 : (kind == LSAMD_ROWS_NORMAL ? 6 : 6);
  if (ld < need) return fail(LSAMD_ERR_ARG, "lsamd_p1_build_rows: ld too small for row kind");
  if (n < 0) return fail(LSAMD_ERR_ARG, "lsamd_p1_build_rows: n < 0");
  if (n == 0) return LSAMD_OK;
  if (!pts || !cols || !coef || !b || !w)
    return fail(LSAMD_ERR_ARG, "lsamd_p1_build_rows: NULL argument");
  if (g->nnodes[0] * g->nnodes[1] * g->nnodes[2] > 2147483647LL)
    return fail(LSAMD_ERR_ARG, "lsamd_p1_build_rows: mesh exceeds int32 node ids");
  hipLaunchKernelGGL(k_p1_build_rows, dim3((unsigned)ceil_div(n, 128)), dim3(128), 0,
                     (hipStream_t)stream, to_p1(*g), (int)kind, pts, n, ld, wscale, bval, cols,
                     coef, b, w, norm);
  LSAMD_CHECK_LAUNCH("k_p1_build_rows");
  return LSAMD_OK;
}

extern "C" int lsamd_p1_operator_table(const lsamd_grid* g, double cgw, double* tab, int32_t* nv) {
  if (int rc = check_mesh(g, "lsamd_p1_operator_table")) return rc;
  if (!tab || !nv) return fail(LSAMD_ERR_ARG, "lsamd_p1_operator_table: NULL argument");
  int v[3];
  build_p1_table(*g, cgw * cgw, tab, v);
  for (int a = 0; a < 3; ++a) nv[a] = v[a];
  return LSAMD_OK;
}

extern "C" int lsamd_p1_evaluate(const lsamd_grid* g, const double* c, const double* pts,
                                 int64_t n, double* values, double* grads, void* stream) {
  if (int rc = injected("p1_evaluate")) return rc;
  if (int rc = check_mesh(g, "lsamd_p1_evaluate")) return rc;
  if (n < 0) return fail(LSAMD_ERR_ARG, "lsamd_p1_evaluate: n < 0");
  if (n == 0) return LSAMD_OK;
  if (!c || !pts) return fail(LSAMD_ERR_ARG, "lsamd_p1_evaluate: NULL input");
  hipLaunchKernelGGL(k_p1_evaluate, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, to_p1(*g), c, pts, n, values, grads);
  LSAMD_CHECK_LAUNCH("k_p1_evaluate");
  return LSAMD_OK;
}
