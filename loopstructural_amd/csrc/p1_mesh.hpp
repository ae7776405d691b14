// Structured 5-tet mesh shared by the P1 kernels and the host-side
// operator table.  Geometry and topology follow
// LC/supports/_3d_structured_tetra.py: each cell of an axis-aligned grid is
// split into 5 tets, the split chosen by the parity of i+j+k (:28-34); tet
// id = local + 5*cell with cells x fastest (:376-419); every non-centre tet
// has one neighbour per axis in the adjacent cell (:626-740).
#pragma once

#include <math.h>
#include <stdint.h>

#include "lsamd_common.hpp"

#define LSAMD_HD __host__ __device__ inline

namespace lsamd {

struct P1Mesh {
  double o[3], h[3];
  int n[3];  // nodes per axis
  int c[3];  // cells per axis
};

inline P1Mesh to_p1(const lsamd_grid& g) {
  P1Mesh m;
  for (int a = 0; a < 3; ++a) {
    m.o[a] = g.origin[a];
    m.h[a] = g.step[a];
    m.n[a] = (int)g.nnodes[a];
    m.c[a] = (int)g.nnodes[a] - 1;
  }
  return m;
}

// 3 bits per entry, entry 0 in the low bits.
constexpr uint64_t pack3() { return 0; }
template <typename... R>
constexpr uint64_t pack3(int a, R... r) {
  return (uint64_t)a | (pack3(r...) << 3);
}
// corner ids (000,100,010,110,001,101,011,111 order) of the 5 tets
constexpr uint64_t kTetEven = pack3(7, 1, 2, 4, 6, 2, 4, 7, 5, 1, 4, 7, 0, 1, 2, 4, 3, 1, 2, 7);
constexpr uint64_t kTetOdd = pack3(0, 6, 5, 3, 7, 3, 5, 6, 4, 0, 5, 6, 2, 0, 3, 6, 1, 0, 3, 5);

LSAMD_HD int tet_corner(int odd, int t, int v) {
  return (int)(((odd ? kTetOdd : kTetEven) >> (3 * (t * 4 + v))) & 7);
}

// Cross-cell neighbour of non-centre tet `loc` (1..4) along axis a: the cell
// step s (-1/+1) and the neighbour's local id (:652-708).  12 bits per loc:
// 3 sign bits (1 = +1) then 3 local ids of 3 bits.
constexpr uint64_t nb_pack(int sx, int sy, int sz, int lx, int ly, int lz) {
  return (uint64_t)((sx > 0) | ((sy > 0) << 1) | ((sz > 0) << 2) | (lx << 3) | (ly << 6) |
                    (lz << 9));
}
constexpr uint64_t kNbEven = nb_pack(-1, 1, 1, 1, 2, 3) | (nb_pack(1, -1, 1, 2, 1, 4) << 12) |
                             (nb_pack(-1, -1, -1, 4, 3, 2) << 24) |
                             (nb_pack(1, 1, -1, 3, 4, 1) << 36);
constexpr uint64_t kNbOdd = nb_pack(1, 1, 1, 1, 2, 4) | (nb_pack(-1, -1, 1, 2, 1, 3) << 12) |
                            (nb_pack(-1, 1, -1, 4, 3, 1) << 24) |
                            (nb_pack(1, -1, -1, 3, 4, 2) << 36);
LSAMD_HD int nb_step(int odd, int loc, int a) {
  return (((odd ? kNbOdd : kNbEven) >> (12 * (loc - 1) + a)) & 1) ? 1 : -1;
}
LSAMD_HD int nb_local(int odd, int loc, int a) {
  return (int)(((odd ? kNbOdd : kNbEven) >> (12 * (loc - 1) + 3 + 3 * a)) & 7);
}

LSAMD_HD int64_t p1_node(const P1Mesh& m, int i, int j, int k) {
  return (int64_t)i + (int64_t)m.n[0] * j + (int64_t)m.n[0] * m.n[1] * k;
}
LSAMD_HD int64_t p1_cell(const P1Mesh& m, int i, int j, int k) {
  return (int64_t)i + (int64_t)m.c[0] * j + (int64_t)m.c[0] * m.c[1] * k;
}
LSAMD_HD void p1_cell_ijk(const P1Mesh& m, int64_t cell, int* ijk) {
  ijk[0] = (int)(cell % m.c[0]);
  ijk[1] = (int)(cell / m.c[0] % m.c[1]);
  ijk[2] = (int)(cell / m.c[0] / m.c[1]);
}

// node (i,j,k) of vertex v of tet `loc` in the cell at ijk
LSAMD_HD void tet_vertex(const int* cell, int loc, int v, int* out) {
  const int odd = (cell[0] + cell[1] + cell[2]) & 1;
  const int cr = tet_corner(odd, loc, v);
  out[0] = cell[0] + (cr & 1);
  out[1] = cell[1] + ((cr >> 1) & 1);
  out[2] = cell[2] + ((cr >> 2) & 1);
}

LSAMD_HD void tet_nodes(const P1Mesh& m, int64_t tet, int64_t* nodes) {
  int cell[3], v[3];
  p1_cell_ijk(m, tet / 5, cell);
  const int loc = (int)(tet % 5);
  for (int q = 0; q < 4; ++q) {
    tet_vertex(cell, loc, q, v);
    nodes[q] = p1_node(m, v[0], v[1], v[2]);
  }
}

// node_indexes_to_position: origin + step * index
LSAMD_HD double node_pos(const P1Mesh& m, int a, int idx) {
  return m.o[a] + m.h[a] * (double)idx;
}
// np.linspace(origin, origin + ncells*step, nnodes)[idx] (the support's nodes)
LSAMD_HD double lin_pos(const P1Mesh& m, int a, int idx) {
  const double stop = m.o[a] + (double)m.c[a] * m.h[a];
  if (idx == m.n[a] - 1) return stop;
  const double step = (stop - m.o[a]) / (double)(m.n[a] - 1);
  return (double)idx * step + m.o[a];
}

// inv(M) @ [[-1,1,0,0],[-1,0,1,0],[-1,0,0,1]] for M with rows v1-v0, v2-v0,
// v3-v0 (get_element_gradients, :421-486): G[d][k] = d lambda_k / d x_d.
// The inverse is the adjugate over the determinant.
LSAMD_HD void inv_grad(const double M[3][3], double G[3][4]) {
  const double c00 = M[1][1] * M[2][2] - M[1][2] * M[2][1];
  const double c01 = M[1][2] * M[2][0] - M[1][0] * M[2][2];
  const double c02 = M[1][0] * M[2][1] - M[1][1] * M[2][0];
  const double det = M[0][0] * c00 + M[0][1] * c01 + M[0][2] * c02;
  double inv[3][3];
  inv[0][0] = c00 / det;
  inv[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) / det;
  inv[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) / det;
  inv[1][0] = c01 / det;
  inv[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) / det;
  inv[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) / det;
  inv[2][0] = c02 / det;
  inv[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) / det;
  inv[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) / det;
  for (int d = 0; d < 3; ++d) {
    G[d][0] = ((-inv[d][0]) - inv[d][1]) - inv[d][2];
    G[d][1] = inv[d][0];
    G[d][2] = inv[d][1];
    G[d][3] = inv[d][2];
  }
}

// vertices (node_indexes_to_position) of a tet given its cell and local id
LSAMD_HD void tet_vertices(const P1Mesh& m, const int* cell, int loc, double V[4][3]) {
  int v[3];
  for (int q = 0; q < 4; ++q) {
    tet_vertex(cell, loc, q, v);
    for (int a = 0; a < 3; ++a) V[q][a] = node_pos(m, a, v[a]);
  }
}

LSAMD_HD void grad_from_vertices(const double V[4][3], double G[3][4]) {
  double M[3][3];
  for (int r = 0; r < 3; ++r)
    for (int a = 0; a < 3; ++a) M[r][a] = V[r + 1][a] - V[0][a];
  inv_grad(M, G);
}

LSAMD_HD void tet_grad(const P1Mesh& m, int64_t tet, double G[3][4]) {
  int cell[3];
  p1_cell_ijk(m, tet / 5, cell);
  double V[4][3];
  tet_vertices(m, cell, (int)(tet % 5), V);
  grad_from_vertices(V, G);
}

// |det(v1-v0, v2-v0, v3-v0)| / 6 from the support's (linspace) nodes (:80-94)
LSAMD_HD double tet_volume(const P1Mesh& m, int64_t tet) {
  int cell[3], v[3];
  p1_cell_ijk(m, tet / 5, cell);
  double V[4][3];
  for (int q = 0; q < 4; ++q) {
    tet_vertex(cell, (int)(tet % 5), q, v);
    for (int a = 0; a < 3; ++a) V[q][a] = lin_pos(m, a, v[a]);
  }
  double M[3][3];
  for (int r = 0; r < 3; ++r)
    for (int a = 0; a < 3; ++a) M[r][a] = V[r + 1][a] - V[0][a];
  const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) +
                     M[0][1] * (M[1][2] * M[2][0] - M[1][0] * M[2][2]) +
                     M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
  return fabs(det) / 6;
}

// ------------------------------------------------------ face pairs ----
// Number of face-sharing tet pairs: 4 inside every cell plus 2 per interior
// cell face.
inline int64_t p1_n_pairs(const P1Mesh& m) {
  const int64_t cx = m.c[0], cy = m.c[1], cz = m.c[2];
  return 4 * cx * cy * cz +
         2 * ((cx - 1) * cy * cz + cx * (cy - 1) * cz + cx * cy * (cz - 1));
}

// Index of the first pair whose larger tet lies in cell (i,j,k).  Pairs are
// ordered like tril(el_pairs).nonzero() (_init_face_table, :120-184): by the
// larger tet id, then the smaller.  A cell owns 4 + 2 per existing lower
// (-x, -y, -z) neighbour cell.
LSAMD_HD int64_t p1_pair_base(const P1Mesh& m, int i, int j, int k) {
  const int64_t cx = m.c[0], cy = m.c[1];
  const int64_t cell = p1_cell(m, i, j, k);
  const int64_t row = (int64_t)j + cy * k;
  const int64_t nx = row * (cx - 1) + (i > 0 ? i - 1 : 0);
  const int64_t ny = (int64_t)k * cx * (cy - 1) + (j > 0 ? cx * (j - 1) + i : 0);
  const int64_t nz = k > 0 ? cx * cy * (k - 1) + cx * (int64_t)j + i : 0;
  return 4 * cell + 2 * (nx + ny + nz);
}

// Enumerate the pairs owned by a cell in reference order; f(t_big, t_small).
template <typename F>
LSAMD_HD void p1_cell_pairs(const P1Mesh& m, const int* cell, F&& f) {
  const int odd = (cell[0] + cell[1] + cell[2]) & 1;
  const int64_t cid = p1_cell(m, cell[0], cell[1], cell[2]);
  const int64_t stride[3] = {1, (int64_t)m.c[0], (int64_t)m.c[0] * m.c[1]};
  for (int loc = 1; loc <= 4; ++loc) {
    const int64_t t = 5 * cid + loc;
    for (int a = 2; a >= 0; --a) {  // lower cells: -z < -y < -x in id
      if (nb_step(odd, loc, a) < 0 && cell[a] > 0)
        f(t, 5 * (cid - stride[a]) + nb_local(odd, loc, a));
    }
    f(t, 5 * cid);
  }
}

// The unnormalised edge-jump row of a pair (minimise_edge_jumps,
// LI/_p1interpolator.py:115-166): [n.grad(T1), -n.grad(T2)] over
// [nodes(T1), nodes(T2)], n = (v1-v0) x (v2-v0) of the shared face with its
// nodes in ascending id order, on the support's nodes.  Returns the 3 shared
// node ids in face[] when non-NULL.
LSAMD_HD void p1_edge_row(const P1Mesh& m, int64_t t1, int64_t t2, int64_t* col, double* a,
                          int64_t* face) {
  tet_nodes(m, t1, col);
  tet_nodes(m, t2, col + 4);
  int64_t f[3];
  int nf = 0;
  for (int q = 0; q < 4; ++q) {
    bool shared = false;
    for (int r = 0; r < 4; ++r) shared |= col[q] == col[4 + r];
    if (shared && nf < 3) f[nf++] = col[q];
  }
  // ascending
  if (f[0] > f[1]) { int64_t s = f[0]; f[0] = f[1]; f[1] = s; }
  if (f[1] > f[2]) { int64_t s = f[1]; f[1] = f[2]; f[2] = s; }
  if (f[0] > f[1]) { int64_t s = f[0]; f[0] = f[1]; f[1] = s; }
  double P[3][3];
  const int64_t nx = m.n[0], nxy = (int64_t)m.n[0] * m.n[1];
  for (int q = 0; q < 3; ++q) {
    const int k = (int)(f[q] / nxy);
    const int64_t r = f[q] - (int64_t)k * nxy;
    const int j = (int)(r / nx);
    const int i = (int)(r - (int64_t)j * nx);
    P[q][0] = lin_pos(m, 0, i);
    P[q][1] = lin_pos(m, 1, j);
    P[q][2] = lin_pos(m, 2, k);
  }
  double u[3], v[3], nv[3];
  for (int d = 0; d < 3; ++d) {
    u[d] = P[1][d] - P[0][d];
    v[d] = P[2][d] - P[0][d];
  }
  nv[0] = u[1] * v[2] - u[2] * v[1];
  nv[1] = u[2] * v[0] - u[0] * v[2];
  nv[2] = u[0] * v[1] - u[1] * v[0];
  double G1[3][4], G2[3][4];
  tet_grad(m, t1, G1);
  tet_grad(m, t2, G2);
  for (int k = 0; k < 4; ++k) {
    a[k] = (nv[0] * G1[0][k] + nv[1] * G1[1][k]) + nv[2] * G1[2][k];
    a[4 + k] = -((nv[0] * G2[0][k] + nv[1] * G2[1][k]) + nv[2] * G2[2][k]);
  }
  if (face) {
    face[0] = f[0];
    face[1] = f[1];
    face[2] = f[2];
  }
}

// L2 norm of 8 entries in numpy's pairwise order
LSAMD_HD double p1_norm8(const double* a) {
  double s[8];
  for (int q = 0; q < 8; ++q) s[q] = a[q] * a[q];
  return sqrt(((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7])));
}

// ---------------------------------------------------- operator table ----
// The edge-jump normal operator couples a node with 21 others: itself,
// +-1 and +-2 along each axis, and the 8 body diagonals.
constexpr int kP1Taps = 21;
LSAMD_HD void p1_tap_offset(int q, int* d) {
  d[0] = d[1] = d[2] = 0;
  if (q == 0) return;
  if (q <= 12) {
    const int a = ((q - 1) % 6) / 2;
    const int s = ((q - 1) & 1) ? -1 : 1;
    d[a] = q <= 6 ? s : 2 * s;
    return;
  }
  const int b = q - 13;
  d[0] = (b & 1) ? 1 : -1;
  d[1] = (b & 2) ? 1 : -1;
  d[2] = (b & 4) ? 1 : -1;
}

// Virtual node index of a real node for the operator table: the operator
// row of a node depends only on its distance (0, 1 or >= 2) to each face and
// on the parity of its index, so axes longer than 7 nodes map onto a 6- or
// 7-node virtual axis of the same parity.
LSAMD_HD int p1_virtual_n(int n) { return n <= 7 ? n : (n % 2 == 0 ? 6 : 7); }
LSAMD_HD int p1_virtual(int i, int n) {
  if (n <= 7 || i < 2) return i;
  const int nv = p1_virtual_n(n);
  if (i >= n - 2) return i - (n - nv);
  return 2 + (i & 1);
}

// Host: coefficient table tab[((vz*nvy + vy)*nvx + vx)*21 + q] of
// w2 * sum over edge rows of a a^T for the virtual mesh of g (see
// p1_virtual), every row normalised by its 8-entry norm.
void build_p1_table(const lsamd_grid& g, double w2, double* tab, int* nv);

}  // namespace lsamd
