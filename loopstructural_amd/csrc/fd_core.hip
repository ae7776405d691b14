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
#include <chrono>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "fd_operator.hpp"
#include "p1_mesh.hpp"

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
  const int given = (g->axis_nodes[0] != nullptr) + (g->axis_nodes[1] != nullptr) +
                    (g->axis_nodes[2] != nullptr);
  if (given != 0 && given != 3)
    return fail(LSAMD_ERR_ARG, "a rectilinear grid needs node arrays on all three axes");
  for (int a = 0; a < 3; ++a) {
    if (g->nnodes[a] < 2) return fail(LSAMD_ERR_ARG, "grid needs >= 2 nodes per axis");
    if (given == 0 && !(g->step[a] != 0.0))
      return fail(LSAMD_ERR_ARG, "grid step must be non-zero");
  }
  if (g->nnodes[0] * g->nnodes[1] * g->nnodes[2] > (int64_t)1 << 40)
    return fail(LSAMD_ERR_ARG, "grid too large");
  return LSAMD_OK;
}

int check_uniform_grid(const lsamd_grid* g, const char* fn) {
  if (int rc = check_grid(g)) return rc;
  if (g->axis_nodes[0])
    return fail(LSAMD_ERR_ARG, std::string(fn) + ": needs a uniform grid (axis_nodes NULL)");
  return LSAMD_OK;
}

// ------------------------------------------------------- CG control block --
struct CgCtl {
  double rho, rho_prev, pq_st, pq_rows, thr, bnorm, acc, alpha, res2;
  int iter, done, maxiter, converged;
};

// ------------------------------------------------------------- stencil ----
// y = (R^T W R + ridge2 I) x with per-block partials of x.y.
//
// Regular grids (>= 7 nodes per axis, uniform row weights) use one launch
// (k_b_tiles, in batch_kernels.hpp) over a table of work items, each
// plane-streaming a 64x8 xy tile through a double-buffered LDS plane with D
// planes of global loads kept in flight in registers:
//   * bulk z-chunks [3 + ZC*b, ...) of tiles interior in x and y: one
//     coefficient set, passed as kernel arguments (st_bulk_body<UNI=true>);
//   * bulk z-chunks of edge tiles: per-lane coefficients of the node's (x, y)
//     class, loaded once before the first prefetch (st_bulk_body<UNI=false>);
//   * output planes [0,3) and [nz-3,nz): coefficients re-read per plane
//     (st_zb_body).
// Other grids (an axis with < 7 nodes, or a per-node weight scale) use
// k_st_general: one node per thread, exact two-pass formula.
constexpr int TX = 64, TY = 8, NT = TX * TY;
constexpr int PX = TX + 4, PY = TY + 4, PSZ = PX * PY;
// Work-item tiles are TXT x (NT / TXT): 64 x 8 (TX, TY above) or 32 x 16.
// Bodies templated on TXT shadow TX, TY, PX, PY and PSZ with their own.
#define LSAMD_TILE_CONSTANTS(TXT)                                           \
  constexpr int TX = (TXT), TY = NT / (TXT), PX = TX + 4, PY = TY + 4;      \
  constexpr int PSZ = PX * PY;                                              \
  static_assert(TX == 64 || TX == 32, "64 x 8 or 32 x 16 tiles");           \
  (void)PY

template <typename T>
struct Coef25 {
  T k[kTaps];
};

// 25-tap update of one output node from its plane sums; see the tap order
// in fd_operator.hpp.  b points at the node in the current LDS plane.
template <typename T, int PX = ::lsamd::PX>
__device__ __forceinline__ T inplane(const T* k, const T* b) {
  return k[0] * b[0] + k[1] * b[1] + k[2] * b[-1] + k[3] * b[PX] + k[4] * b[-PX] + k[5] * b[2] +
         k[6] * b[-2] + k[7] * b[2 * PX] + k[8] * b[-2 * PX] + k[9] * b[2 * PX + 2] +
         k[10] * b[2 * PX - 2] + k[11] * b[-2 * PX + 2] + k[12] * b[-2 * PX - 2];
}

// Output planes [3 + bz*ZC, min(3 + (bz+1)*ZC, nz-3)) of the 64x8 xy tile tb;
// returns this thread's share of x.y.
template <typename T, int ZC, int D, bool UNI, int TXT = 64>
__device__ __forceinline__ double st_bulk_body(const T* __restrict__ x, T* __restrict__ y,
                                               const GridDev& g, const T* __restrict__ tab,
                                               const Coef25<T>& ku, int2 tb, int bz) {
  LSAMD_TILE_CONSTANTS(TXT);
  constexpr int NS = ZC + 4;
  __shared__ T pl[2][PSZ];
  const int x0 = tb.x * TX, y0 = tb.y * TY, z0 = 3 + bz * ZC;
  const int zend = min(z0 + ZC, g.nz - 3);
  const int lx = threadIdx.x & (TX - 1), ly = threadIdx.x / TX;
  const int ix = x0 + lx, iy = y0 + ly;
  const bool colok = ix < g.nx && iy < g.ny;
  const int64_t nxy = (int64_t)g.nx * g.ny;
  T k[kTaps];
  if constexpr (UNI) {
#pragma unroll
    for (int q = 0; q < kTaps; ++q) k[q] = ku.k[q];
  } else {
    const int cx = cls7(min(ix, g.nx - 1), g.nx), cy = cls7(min(iy, g.ny - 1), g.ny);
    const T* tc = tab + (int64_t)((3 * 7 + cy) * 7 + cx) * kTaps;
#pragma unroll
    for (int q = 0; q < kTaps; ++q) k[q] = tc[q];
  }
  // two load slots per thread in the (64+4) x (8+4) plane tile; addresses
  // clamped into the grid, out-of-grid values zeroed after the load
  const int e1 = threadIdx.x + NT;
  const bool has1 = e1 < PSZ;
  int64_t go0, go1;
  bool ok0, ok1;
  {
    const int ey = threadIdx.x / PX, ex = threadIdx.x - ey * PX;
    const int gx = x0 - 2 + ex, gy = y0 - 2 + ey;
    ok0 = gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny;
    go0 = (int64_t)min(max(gx, 0), g.nx - 1) + (int64_t)min(max(gy, 0), g.ny - 1) * g.nx;
  }
  {
    const int ee = has1 ? e1 : 0;
    const int ey = ee / PX, ex = ee - ey * PX;
    const int gx = x0 - 2 + ex, gy = y0 - 2 + ey;
    ok1 = has1 && gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny;
    go1 = (int64_t)min(max(gx, 0), g.nx - 1) + (int64_t)min(max(gy, 0), g.ny - 1) * g.nx;
  }
  const int pbeg = z0 - 2;  // >= 1
  T qa[D], qb[D];
  auto ld = [&](int p, T& va, T& vb) {
    const int64_t po = (int64_t)min(p, g.nz - 1) * nxy;
    const T a = x[go0 + po], b = x[go1 + po];
    va = ok0 ? a : (T)0;
    vb = ok1 ? b : (T)0;
  };
#pragma unroll
  for (int d = 0; d < D; ++d) ld(pbeg + d, qa[d], qb[d]);
  // ring of the four previous planes: c, x(+2,0), x(-2,0), x(0,+2), x(0,-2)
  T r4[5] = {0, 0, 0, 0, 0}, r3[5] = {0, 0, 0, 0, 0}, r2[5] = {0, 0, 0, 0, 0},
    r1[5] = {0, 0, 0, 0, 0};
  T v2 = 0, v1 = 0;
  double dot = 0.0;
  pl[0][threadIdx.x] = qa[0];
  if (has1) pl[0][e1] = qb[0];
  __syncthreads();
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    if (s + D < NS) ld(pbeg + s + D, qa[s % D], qb[s % D]);
    const T* b = pl[s & 1] + (ly + 2) * PX + lx + 2;
    const T c = b[0], xp2 = b[2], xm2 = b[-2], yp2 = b[2 * PX], ym2 = b[-2 * PX];
    const T v0 = inplane<T, PX>(k, b);
    if (s >= 4) {
      const int z = pbeg + s - 2;
      const T yv = v2 + k[13] * r3[0] + k[14] * r1[0] + k[15] * r4[0] + k[16] * r4[1] +
                   k[17] * r4[2] + k[18] * r4[3] + k[19] * r4[4] + k[20] * c + k[21] * xp2 +
                   k[22] * xm2 + k[23] * yp2 + k[24] * ym2;
      if (colok && z < zend) {
        // streaming store: y is not read again by this kernel
        __builtin_nontemporal_store(yv, &y[(int64_t)ix + (int64_t)iy * g.nx + (int64_t)z * nxy]);
        dot += (double)r2[0] * (double)yv;
      }
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      r4[q] = r3[q];
      r3[q] = r2[q];
      r2[q] = r1[q];
    }
    r1[0] = c;
    r1[1] = xp2;
    r1[2] = xm2;
    r1[3] = yp2;
    r1[4] = ym2;
    v2 = v1;
    v1 = v0;
    if (s + 1 < NS) {
      pl[(s + 1) & 1][threadIdx.x] = qa[(s + 1) % D];
      if (has1) pl[(s + 1) & 1][e1] = qb[(s + 1) % D];
    }
    __syncthreads();
  }
  return dot;
}

// ---- planes staged by direct global -> LDS loads (LDS DMA) ----
//
// The same work item and arithmetic as st_bulk_body, with the (64+4) x (8+4)
// plane tiles copied by global_load_lds_dword straight into a ring of DD LDS
// buffers: no VGPRs hold planes in flight, so DD - 1 planes per workgroup are
// in flight at once.  Each wave copies SL slots of 64 dwords per plane (a
// double takes two dwords), slots past the tile re-read its first dword.
// Out-of-grid halo positions read the clamped in-grid node instead of zero:
// the class coefficients of every tap leaving the grid are exactly zero, so
// each output is bit-identical to st_bulk_body's.  Outputs are written by
// buffer stores that every wave issues once per output plane (lanes without
// an output get an offset past the buffer, which the hardware drops), so
// the count of memory operations issued after each plane copy is known at
// compile time and vmcnt(N) proves that the plane has landed (vector memory
// operations complete in issue order).  The ring buffers are distinct LDS
// arrays so the compiler's own wait insertion can tell the copy into one
// buffer from reads of another.
#define LSAMD_LDS_PTR(p) ((__attribute__((address_space(3))) void*)(p))

template <typename T>
constexpr int kDmaWords = PSZ * (int)sizeof(T) / 4;  // dwords of one plane tile
template <typename T>
constexpr int kDmaSlots = (kDmaWords<T> + NT - 1) / NT;  // 64-dword slots per wave
template <typename T>
constexpr int kDmaBuf = kDmaSlots<T> * NT;  // dwords of one ring buffer
template <typename T>
constexpr int kDmaDepth = sizeof(T) == 4 ? 7 : 6;  // ring buffers (LDS: 28 KB / 48 KB)

// buffer resource of one plane (bytes in num_records; raw, no stride)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(base, 0, (int)bytes, 0x00020000);
}

// non-temporal buffer store of one element (dropped past num_records)
template <typename T>
__device__ __forceinline__ void buffer_store_nt(T v, __amdgpu_buffer_rsrc_t r, unsigned off) {
  if constexpr (sizeof(T) == 4)
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)off, 0, 2);
  else
    __builtin_amdgcn_raw_buffer_store_b64(
        __builtin_bit_cast(__attribute__((__vector_size__(8))) unsigned, v), r, (int)off, 0, 2);
}

// s_waitcnt vmcnt(N) lgkmcnt(0) (expcnt untouched)
template <int N>
__device__ __forceinline__ void wait_vm_lgkm() {
  static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4));
}

__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// ring buffer I of element type T: one LDS array per (T, I) in a kernel,
// shared by every body of the kernel that uses it
template <typename T, int I>
__device__ __forceinline__ unsigned* dma_ring() {
  __shared__ __align__(16) unsigned buf[kDmaBuf<T>];
  return buf;
}

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

template <typename T, int ZC, bool UNI>
__device__ __forceinline__ double st_bulk_dma(const T* __restrict__ x, T* __restrict__ y,
                                              const GridDev& g, const T* __restrict__ tab,
                                              const Coef25<T>& ku, int2 tb, int bz) {
  constexpr int NS = ZC + 4;
  constexpr int DD = kDmaDepth<T>;
  constexpr int SL = kDmaSlots<T>;
  constexpr int WPE = (int)sizeof(T) / 4;  // dwords per element
  static_assert(DD >= 3 && DD <= 8, "ring depth");
  auto ring = [](auto I) -> unsigned* { return dma_ring<T, decltype(I)::value % DD>(); };
  const int x0 = tb.x * TX, y0 = tb.y * TY, z0 = 3 + bz * ZC;
  const int zend = min(z0 + ZC, g.nz - 3);
  const int lx = threadIdx.x & (TX - 1), ly = threadIdx.x / TX;
  const int ix = x0 + lx, iy = y0 + ly;
  const bool colok = ix < g.nx && iy < g.ny;
  const int64_t nxy = (int64_t)g.nx * g.ny;
  T k[kTaps];
  if constexpr (UNI) {
#pragma unroll
    for (int q = 0; q < kTaps; ++q) k[q] = ku.k[q];
  } else {
    const int cx = cls7(min(ix, g.nx - 1), g.nx), cy = cls7(min(iy, g.ny - 1), g.ny);
    const T* tc = tab + (int64_t)((3 * 7 + cy) * 7 + cx) * kTaps;
#pragma unroll
    for (int q = 0; q < kTaps; ++q) k[q] = tc[q];
  }
  // copy slots of this lane: byte offset inside a plane of the tile (a plane
  // is at most 4 GiB), added to the plane's wave-uniform base address
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ln = threadIdx.x & 63;
  unsigned goff[SL];
#pragma unroll
  for (int j = 0; j < SL; ++j) {
    const int e = (j * (NT / 64) + wv) * 64 + ln;
    const int ee = e < kDmaWords<T> ? e : 0;
    const int el = ee / WPE, half = ee - el * WPE;
    const int ey = el / PX, ex = el - ey * PX;
    const int gx = min(max(x0 - 2 + ex, 0), g.nx - 1), gy = min(max(y0 - 2 + ey, 0), g.ny - 1);
    goff[j] = (unsigned)((((int64_t)gx + (int64_t)gy * g.nx) * WPE + half) * 4);
  }
  const char* xb = (const char*)x;
  const int pbeg = z0 - 2;  // >= 1
  auto copy_plane = [&](auto I) {
    constexpr int q = decltype(I)::value;
    unsigned* buf = ring(I);
    const char* plane = xb + (int64_t)min(pbeg + q, g.nz - 1) * nxy * (int64_t)sizeof(T);
#pragma unroll
    for (int j = 0; j < SL; ++j)
      __builtin_amdgcn_global_load_lds((const void*)(plane + goff[j]),
                                       LSAMD_LDS_PTR(buf + (j * (NT / 64) + wv) * 64), 4, 0, 0);
  };
  // the coefficient loads complete before the copies start
  wait_vm_lgkm<0>();
  asm volatile("" ::: "memory");
  static_for<0, (DD - 1 < NS ? DD - 1 : NS)>(copy_plane);
  T r4[5] = {0, 0, 0, 0, 0}, r3[5] = {0, 0, 0, 0, 0}, r2[5] = {0, 0, 0, 0, 0},
    r1[5] = {0, 0, 0, 0, 0};
  T v2 = 0, v1 = 0;
  double dot = 0.0;
  const unsigned pbytes = (unsigned)(nxy * (int64_t)sizeof(T));
  // lanes without an output store at kDropOffset: past any plane (planes are
  // below 2 GiB) and far from wrapping, so the hardware drops the store
  constexpr unsigned kDropOffset = 0x80000000u;
  const unsigned yoff = colok ? (unsigned)(((int64_t)ix + (int64_t)iy * g.nx) * sizeof(T))
                              : kDropOffset;
  static_for<0, NS>([&](auto I) {
    constexpr int s = decltype(I)::value;
    // Plane s has landed once at most the operations issued after its copy
    // are outstanding: the copies of the later planes issued so far and the
    // output stores of the iterations since (one per iteration from s = 4).
    // The barrier also ends every wave's reads of plane s - 1, whose buffer
    // the next copy overwrites.
    constexpr int later = (DD - 2 < NS - 1 - s ? DD - 2 : NS - 1 - s);
    constexpr int t0 = s - DD + 1 > 0 ? s - DD + 1 : 0;  // iteration that copied plane s
    constexpr int st0 = t0 > 4 ? t0 : 4;
    constexpr int stores = s > st0 ? s - st0 : 0;
    wait_vm_lgkm<SL * later + stores>();
    lds_barrier();
    if constexpr (s + DD - 1 < NS) copy_plane(std::integral_constant<int, s + DD - 1>{});
    const T* b = (const T*)ring(I) + (ly + 2) * PX + lx + 2;
    const T c = b[0], xp2 = b[2], xm2 = b[-2], yp2 = b[2 * PX], ym2 = b[-2 * PX];
    const T v0 = inplane(k, b);
    if constexpr (s >= 4) {
      const int z = pbeg + s - 2;
      const T yv = v2 + k[13] * r3[0] + k[14] * r1[0] + k[15] * r4[0] + k[16] * r4[1] +
                   k[17] * r4[2] + k[18] * r4[3] + k[19] * r4[4] + k[20] * c + k[21] * xp2 +
                   k[22] * xm2 + k[23] * yp2 + k[24] * ym2;
      // branch-free, so that every wave issues exactly one store here
      const bool out = colok && z < zend;
      buffer_store_nt<T>(yv, plane_rsrc(y + (int64_t)z * nxy, pbytes), out ? yoff : kDropOffset);
      dot += (double)(out ? r2[0] : (T)0) * (double)yv;
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      r4[q] = r3[q];
      r3[q] = r2[q];
      r2[q] = r1[q];
    }
    r1[0] = c;
    r1[1] = xp2;
    r1[2] = xm2;
    r1[3] = yp2;
    r1[4] = ym2;
    v2 = v1;
    v1 = v0;
  });
  return dot;
}

// Output planes [0,3) (bz = 0) or [nz-3, nz) (bz = 1) of xy tile (tx, ty);
// coefficients re-read per plane.  Returns this thread's share of x.y.
template <typename T, bool RING = false, int TXT = 64>
__device__ __forceinline__ double st_zb_body(const T* __restrict__ x, T* __restrict__ y,
                                             const GridDev& g, const T* __restrict__ tab, int tx,
                                             int ty, int bz) {
  LSAMD_TILE_CONSTANTS(TXT);
  // RING: the two plane buffers are the first two LDS-DMA ring buffers of
  // the kernel (a block runs one body, so they are free here)
  __shared__ T own[RING ? 1 : 2][RING ? 1 : PSZ];
  T* pl[2] = {RING ? (T*)dma_ring<T, 0>() : own[0], RING ? (T*)dma_ring<T, 1>() : own[RING ? 0 : 1]};
  const int x0 = tx * TX, y0 = ty * TY, z0 = bz == 0 ? 0 : g.nz - 3;
  const int lx = threadIdx.x & (TX - 1), ly = threadIdx.x / TX;
  const int ix = x0 + lx, iy = y0 + ly;
  const bool colok = ix < g.nx && iy < g.ny;
  const int64_t nxy = (int64_t)g.nx * g.ny;
  const int e1 = threadIdx.x + NT;
  int64_t go0, go1;
  bool ok0, ok1;
  {
    const int ey = threadIdx.x / PX, ex = threadIdx.x - ey * PX;
    const int gx = x0 - 2 + ex, gy = y0 - 2 + ey;
    ok0 = gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny;
    go0 = ok0 ? (int64_t)gx + (int64_t)gy * g.nx : 0;
  }
  {
    const int ey = e1 / PX, ex = e1 - ey * PX;
    const int gx = x0 - 2 + ex, gy = y0 - 2 + ey;
    ok1 = e1 < PSZ && gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny;
    go1 = ok1 ? (int64_t)gx + (int64_t)gy * g.nx : 0;
  }
  const int cx = cls7(min(ix, g.nx - 1), g.nx), cy = cls7(min(iy, g.ny - 1), g.ny);
  const T* tc = tab + (int64_t)(cy * 7 + cx) * kTaps;
  T r4[5] = {0, 0, 0, 0, 0}, r3[5] = {0, 0, 0, 0, 0}, r2[5] = {0, 0, 0, 0, 0},
    r1[5] = {0, 0, 0, 0, 0};
  T v2 = 0, v1 = 0;
  double dot = 0.0;
  for (int s = 0; s < 7; ++s) {
    const int p = z0 - 2 + s;
    const bool pin = p >= 0 && p < g.nz;
    const int64_t po = (int64_t)(pin ? p : 0) * nxy;
    pl[s & 1][threadIdx.x] = (pin && ok0) ? x[go0 + po] : (T)0;
    if (e1 < PSZ) pl[s & 1][e1] = (pin && ok1) ? x[go1 + po] : (T)0;
    __syncthreads();
    const T* b = pl[s & 1] + (ly + 2) * PX + lx + 2;
    const T* t = tc + (int64_t)(pin ? cls7(p, g.nz) : 3) * 49 * kTaps;
    const T c = b[0], xp2 = b[2], xm2 = b[-2], yp2 = b[2 * PX], ym2 = b[-2 * PX];
    const T v0 = inplane<T, PX>(t, b);
    const int z = p - 2;
    if (s >= 4 && z >= 0 && z < g.nz) {
      const T* o = tc + (int64_t)cls7(z, g.nz) * 49 * kTaps;
      const T yv = v2 + o[13] * r3[0] + o[14] * r1[0] + o[15] * r4[0] + o[16] * r4[1] +
                   o[17] * r4[2] + o[18] * r4[3] + o[19] * r4[4] + o[20] * c + o[21] * xp2 +
                   o[22] * xm2 + o[23] * yp2 + o[24] * ym2;
      if (colok) {
        // streaming store: y is not read again by this kernel
        __builtin_nontemporal_store(yv, &y[(int64_t)ix + (int64_t)iy * g.nx + (int64_t)z * nxy]);
        dot += (double)r2[0] * (double)yv;
      }
    }
    for (int q = 0; q < 5; ++q) {
      r4[q] = r3[q];
      r3[q] = r2[q];
      r2[q] = r1[q];
    }
    r1[0] = c;
    r1[1] = xp2;
    r1[2] = xm2;
    r1[3] = yp2;
    r1[4] = ym2;
    v2 = v1;
    v1 = v0;
    __syncthreads();
  }
  return dot;
}

// any grid: one node per thread, exact two-pass formula
template <typename T>
__global__ __launch_bounds__(256) void k_st_general(const T* __restrict__ x, T* __restrict__ y,
                                                    GridDev g, RegW<T> w, const double* scale2,
                                                    double* __restrict__ part, const CgCtl* ctl) {
  if (ctl->done) return;
  __shared__ double shm[4];
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double dot = 0.0;
  if (id < n) {
    const int64_t nxy = (int64_t)g.nx * g.ny;
    const int k = (int)(id / nxy);
    const int64_t r = id - (int64_t)k * nxy;
    const int j = (int)(r / g.nx);
    const int i = (int)(r - (int64_t)j * g.nx);
    const T yv = general_node(x, g, w, scale2, i, j, k);
    y[id] = yv;
    dot = (double)x[id] * (double)yv;
  }
  const double bs = block_sum(dot, shm);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// P1 edge-jump regulariser: 21-tap table stencil, one node per thread.  The
// table row of a node is chosen by its virtual index per axis (p1_virtual);
// taps falling outside the grid have zero coefficients and are not read.
template <typename T>
__global__ __launch_bounds__(256) void k_p1_stencil(const T* __restrict__ x, T* __restrict__ y,
                                                    GridDev g, const T* __restrict__ tab, int nvx,
                                                    int nvy, T ridge2, double* __restrict__ part,
                                                    const CgCtl* ctl) {
  if (ctl->done) return;
  __shared__ double shm[4];
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double dot = 0.0;
  if (id < n) {
    const int64_t nxy = (int64_t)g.nx * g.ny;
    const int k = (int)(id / nxy);
    const int64_t r = id - (int64_t)k * nxy;
    const int j = (int)(r / g.nx);
    const int i = (int)(r - (int64_t)j * g.nx);
    const int vx = p1_virtual(i, g.nx), vy = p1_virtual(j, g.ny), vz = p1_virtual(k, g.nz);
    const T* c = tab + (((int64_t)vz * nvy + vy) * nvx + vx) * kP1Taps;
    const T xc = x[id];
    T acc = c[0] * xc;
#pragma unroll
    for (int q = 1; q < kP1Taps; ++q) {
      int d[3];
      p1_tap_offset(q, d);
      const bool ok = (unsigned)(i + d[0]) < (unsigned)g.nx && (unsigned)(j + d[1]) < (unsigned)g.ny &&
                      (unsigned)(k + d[2]) < (unsigned)g.nz;
      const int64_t o = ok ? id + d[0] + (int64_t)g.nx * d[1] + nxy * d[2] : id;
      const T xv = x[o];
      acc += c[q] * (ok ? xv : (T)0);
    }
    acc += ridge2 * xc;
    y[id] = acc;
    dot = (double)xc * (double)acc;
  }
  const double bs = block_sum(dot, shm);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// P1 edge-jump regulariser, plane-streaming.  A block covers a TXV x (512 /
// TXV) xy tile and ZC output planes; the planes from two below to two above
// the chunk stream through a double-buffered LDS tile with a 2-node xy halo
// (D planes of global loads in flight in registers, out-of-grid values
// zeroed).  Each plane, as it arrives, adds to the five output planes of the
// column that it touches: its in-plane taps (+-1, +-2 along x and y, and the
// centre) to the output in the same plane, its centre and four diagonal
// nodes to the outputs one plane above and below, its centre to those two
// planes away.  The output two planes below is then complete and stored.
//
// Coefficients.  A node's table row is fixed by its virtual (x, y) index and
// its virtual z index, which is 2 + (z & 1) for every z at least two nodes
// from a z face.  A z chunk whose outputs all lie there (ZIN) takes its rows
// from the table slab of virtual z 2 and 3, staged in LDS for every virtual
// (x, y): each thread holds the two rows of its column (even and odd z), in
// registers for FP32 and read from LDS for FP64, and the plane loop runs in
// even/odd pairs.  The first and last chunks read rows from the table in
// global memory (cache resident).  One x.y partial per block.
template <typename T>
__device__ __forceinline__ const T* p1_grow(const T* __restrict__ tab, int vx, int vy, int vz,
                                            int nvx, int nvy) {
  return tab + (((int64_t)vz * nvy + vy) * nvx + vx) * kP1Taps;
}

constexpr int kP1SlabMax = 2 * 7 * 7 * kP1Taps;  // rows of virtual z 2, 3

template <typename T, int TXV, int ZC, bool ZIN>
__device__ __forceinline__ double p1_body(const T* __restrict__ x, T* __restrict__ y,
                                          const GridDev& g, const T* __restrict__ tab,
                                          const T* slab, T* plbuf, int nvx, int nvy, T ridge2,
                                          int x0, int y0, int z0, int z1) {
  constexpr int TYV = 512 / TXV, QX = TXV + 4, QY = TYV + 4, QS = QX * QY;
  constexpr int D = ZIN ? 4 : 2;
  constexpr bool CREG = ZIN && sizeof(T) == 4;  // FP32 rows in registers
  T* const pl[2] = {plbuf, plbuf + QS};
  const int lx = threadIdx.x % TXV, ly = threadIdx.x / TXV;
  const int i = x0 + lx, j = y0 + ly;
  const bool col = i < g.nx && j < g.ny;
  const int64_t nxy = (int64_t)g.nx * g.ny;
  // two load slots per thread in the (TXV+4) x (TYV+4) plane tile
  const int e1 = threadIdx.x + 512;
  const bool has1 = e1 < QS;
  int64_t go0, go1;
  bool ok0, ok1;
  {
    const int ey = threadIdx.x / QX, ex = threadIdx.x - ey * QX;
    const int gx = x0 - 2 + ex, gy = y0 - 2 + ey;
    ok0 = gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny;
    go0 = ok0 ? (int64_t)gx + (int64_t)gy * g.nx : 0;
  }
  {
    const int ee = has1 ? e1 : 0;
    const int ey = ee / QX, ex = ee - ey * QX;
    const int gx = x0 - 2 + ex, gy = y0 - 2 + ey;
    ok1 = has1 && gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny;
    go1 = ok1 ? (int64_t)gx + (int64_t)gy * g.nx : 0;
  }
  const int vx = p1_virtual(min(i, g.nx - 1), g.nx), vy = p1_virtual(min(j, g.ny - 1), g.ny);
  // ZIN: the rows of this column for even and odd z
  const T* rowE = slab + ((int64_t)(0 * nvy + vy) * nvx + vx) * kP1Taps;
  const T* rowO = slab + ((int64_t)(1 * nvy + vy) * nvx + vx) * kP1Taps;
  T ce[CREG ? kP1Taps : 1], co[CREG ? kP1Taps : 1];
  if constexpr (CREG) {
#pragma unroll
    for (int q = 0; q < kP1Taps; ++q) {
      ce[q] = rowE[q];
      co[q] = rowO[q];
    }
  }
  // planes pb .. pb + NS - 1; a ZIN chunk runs ZC + 4 planes (a short last
  // chunk reads zeroed planes past the grid and writes no output there)
  const int pb = z0 - 2, NS = ZIN ? ZC + 4 : z1 - z0 + 4;
  T qa[D], qb[D];
  auto ld = [&](int p, T& va, T& vb) {
    const bool pin = p >= 0 && p < g.nz;
    const int64_t po = (int64_t)(pin ? p : 0) * nxy;
    const T a = x[go0 + po], b = x[go1 + po];
    va = (pin && ok0) ? a : (T)0;
    vb = (pin && ok1) ? b : (T)0;
  };
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (d < NS) ld(pb + d, qa[d], qb[d]);
  // pending outputs z = p-2 .. p+2 and the centre values of planes p-2, p-1
  T a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0, h2 = 0, h1 = 0;
  double dot = 0.0;
  pl[0][threadIdx.x] = qa[0];
  if (has1) pl[0][e1] = qb[0];
  __syncthreads();
  // one plane; ODD: parity of p (ZIN only)
  auto plane = [&](int s, auto odd) {
    constexpr bool ODD = decltype(odd)::value;
    const int p = pb + s;
    if (s + D < NS) ld(p + D, qa[(s + D) % D], qb[(s + D) % D]);
    const T* bb = pl[s & 1] + (ly + 2) * QX + lx + 2;
    const T c = bb[0];
    const bool pin = p >= 0 && p < g.nz;
    if (col && pin) {
      // Rs: row of an output in this plane (or two away), Rn: one plane away
      auto Rs = [&](int q, int z) -> T {
        if constexpr (CREG) return ODD ? co[q] : ce[q];
        else if constexpr (ZIN) return (ODD ? rowO : rowE)[q];
        else return p1_grow(tab, vx, vy, p1_virtual(z, g.nz), nvx, nvy)[q];
      };
      auto Rn = [&](int q, int z) -> T {
        if constexpr (CREG) return ODD ? ce[q] : co[q];
        else if constexpr (ZIN) return (ODD ? rowE : rowO)[q];
        else return p1_grow(tab, vx, vy, p1_virtual(z, g.nz), nvx, nvy)[q];
      };
      if (p >= z0 && p < z1)
        a2 += Rs(0, p) * c + Rs(1, p) * bb[1] + Rs(2, p) * bb[-1] + Rs(3, p) * bb[QX] +
              Rs(4, p) * bb[-QX] + Rs(7, p) * bb[2] + Rs(8, p) * bb[-2] + Rs(9, p) * bb[2 * QX] +
              Rs(10, p) * bb[-2 * QX] + ridge2 * c;
      const T dmm = bb[-1 - QX], dpm = bb[1 - QX], dmp = bb[-1 + QX], dpp = bb[1 + QX];
      if (p - 1 >= z0 && p - 1 < z1)  // this plane is dz = +1 of output p-1
        a1 += Rn(5, p - 1) * c + Rn(17, p - 1) * dmm + Rn(18, p - 1) * dpm +
              Rn(19, p - 1) * dmp + Rn(20, p - 1) * dpp;
      if (p + 1 >= z0 && p + 1 < z1)  // dz = -1 of output p+1
        a3 += Rn(6, p + 1) * c + Rn(13, p + 1) * dmm + Rn(14, p + 1) * dpm +
              Rn(15, p + 1) * dmp + Rn(16, p + 1) * dpp;
      if (p - 2 >= z0 && p - 2 < z1) a0 += Rs(11, p - 2) * c;
      if (p + 2 >= z0 && p + 2 < z1) a4 += Rs(12, p + 2) * c;
    }
    const int zo = p - 2;  // complete now
    if (col && zo >= z0 && zo < z1) {
      __builtin_nontemporal_store(a0, &y[(int64_t)i + (int64_t)j * g.nx + (int64_t)zo * nxy]);
      dot += (double)h2 * (double)a0;
    }
    a0 = a1;
    a1 = a2;
    a2 = a3;
    a3 = a4;
    a4 = 0;
    h2 = h1;
    h1 = pin ? c : (T)0;
    if (s + 1 < NS) {
      pl[(s + 1) & 1][threadIdx.x] = qa[(s + 1) % D];
      if (has1) pl[(s + 1) & 1][e1] = qb[(s + 1) % D];
    }
    __syncthreads();
  };
  if constexpr (ZIN) {
    // z0 even so plane s has the parity of s; NS = ZC + 4 is even
#pragma unroll
    for (int s = 0; s < NS; s += 2) {
      plane(s, std::false_type{});
      plane(s + 1, std::true_type{});
    }
  } else {
    for (int s = 0; s < NS; ++s) plane(s, std::false_type{});
  }
  return dot;
}

// z chunk bz of a P1 grid: on a z axis of more than 7 nodes, [0, 2), then
// chunks of zc output planes over [2, nz - 2) (the last one shorter), then
// [nz - 2, nz); otherwise chunks of zc from 0.  Every chunk that is neither
// first nor last starts on an even plane.
LSAMD_HD int p1_nzchunks(int nz, int zc) {
  return nz > 7 ? 2 + (int)((nz - 4 + zc - 1) / zc) : (nz + zc - 1) / zc;
}
LSAMD_HD void p1_zchunk(int nz, int zc, int bz, int* z0, int* z1) {
  if (nz <= 7) {
    *z0 = bz * zc;
    *z1 = min(*z0 + zc, nz);
    return;
  }
  const int last = p1_nzchunks(nz, zc) - 1;
  if (bz == 0) {
    *z0 = 0;
    *z1 = 2;
  } else if (bz == last) {
    *z0 = nz - 2;
    *z1 = nz;
  } else {
    *z0 = 2 + (bz - 1) * zc;
    *z1 = min(*z0 + zc, nz - 2);
  }
}

template <typename T, int TXV, int ZC>
__global__ __launch_bounds__(512) void k_p1_tiles(const T* __restrict__ x, T* __restrict__ y,
                                                  GridDev g, const T* __restrict__ tab, int nvx,
                                                  int nvy, int nvz, T ridge2, int ntx, int nzc,
                                                  double* __restrict__ part, const CgCtl* ctl) {
  static_assert(ZC % 2 == 0, "even z chunks keep the plane parity");
  if (ctl->done) return;
  constexpr int TYV = 512 / TXV;
  __shared__ T slab[kP1SlabMax];
  __shared__ T plbuf[2 * (TXV + 4) * (TYV + 4)];
  __shared__ double shm[8];
  const int bz = (int)(blockIdx.x % (unsigned)nzc);
  const int tt = (int)(blockIdx.x / (unsigned)nzc);
  const int x0 = (tt % ntx) * TXV, y0 = (tt / ntx) * TYV;
  int z0, z1;
  p1_zchunk(g.nz, ZC, bz, &z0, &z1);
  // outputs at least two nodes from both z faces, on a z axis long enough
  // for the parity rule (virtual z 2 and 3 exist)
  const bool zin = nvz > 3 && g.nz > 7 && z0 >= 2 && z1 <= g.nz - 2;
  double dot;
  if (zin) {
    const int ns = 2 * nvy * nvx * kP1Taps;
    const T* src = tab + (int64_t)2 * nvy * nvx * kP1Taps;
    for (int e = threadIdx.x; e < ns; e += 512) slab[e] = src[e];
    __syncthreads();
    dot = p1_body<T, TXV, ZC, true>(x, y, g, tab, slab, plbuf, nvx, nvy, ridge2, x0, y0, z0,
                                    z1);
  } else {
    dot = p1_body<T, TXV, ZC, false>(x, y, g, tab, slab, plbuf, nvx, nvy, ridge2, x0, y0, z0,
                                     z1);
  }
  const double bsum = block_sum(dot, shm);
  if (threadIdx.x == 0) part[blockIdx.x] = bsum;
}

__global__ void k_p1_diag(GridDev g, const double* __restrict__ tab, int nvx, int nvy,
                          double ridge2, double* __restrict__ diag) {
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= n) return;
  const int64_t nxy = (int64_t)g.nx * g.ny;
  const int k = (int)(id / nxy);
  const int64_t r = id - (int64_t)k * nxy;
  const int j = (int)(r / g.nx);
  const int i = (int)(r - (int64_t)j * g.nx);
  const int vx = p1_virtual(i, g.nx), vy = p1_virtual(j, g.ny), vz = p1_virtual(k, g.nz);
  diag[id] = tab[(((int64_t)vz * nvy + vy) * nvx + vx) * kP1Taps] + ridge2;
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
  if (ctl->done) return;
  __shared__ double sh[4];
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
}

// y[node] += sum over the node's entries of coef_e t[row_e]  (node-major,
// fixed order: deterministic, no atomics).  accumulate=0 overwrites.
// With fin != 0, block 0 then reduces the stencil and row partials (written
// by the two previous kernels of the stream) into pq and alpha = rho / pq.
template <typename T, typename TO>
__global__ __launch_bounds__(256) void k_rows_scatter(RowsDev R, const T* __restrict__ t,
                                                      TO* __restrict__ y, int accumulate,
                                                      CgCtl* ctl, int fin,
                                                      const double* __restrict__ part_st,
                                                      int n_st,
                                                      const double* __restrict__ part_rows,
                                                      int n_rows) {
  if (ctl && ctl->done) return;
  const T* coef = (const T*)R.coef;
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d < R.ndn) {
    T acc = 0;
    for (int64_t s = R.seg_ptr[d]; s < R.seg_ptr[d + 1]; ++s)
      acc += coef[R.seg_entry[s]] * t[R.seg_row[s]];
    const int64_t node = R.dnode[d];
    if (accumulate)
      y[node] += (TO)acc;
    else
      y[node] = (TO)acc;
  }
  if (fin && blockIdx.x == 0) {
    __shared__ double sh[4];
    const double a = reduce_partials(part_st, n_st, sh);
    const double b = reduce_partials(part_rows, n_rows, sh);
    if (threadIdx.x == 0) {
      const double pq = a + b;
      ctl->pq_st = a;
      ctl->pq_rows = b;
      ctl->alpha = ctl->rho / pq;
    }
  }
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
__global__ void k_fill_one(double* a, float* b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    a[i] = 1.0;
    b[i] = 1.0f;
  }
}
void launch_eval_entries(const GridDev& g, const double* pts, const double* gv, int64_t n,
                         int32_t* key, int32_t* idx, double* val, hipStream_t st);

// one thread per run of equal keys in the sorted entries: out[key] = sum of
// the run's values in entry order
__global__ void k_eval_segsum(const int32_t* __restrict__ key, const int32_t* __restrict__ idx,
                              const double* __restrict__ val, int64_t ne,
                              double* __restrict__ out) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ne) return;
  const int32_t k = key[s];
  if (s > 0 && key[s - 1] == k) return;
  double acc = 0.0;
  for (int64_t t = s; t < ne && key[t] == k; ++t) acc += val[idx[t]];
  out[k] += acc;
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

// x = 0 ; r = rhs ; partials of r.(minv r)
template <typename T>
__global__ __launch_bounds__(VB) void k_cg_init(const double* __restrict__ rhs,
                                                const T* __restrict__ minv, T* __restrict__ x,
                                                T* __restrict__ r, int64_t n, double* part) {
  __shared__ double sh[VB / 64];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
    const T rv = (T)rhs[i];
    x[i] = (T)0;
    r[i] = rv;
    acc += (double)rv * (double)(minv[i] * rv);
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// one block: rho = sum(part); thr = max(atol, rtol sqrt(rho)); iter = 0
__global__ __launch_bounds__(VB) void k_cg_init_fin(const double* __restrict__ part, int np,
                                                    double rtol, double atol, int maxiter,
                                                    CgCtl* ctl) {
  __shared__ double sh[VB / 64];
  const double tot = reduce_partials(part, np, sh);
  if (threadIdx.x == 0) {
    const double bn = sqrt(tot);
    const double thr = fmax(atol, rtol * bn);
    ctl->rho = tot;
    ctl->rho_prev = 1.0;
    ctl->bnorm = bn;
    ctl->thr = thr;
    ctl->alpha = 0.0;
    ctl->iter = 0;
    ctl->maxiter = maxiter;
    ctl->converged = (sqrt(tot) < thr) || bn == 0.0;
    ctl->done = ctl->converged;
  }
}

// p = minv r + beta p  (beta = rho/rho_prev, p = minv r on the first step)
template <typename T>
__global__ __launch_bounds__(VB) void k_cg_dir(const T* __restrict__ r, const T* __restrict__ minv,
                                               T* __restrict__ p, int64_t n, CgCtl* ctl) {
  if (ctl->done) return;
  const int first = ctl->iter == 0;
  const T beta = first ? (T)0 : (T)(ctl->rho / ctl->rho_prev);
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
    const T z = minv[i] * r[i];
    p[i] = first ? z : z + beta * p[i];
  }
}

// x += alpha p ; r -= alpha q ; partials of r.(minv r)
template <typename T>
__global__ __launch_bounds__(VB) void k_cg_upd(T* __restrict__ x, T* __restrict__ r,
                                               const T* __restrict__ p, const T* __restrict__ q,
                                               const T* __restrict__ minv, int64_t n,
                                               double* part, CgCtl* ctl) {
  if (ctl->done) return;
  __shared__ double sh[VB / 64];
  const T alpha = (T)ctl->alpha;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
    x[i] += alpha * p[i];
    const T rv = r[i] - alpha * q[i];
    r[i] = rv;
    acc += (double)rv * (double)(minv[i] * rv);
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// one block: rho_prev = rho ; rho = sum(part) ; iter++ ; stopping test
__global__ __launch_bounds__(VB) void k_cg_upd_fin(const double* __restrict__ part, int np,
                                                   CgCtl* ctl) {
  if (ctl->done) return;
  __shared__ double sh[VB / 64];
  const double tot = reduce_partials(part, np, sh);
  if (threadIdx.x == 0) {
    ctl->rho_prev = ctl->rho;
    ctl->rho = tot;
    const int it = ctl->iter + 1;
    ctl->iter = it;
    const int conv = sqrt(tot) < ctl->thr;
    ctl->converged = conv;
    ctl->done = conv || it >= ctl->maxiter;
  }
}

// r64 = rhs - q64 (q64 = N c) ; partials of |r|_M^2
__global__ __launch_bounds__(VB) void k_residual(const double* __restrict__ rhs,
                                                 const double* __restrict__ q,
                                                 const double* __restrict__ minv,
                                                 double* __restrict__ r, int64_t n,
                                                 double* part) {
  __shared__ double sh[VB / 64];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * VB + threadIdx.x; i < n; i += (int64_t)gridDim.x * VB) {
    const double rv = rhs[i] - q[i];
    r[i] = rv;
    acc += rv * (minv[i] * rv);
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// one block: ctl->acc = sum(part)
__global__ __launch_bounds__(VB) void k_sum_fin(const double* __restrict__ part, int np,
                                                CgCtl* ctl) {
  __shared__ double sh[VB / 64];
  const double tot = reduce_partials(part, np, sh);
  if (threadIdx.x == 0) ctl->acc = tot;
}

// Squared distance from every grid node (np.linspace coordinates) to the
// nearest of m points; points staged through LDS in tiles.
__global__ __launch_bounds__(256) void k_nearest_d2(GridDev g, const double* __restrict__ pts,
                                                    int64_t m, double* __restrict__ d2) {
  __shared__ double sp[3][256];
  const int64_t n = (int64_t)g.nx * g.ny * g.nz;
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nxy = (int64_t)g.nx * g.ny;
  const int k = (int)(id / nxy);
  const int64_t r = id - (int64_t)k * nxy;
  const int j = (int)(r / g.nx);
  const int i = (int)(r - (int64_t)j * g.nx);
  // np.linspace(o, o + (n-1) h, n): x_i = i * ((max - o)/(n-1)) + o, last = max
  auto lin = [](int i, int nn, double o, double h) {
    const double mx = o + (double)(nn - 1) * h;
    if (i == nn - 1) return mx;
    const double st = (mx - o) / (double)(nn - 1);
    return (double)i * st + o;
  };
  const bool rect = g.xn != nullptr;
  const double x = rect ? g.xn[i < g.nx ? i : g.nx - 1] : lin(i, g.nx, g.ox, g.hx);
  const double y = rect ? g.yn[j < g.ny ? j : g.ny - 1] : lin(j, g.ny, g.oy, g.hy);
  const double z = rect ? g.zn[k < g.nz ? k : g.nz - 1] : lin(k, g.nz, g.oz, g.hz);
  double best = __builtin_inf();
  for (int64_t t0 = 0; t0 < m; t0 += 256) {
    __syncthreads();
    const int64_t q = t0 + threadIdx.x;
    if (q < m) {
      sp[0][threadIdx.x] = pts[3 * q];
      sp[1][threadIdx.x] = pts[3 * q + 1];
      sp[2][threadIdx.x] = pts[3 * q + 2];
    }
    __syncthreads();
    const int cnt = (int)min((int64_t)256, m - t0);
    for (int u = 0; u < cnt; ++u) {
      const double dx = x - sp[0][u], dy = y - sp[1][u], dz = z - sp[2][u];
      const double dd = (dx * dx + dy * dy) + dz * dz;
      best = dd < best ? dd : best;
    }
  }
  if (id < n) d2[id] = best;
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

// entry columns must address nodes of the system ([lo, hi)); *bad set otherwise
__global__ void k_check_cols(const int32_t* __restrict__ cols, int64_t nnz, int64_t lo, int64_t hi,
                             int* __restrict__ bad) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nnz && (cols[e] < lo || cols[e] >= hi)) *bad = 1;
}

#include "batch_kernels.hpp"
#include "lsmr.hpp"
#include "mg.hpp"
#include "mrhs.hpp"
#include "bmg.hpp"

// rows of system s may only address its own nodes; one workgroup per system
namespace lsamd {
__global__ void k_b_check_rows(RowsDev R, const BSys* __restrict__ sys, int* __restrict__ bad) {
  const BSys S = sys[blockIdx.x];
  for (int64_t r = S.r0 + threadIdx.x; r < S.r1; r += blockDim.x)
    for (int64_t e = R.row_ptr[r]; e < R.row_ptr[r + 1]; ++e)
      if (R.cols[e] < S.off || R.cols[e] >= S.off + S.n) *bad = 1;
}
}  // namespace lsamd

// ================================================================ host side ==
using namespace lsamd;

namespace lsamd {
struct Mg;
void mg_destroy(Mg* m);
void mg_invalidate(Mg* m);
struct Ls;
void ls_destroy(Ls* l);
struct MrWork;
void mr_destroy(MrWork* w);
void mr_invalidate(MrWork* w);
}  // namespace lsamd

struct lsamd_fd_system {
  lsamd::Mg* mg = nullptr;  // multigrid hierarchy (built on first MG solve after set_rows)
  lsamd::Ls* ls = nullptr;  // LSMR work (built on the first LSMR solve after set_rows)
  // multi-rhs multigrid work per group width 4/8/16/32 (kept across set_rows)
  lsamd::MrWork* mr[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  void* tzero = nullptr;  // zeroed control blocks of the timing entry points
  // buffers of the row set (set_rows), kept across calls: (pointer, bytes)
  std::vector<std::pair<void*, size_t>> rpool;
  lsamd_grid grid;
  GridDev gd;
  lsamd_fd_reg reg;
  int64_t n = 0;
  double* scale2 = nullptr;
  // stencil plan
  bool regular = false;   // class-table kernels usable
  int nbx = 0, nby = 0;   // 64x8 xy tiles
  lsamd::BSys* sys1 = nullptr;           // this system as a one-entry table
  // work items per z-chunk 8/16/32 on 64 x 8 tiles, then on 32 x 16 tiles
  int4* stiles[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int nstile[6] = {0, 0, 0, 0, 0, 0};
  double* tab64 = nullptr;
  float* tab32 = nullptr;
  Coef25<double> ku64;
  Coef25<float> ku32;
  int op = 0;                  // regulariser: 0 FD second derivatives, 1 P1 edge jumps
  int p1nv[3] = {0, 0, 0};     // P1 virtual axis lengths of the table
  double* p1tab64 = nullptr;
  float* p1tab32 = nullptr;
  double p1_ridge2 = 0.0;
  // batched ensemble (op == 2): per-system descriptors and node-block table
  int nsys = 0;
  std::vector<BSys> bsys_h;
  BSys* bsys = nullptr;
  int* blk_sys = nullptr;
  int nblk = 0;
  // stencil tiles of all systems, one table per batch form f: 64 x 8 tiles
  // (f & 1 == 0) or 32 x 16, z-chunk 16 (f & 2 == 0) or 32
  int4* btab[4] = {};
  int nbt[4] = {};
  std::vector<int2> trange[4];  // per system: its range in each table
  int bform = -1;               // form in use (-1: not yet tuned, form 0 until then)
  Coef25<double> bku64;         // bulk coefficients (shared by all systems)
  Coef25<float> bku32;
  CgCtl* bctl = nullptr;        // per-system PCG control
  CgCtl* bctl_apply = nullptr;  // all-zero (no system done): standalone apply
  CgCtl* bctl64 = nullptr;      // FP64 thresholds of the mixed-precision solve
  CgCtl* bctl_host = nullptr;   // pinned mirror
  double* u_rows = nullptr;     // per data row: t_r (a_r . p)
  int* odone = nullptr;         // per system: outer refinement converged
  double* bres = nullptr;       // per system: outer residual
  hipGraphExec_t bgraph[2] = {nullptr, nullptr};
  int bgraph_chunk[2] = {0, 0};
  int zc[2] = {0, 0};          // tuned z-chunk per precision (0: not tuned yet)
  float tune_us[2][9] = {{0}};  // timings of the stencil forms kStCand
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
  const void* minv_graph[2] = {nullptr, nullptr};
  double* ones64 = nullptr;
  float* ones32 = nullptr;
  int graph_chunk[2] = {0, 0};
  std::vector<void*> owned;
};

namespace {

// Calls that run on the legacy null stream (synchronous hipMemcpy and
// hipMemset, device synchronisation, launches on stream 0) would make that
// stream depend on a stream another host thread is capturing into a graph,
// which HIP rejects.  Handles driven from several host threads at once (one
// stream each) therefore take this lock around every such call and around
// every capture region, so the two never overlap.
std::recursive_mutex g_legacy_mu;

hipError_t legacy_memcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  std::lock_guard<std::recursive_mutex> lock(g_legacy_mu);
  return hipMemcpy(dst, src, bytes, kind);
}

hipError_t legacy_memset(void* dst, int value, size_t bytes) {
  std::lock_guard<std::recursive_mutex> lock(g_legacy_mu);
  return hipMemset(dst, value, bytes);
}

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

// Buffer `slot` of the row set: kept across set_rows calls and replaced by a
// larger one (25% headroom) only when the new row set does not fit, so that a
// handle reused for many row sets of similar size stops allocating.
enum RowSlot {
  kRsRowPtr, kRsCols, kRsCoef64, kRsCoef32, kRsB, kRsW, kRsW2_64, kRsW2_32, kRsT64, kRsT32,
  kRsPartRows, kRsURows, kRsBad, kRsErow, kRsKeys, kRsVals, kRsSegEntry, kRsSegRow, kRsUniq,
  kRsCounts, kRsNrun, kRsTmp, kRsOff32, kRsSegPtr, kRsDnode, kRsAdjoint, kRsSlots
};

template <typename P>
int rbuf(lsamd_fd_system* s, int slot, P** p, int64_t count) {
  if (s->rpool.empty()) s->rpool.assign(kRsSlots, {nullptr, 0});
  const size_t bytes = std::max<size_t>(8, (size_t)std::max<int64_t>(count, 1) * sizeof(P));
  auto& e = s->rpool[slot];
  if (e.second < bytes) {
    dfree(s, e.first);
    e = {nullptr, 0};
    void* q = nullptr;
    const size_t want = bytes + bytes / 4;
    if (int rc = dmalloc(s, &q, want)) return rc;
    e = {q, want};
  }
  *p = (P*)e.first;
  return LSAMD_OK;
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

// regulariser (and ridge) part of diag(N)
void launch_reg_diag(lsamd_fd_system* s, hipStream_t st) {
  if (s->op == 2)
    hipLaunchKernelGGL(k_b_diag, dim3((unsigned)s->nblk), dim3(256), 0, st,
                       (const BSys*)s->bsys, (const int*)s->blk_sys, (const double*)s->tab64,
                       s->diag);
  else if (s->op == 1)
    hipLaunchKernelGGL(k_p1_diag, dim3((unsigned)ceil_div(s->n, 256)), dim3(256), 0, st, s->gd,
                       (const double*)s->p1tab64, s->p1nv[0], s->p1nv[1], s->p1_ridge2, s->diag);
  else
    hipLaunchKernelGGL(k_stencil_diag, dim3((unsigned)ceil_div(s->n, 256)), dim3(256), 0, st,
                       s->gd, make_regw<double>(s->reg), s->scale2, s->diag);
}

int vec_blocks(int64_t n) {
  int64_t b = ceil_div(n, VB);
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (int)b;
}

constexpr int kZcCand[3] = {8, 16, 32};

// Stencil forms of a regular system: the z-chunk, plus kDmaZc when the bulk
// planes are staged by direct global -> LDS copies (planes below 2 GiB).
constexpr int kDmaZc = 1000;
constexpr int kNStCand = 9;
constexpr int kTile32 = 2000;  // + z-chunk: register-staged planes on 32 x 16 tiles
constexpr int kStCand[kNStCand] = {8,          16,          32,          kDmaZc + 8,    kDmaZc + 16,
                                   kDmaZc + 32, kTile32 + 8, kTile32 + 16, kTile32 + 32};

bool dma_fits(const GridDev& g, size_t elem) {
  return (int64_t)g.nx * g.ny * (int64_t)elem < ((int64_t)1 << 31);
}

int zc_index(int zc) {
  const int base = zc >= kTile32 ? 3 : 0;
  zc %= kDmaZc;
  return base + (zc == 8 ? 0 : (zc == 16 ? 1 : 2));
}

// P1 plane-streaming stencil configurations (tile width, z-chunk); the
// stencil's `zc` slot of a P1 system holds 1 + the configuration index, or
// kP1Node for the one-node-per-thread kernel (LSAMD_P1_STENCIL=node).
constexpr int kP1NCfg = 6;
constexpr int kP1Cfg[kP1NCfg][2] = {{64, 8}, {64, 16}, {32, 8}, {32, 16}, {64, 32}, {32, 32}};
constexpr int kP1Node = 100;
constexpr int kP1Default = 2;  // (64, 16) before tuning

void p1_tiling(const lsamd_fd_system* s, int cfg, int* ntx, int* nty, int* nzc) {
  const int tx = kP1Cfg[cfg][0], zc = kP1Cfg[cfg][1];
  *ntx = (int)ceil_div(s->gd.nx, tx);
  *nty = (int)ceil_div(s->gd.ny, 512 / tx);
  *nzc = p1_nzchunks(s->gd.nz, zc);
}

int64_t p1_blocks(const lsamd_fd_system* s, int cfg) {
  int ntx, nty, nzc;
  p1_tiling(s, cfg, &ntx, &nty, &nzc);
  return (int64_t)ntx * nty * nzc;
}

// largest partial count over the P1 stencil forms (allocation)
int64_t p1_max_partials(const lsamd_fd_system* s) {
  int64_t m = ceil_div(s->n, 256);
  for (int c = 0; c < kP1NCfg; ++c) m = std::max(m, p1_blocks(s, c));
  return m;
}

int64_t stencil_partials(const lsamd_fd_system* s, int zc) {
  if (s->op == 2)
    return std::max<int64_t>(s->nblk, *std::max_element(s->nbt, s->nbt + 4));
  if (s->op == 1) {
    if (zc == kP1Node) return ceil_div(s->n, 256);
    return p1_blocks(s, (zc >= 1 && zc <= kP1NCfg) ? zc - 1 : kP1Default - 1);
  }
  if (!s->regular) return ceil_div(s->n, 256);
  return s->nstile[zc_index(zc)];
}

// the regular single-system stencil: one launch over the system's work items
template <typename T, int ZC, bool DMA = false, int TXT = 64>
void launch_bulk(lsamd_fd_system* s, const T* x, T* y, const CgCtl* ctl, hipStream_t st) {
  const bool f32 = sizeof(T) == 4;
  const T* tab = f32 ? (const T*)s->tab32 : (const T*)s->tab64;
  Coef25<T> ku;
  for (int q = 0; q < kTaps; ++q) ku.k[q] = f32 ? (T)s->ku32.k[q] : (T)s->ku64.k[q];
  const int zi = zc_index(ZC) + (TXT == 32 ? 3 : 0);
  hipLaunchKernelGGL((k_b_tiles<T, ZC, DMA, TXT>), dim3((unsigned)s->nstile[zi]), dim3(NT), 0, st,
                     x, y, (const BSys*)s->sys1, (const int4*)s->stiles[zi], tab, ku, s->part,
                     ctl);
}

template <typename T>
void launch_b_tiles(lsamd_fd_system* s, const T* x, T* y, const CgCtl* ctl, hipStream_t st) {
  const bool f32 = sizeof(T) == 4;
  Coef25<T> ku;
  for (int q = 0; q < kTaps; ++q) ku.k[q] = f32 ? (T)s->bku32.k[q] : (T)s->bku64.k[q];
  const T* tab = f32 ? (const T*)s->tab32 : (const T*)s->tab64;
  const int f = std::max(0, s->bform);
  const dim3 grid((unsigned)s->nbt[f]);
  const BSys* sys = (const BSys*)s->bsys;
  const int4* tb = (const int4*)s->btab[f];
  switch (f) {
    case 0:
      hipLaunchKernelGGL((k_b_tiles<T, kBZc>), grid, dim3(NT), 0, st, x, y, sys, tb, tab, ku,
                         s->part, ctl);
      break;
    case 1:
      hipLaunchKernelGGL((k_b_tiles<T, kBZc, false, 32>), grid, dim3(NT), 0, st, x, y, sys, tb,
                         tab, ku, s->part, ctl);
      break;
    case 2:
      hipLaunchKernelGGL((k_b_tiles<T, 2 * kBZc>), grid, dim3(NT), 0, st, x, y, sys, tb, tab, ku,
                         s->part, ctl);
      break;
    default:
      hipLaunchKernelGGL((k_b_tiles<T, 2 * kBZc, false, 32>), grid, dim3(NT), 0, st, x, y, sys,
                         tb, tab, ku, s->part, ctl);
  }
}

// Stencil form of a batch: the tables of 64 x 8 and 32 x 16 tiles with
// z-chunks 16 and 32 timed on the FP32 stencil (HIP events, best of three
// rounds of five launches), once; the systems' partial ranges follow the
// chosen table.  LSAMD_BATCH_TILE=64|32 pins the tile width and
// LSAMD_BATCH_ZC=16|32 the z-chunk.
int ensure_batch_tuned(lsamd_fd_system* s, hipStream_t st) {
  if (s->bform >= 0) return LSAMD_OK;
  const char* e = getenv("LSAMD_BATCH_TILE");
  const char* ez = getenv("LSAMD_BATCH_ZC");
  const int tx = e ? atoi(e) : 0, zc = ez ? atoi(ez) : 0;
  std::vector<int> cand;
  for (int f = 0; f < 4; ++f)
    if ((tx != 64 && tx != 32) || tx == ((f & 1) ? 32 : 64))
      if ((zc != 16 && zc != 32) || zc == ((f & 2) ? 2 * kBZc : kBZc)) cand.push_back(f);
  int pick = cand[0];
  if (cand.size() > 1) {
    hipEvent_t e0, e1;
    LSAMD_HIP(hipEventCreate(&e0));
    LSAMD_HIP(hipEventCreate(&e1));
    float best[4] = {1e30f, 1e30f, 1e30f, 1e30f};
    for (int round = 0; round < 3; ++round)
      for (int c : cand) {
        s->bform = c;
        LSAMD_HIP(hipEventRecord(e0, st));
        for (int r = 0; r < 5; ++r) launch_b_tiles<float>(s, s->p32, s->q32, s->bctl_apply, st);
        LSAMD_HIP(hipEventRecord(e1, st));
        LSAMD_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        LSAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
        best[c] = std::min(best[c], ms);
      }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    LSAMD_CHECK_LAUNCH("batch tile tuning");
    for (int c : cand)
      if (best[c] < best[pick]) pick = c;
  }
  s->bform = pick;
  const std::vector<int2>& r = s->trange[pick];
  for (int i = 0; i < s->nsys; ++i) {
    s->bsys_h[i].t0 = r[i].x;
    s->bsys_h[i].t1 = r[i].y;
  }
  LSAMD_HIP(hipMemcpyAsync(s->bsys, s->bsys_h.data(), s->nsys * sizeof(BSys),
                           hipMemcpyHostToDevice, st));
  LSAMD_HIP(hipStreamSynchronize(st));
  return LSAMD_OK;
}

// Launch the regulariser stencil; returns the number of block partials.
template <typename T>
int64_t launch_stencil(lsamd_fd_system* s, const T* x, T* y, const CgCtl* ctl, hipStream_t st,
                       int zc) {
  if (s->op == 2) {
    launch_b_tiles<T>(s, x, y, s->bctl_apply, st);
    return s->nbt[std::max(0, s->bform)];
  }
  if (s->op == 1) {
    const bool f32 = sizeof(T) == 4;
    const T* tab = f32 ? (const T*)s->p1tab32 : (const T*)s->p1tab64;
    if (zc == kP1Node) {
      hipLaunchKernelGGL(k_p1_stencil<T>, dim3((unsigned)ceil_div(s->n, 256)), dim3(256), 0, st,
                         x, y, s->gd, tab, s->p1nv[0], s->p1nv[1], (T)s->p1_ridge2, s->part, ctl);
      return ceil_div(s->n, 256);
    }
    const int cfg = (zc >= 1 && zc <= kP1NCfg) ? zc - 1 : kP1Default - 1;
    int ntx, nty, nzc;
    p1_tiling(s, cfg, &ntx, &nty, &nzc);
    const unsigned nb = (unsigned)((int64_t)ntx * nty * nzc);
    const T r2 = (T)s->p1_ridge2;
    const int* nv = s->p1nv;
    switch (cfg) {
      case 0:
        hipLaunchKernelGGL((k_p1_tiles<T, 64, 8>), dim3(nb), dim3(512), 0, st, x, y, s->gd, tab,
                           nv[0], nv[1], nv[2], r2, ntx, nzc, s->part, ctl);
        break;
      case 1:
        hipLaunchKernelGGL((k_p1_tiles<T, 64, 16>), dim3(nb), dim3(512), 0, st, x, y, s->gd, tab,
                           nv[0], nv[1], nv[2], r2, ntx, nzc, s->part, ctl);
        break;
      case 2:
        hipLaunchKernelGGL((k_p1_tiles<T, 32, 8>), dim3(nb), dim3(512), 0, st, x, y, s->gd, tab,
                           nv[0], nv[1], nv[2], r2, ntx, nzc, s->part, ctl);
        break;
      case 3:
        hipLaunchKernelGGL((k_p1_tiles<T, 32, 16>), dim3(nb), dim3(512), 0, st, x, y, s->gd, tab,
                           nv[0], nv[1], nv[2], r2, ntx, nzc, s->part, ctl);
        break;
      case 4:
        hipLaunchKernelGGL((k_p1_tiles<T, 64, 32>), dim3(nb), dim3(512), 0, st, x, y, s->gd, tab,
                           nv[0], nv[1], nv[2], r2, ntx, nzc, s->part, ctl);
        break;
      default:
        hipLaunchKernelGGL((k_p1_tiles<T, 32, 32>), dim3(nb), dim3(512), 0, st, x, y, s->gd, tab,
                           nv[0], nv[1], nv[2], r2, ntx, nzc, s->part, ctl);
        break;
    }
    return nb;
  }
  if (!s->regular) {
    hipLaunchKernelGGL(k_st_general<T>, dim3((unsigned)ceil_div(s->n, 256)), dim3(256), 0, st, x,
                       y, s->gd, make_regw<T>(s->reg), (const double*)s->scale2, s->part, ctl);
    return ceil_div(s->n, 256);
  }
  switch (zc) {
    case 8: launch_bulk<T, 8>(s, x, y, ctl, st); break;
    case 16: launch_bulk<T, 16>(s, x, y, ctl, st); break;
    case kDmaZc + 8: launch_bulk<T, 8, true>(s, x, y, ctl, st); break;
    case kDmaZc + 16: launch_bulk<T, 16, true>(s, x, y, ctl, st); break;
    case kDmaZc + 32: launch_bulk<T, 32, true>(s, x, y, ctl, st); break;
    case kTile32 + 8: launch_bulk<T, 8, false, 32>(s, x, y, ctl, st); break;
    case kTile32 + 16: launch_bulk<T, 16, false, 32>(s, x, y, ctl, st); break;
    case kTile32 + 32: launch_bulk<T, 32, false, 32>(s, x, y, ctl, st); break;
    default: launch_bulk<T, 32>(s, x, y, ctl, st); break;
  }
  return stencil_partials(s, zc);
}

// P1 autotuner: the tile shape (64 x 8 or 32 x 16) and z-chunk (8 or 16
// planes) of the plane-streaming stencil, timed like the FD candidates.
template <typename T>
int tune_p1(lsamd_fd_system* s, T* x, T* y, hipStream_t st) {
  const int pi = sizeof(T) == 4 ? 1 : 0;
  if (const char* e = getenv("LSAMD_P1_STENCIL")) {
    if (strcmp(e, "node") == 0) {
      s->zc[pi] = kP1Node;
      return LSAMD_OK;
    }
  }
  hipEvent_t e0, e1;
  LSAMD_HIP(hipEventCreate(&e0));
  LSAMD_HIP(hipEventCreate(&e1));
  float best = 1e30f;
  int bestc = kP1Default;
  for (int round = 0; round < 3; ++round)
    for (int c = 1; c <= kP1NCfg; ++c) {
      launch_stencil<T>(s, x, y, s->ctl_apply, st, c);
      LSAMD_HIP(hipEventRecord(e0, st));
      for (int r = 0; r < 5; ++r) launch_stencil<T>(s, x, y, s->ctl_apply, st, c);
      LSAMD_HIP(hipEventRecord(e1, st));
      LSAMD_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      LSAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        bestc = c;
      }
    }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  LSAMD_CHECK_LAUNCH("P1 stencil autotune");
  s->zc[pi] = bestc;
  return LSAMD_OK;
}

// Autotuner: time each stencil form (z-chunk 8/16/32 with register-staged
// planes, 8/16 with LDS-DMA staged planes; HIP events) and keep the fastest.
// LSAMD_STENCIL_ZC in the environment pins the choice (a kStCand value).
template <typename T>
int ensure_tuned(lsamd_fd_system* s, T* x, T* y, hipStream_t st) {
  const int pi = sizeof(T) == 4 ? 1 : 0;
  if (s->zc[pi]) return LSAMD_OK;
  if (s->op == 1) return tune_p1<T>(s, x, y, st);
  if (!s->regular) {
    s->zc[pi] = 1;
    return LSAMD_OK;
  }
  if (const char* e = getenv("LSAMD_STENCIL_ZC")) {
    const int v = atoi(e);
    for (int c = 0; c < kNStCand; ++c)
      if (v == kStCand[c] && (v < kDmaZc || v > kTile32 || dma_fits(s->gd, sizeof(T)))) {
        s->zc[pi] = v;
        return LSAMD_OK;
      }
  }
  hipEvent_t e0, e1;
  LSAMD_HIP(hipEventCreate(&e0));
  LSAMD_HIP(hipEventCreate(&e1));
  // three interleaved rounds of 5 back-to-back launches per candidate; the
  // fastest round of each candidate counts (clock ramps and cache state
  // perturb single rounds)
  float best = 1e30f;
  int bestzc = 16;
  for (int c = 0; c < kNStCand; ++c) s->tune_us[pi][c] = -1.f;
  for (int round = 0; round < 3; ++round)
    for (int c = 0; c < kNStCand; ++c) {
      const int zc = kStCand[c];
      // a chunk longer than the bulk gains nothing; LDS-DMA forms address a
      // plane with 32-bit byte offsets
      if (zc % kDmaZc > 8 && s->gd.nz - 6 < zc % kDmaZc) continue;
      if (zc > kDmaZc && zc < kTile32 && !dma_fits(s->gd, sizeof(T))) continue;
      launch_stencil<T>(s, x, y, s->ctl_apply, st, zc);
      LSAMD_HIP(hipEventRecord(e0, st));
      for (int r = 0; r < 5; ++r) launch_stencil<T>(s, x, y, s->ctl_apply, st, zc);
      LSAMD_HIP(hipEventRecord(e1, st));
      LSAMD_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      LSAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
      const float us = ms * 1e3f / 5;
      if (s->tune_us[pi][c] < 0.f || us < s->tune_us[pi][c]) s->tune_us[pi][c] = us;
      if (us < best) {
        best = us;
        bestzc = zc;
      }
    }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  LSAMD_CHECK_LAUNCH("stencil autotune");
  s->zc[pi] = bestzc;
  return LSAMD_OK;
}

// q = N x (stencil + rows); the scatter kernel's block 0 reduces the two
// partial-dot arrays into pq and alpha = rho / pq.
template <typename T>
int launch_apply(lsamd_fd_system* s, const T* x, T* y, T* t, CgCtl* ctl, hipStream_t st) {
  const bool f32 = sizeof(T) == 4;
  RowsDev R = rows_dev(s, f32);
  const int nrb = s->nrows > 0 ? (int)ceil_div(s->nrows, 256) : 0;
  if (nrb > 0) {
    hipLaunchKernelGGL(k_rows_t<T>, dim3((unsigned)nrb), dim3(256), 0, st, x, R, t, s->part_rows,
                       ctl);
  }
  const int64_t nst = launch_stencil<T>(s, x, y, ctl, st, s->zc[f32 ? 1 : 0]);
  const int nsb = (int)std::max<int64_t>(1, ceil_div(s->ndn, 256));
  hipLaunchKernelGGL((k_rows_scatter<T, T>), dim3((unsigned)nsb), dim3(256), 0, st, R,
                     (const T*)t, y, 1, ctl, 1, (const double*)s->part, (int)nst,
                     (const double*)s->part_rows, nrb);
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
  hipLaunchKernelGGL(k_cg_upd_fin, dim3(1), dim3(VB), 0, st, (const double*)s->part, vb, ctl);
  LSAMD_CHECK_LAUNCH("cg iteration");
  return LSAMD_OK;
}

template <typename T>
int launch_init(lsamd_fd_system* s, const double* b, const T* minv, T* x, T* r, double rtol,
                double atol, int maxiter, hipStream_t st) {
  const int vb = vec_blocks(s->n);
  hipLaunchKernelGGL(k_cg_init<T>, dim3(vb), dim3(VB), 0, st, b, minv, x, r, s->n, s->part);
  hipLaunchKernelGGL(k_cg_init_fin, dim3(1), dim3(VB), 0, st, (const double*)s->part, vb, rtol,
                     atol, maxiter, s->ctl);
  LSAMD_CHECK_LAUNCH("cg init");
  return LSAMD_OK;
}

int ensure_work(lsamd_fd_system* s, bool f32) {
  const int64_t n = s->n;
  if (!f32 && !s->x64) {
    if (int rc = dalloc(s, &s->x64, n)) return rc;
    if (int rc = dalloc(s, &s->r64, n)) return rc;
    if (int rc = dalloc(s, &s->p64, n)) return rc;
    if (int rc = dalloc(s, &s->q64, n)) return rc;
    for (double* v : {s->x64, s->r64, s->p64, s->q64}) LSAMD_HIP(legacy_memset(v, 0, n * 8));
  }
  if (f32 && !s->x32) {
    if (int rc = dalloc(s, &s->x32, n)) return rc;
    if (int rc = dalloc(s, &s->r32, n)) return rc;
    if (int rc = dalloc(s, &s->p32, n)) return rc;
    if (int rc = dalloc(s, &s->q32, n)) return rc;
    for (float* v : {s->x32, s->r32, s->p32, s->q32}) LSAMD_HIP(legacy_memset(v, 0, n * 4));
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
    std::lock_guard<std::recursive_mutex> capture_lock(g_legacy_mu);
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

// Work buffers, control blocks and the regulariser-only diagonal shared by
// every system kind.
int init_common(lsamd_fd_system* s) {
  int64_t pm = 8192;
  if (s->op == 1)
    pm = std::max(pm, p1_max_partials(s));
  else
    for (int c = 0; c < kNStCand; ++c) pm = std::max(pm, stencil_partials(s, kStCand[c]));
  s->part_n = pm;
  if (int rc = dalloc(s, &s->part, s->part_n)) return rc;
  if (int rc = dalloc(s, &s->ctl, 1)) return rc;
  if (int rc = dalloc(s, &s->ctl_apply, 1)) return rc;
  LSAMD_HIP(legacy_memset(s->ctl, 0, sizeof(CgCtl)));
  LSAMD_HIP(legacy_memset(s->ctl_apply, 0, sizeof(CgCtl)));
  if (hipHostMalloc((void**)&s->ctl_host, sizeof(CgCtl), hipHostMallocDefault) != hipSuccess)
    return fail(LSAMD_ERR_ALLOC, "hipHostMalloc failed");
  if (int rc = dalloc(s, &s->rhs, s->n)) return rc;
  if (int rc = dalloc(s, &s->diag, s->n)) return rc;
  if (int rc = dalloc(s, &s->minv64, s->n)) return rc;
  if (int rc = dalloc(s, &s->minv32, s->n)) return rc;
  // regulariser-only system until rows are set
  std::lock_guard<std::recursive_mutex> lock(g_legacy_mu);
  launch_reg_diag(s, 0);
  hipLaunchKernelGGL(k_minv, dim3((unsigned)ceil_div(s->n, 256)), dim3(256), 0, 0, s->diag,
                     s->minv64, s->minv32, s->n);
  hipLaunchKernelGGL(k_zero_d, dim3((unsigned)ceil_div(s->n, 256)), dim3(256), 0, 0, s->rhs, s->n);
  LSAMD_CHECK_LAUNCH("system_create");
  LSAMD_HIP(hipDeviceSynchronize());
  return LSAMD_OK;
}


// ------------------------------------------------------------- batch ----
template <typename T>
int launch_b_iteration(lsamd_fd_system* s, T* x, T* r, T* p, T* q, T* t, const T* minv,
                       hipStream_t st) {
  const bool f32 = sizeof(T) == 4;
  const BSys* sys = s->bsys;
  const int* bs = s->blk_sys;
  hipLaunchKernelGGL(k_b_dir<T>, dim3((unsigned)s->nblk), dim3(256), 0, st, (const T*)r, minv, p,
                     sys, bs, (const CgCtl*)s->bctl);
  RowsDev R = rows_dev(s, f32);
  if (s->nrows > 0)
    hipLaunchKernelGGL(k_b_rows_t<T>, dim3((unsigned)ceil_div(s->nrows, 256)), dim3(256), 0, st,
                       (const T*)p, R, t, s->u_rows);
  launch_b_tiles<T>(s, (const T*)p, q, s->bctl, st);
  if (s->nrows > 0)
    hipLaunchKernelGGL((k_rows_scatter<T, T>), dim3((unsigned)std::max<int64_t>(1, ceil_div(s->ndn, 256))),
                       dim3(256), 0, st, R, (const T*)t, q, 1, s->ctl_apply, 0,
                       (const double*)nullptr, 0, (const double*)nullptr, 0);
  hipLaunchKernelGGL(k_b_alpha, dim3((unsigned)s->nsys), dim3(256), 0, st, (const double*)s->part,
                     (const double*)s->u_rows, sys, s->bctl);
  hipLaunchKernelGGL(k_b_upd<T>, dim3((unsigned)s->nblk), dim3(256), 0, st, x, r, (const T*)p,
                     (const T*)q, minv, sys, bs, s->part, (const CgCtl*)s->bctl);
  hipLaunchKernelGGL(k_b_rho, dim3((unsigned)s->nsys), dim3(256), 0, st, (const double*)s->part,
                     sys, s->bctl);
  LSAMD_CHECK_LAUNCH("batch iteration");
  return LSAMD_OK;
}

template <typename T>
int launch_b_init(lsamd_fd_system* s, const double* b, const T* minv, T* x, T* r, double rtol,
                  double atol, int maxiter, const int* skip, hipStream_t st) {
  hipLaunchKernelGGL(k_b_init<T>, dim3((unsigned)s->nblk), dim3(256), 0, st, b, minv, x, r,
                     (const BSys*)s->bsys, (const int*)s->blk_sys, skip, s->part);
  hipLaunchKernelGGL(k_b_init_fin, dim3((unsigned)s->nsys), dim3(256), 0, st,
                     (const double*)s->part, (const BSys*)s->bsys, rtol, atol, maxiter, skip,
                     s->bctl);
  LSAMD_CHECK_LAUNCH("batch init");
  return LSAMD_OK;
}

int b_all_done(lsamd_fd_system* s, hipStream_t st, bool* done) {
  LSAMD_HIP(hipMemcpyAsync(s->bctl_host, s->bctl, s->nsys * sizeof(CgCtl), hipMemcpyDeviceToHost,
                           st));
  LSAMD_HIP(hipStreamSynchronize(st));
  bool all = true;
  for (int i = 0; i < s->nsys && all; ++i) all = s->bctl_host[i].done != 0;
  *done = all;
  return LSAMD_OK;
}

// chunks of `chunk` batched iterations (one captured graph) until every
// system is done
template <typename T>
int run_b_pcg(lsamd_fd_system* s, int chunk, T* x, T* r, T* p, T* q, T* t, const T* minv,
              hipStream_t st) {
  const int gi = sizeof(T) == 4 ? 1 : 0;
  if (s->bgraph[gi] && s->bgraph_chunk[gi] != chunk) {
    hipGraphExecDestroy(s->bgraph[gi]);
    s->bgraph[gi] = nullptr;
  }
  if (!s->bgraph[gi]) {
    std::lock_guard<std::recursive_mutex> capture_lock(g_legacy_mu);
    hipStream_t cs;
    LSAMD_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t graph;
    LSAMD_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    int rc = LSAMD_OK;
    for (int i = 0; i < chunk && rc == LSAMD_OK; ++i)
      rc = launch_b_iteration<T>(s, x, r, p, q, t, minv, cs);
    hipError_t e = hipStreamEndCapture(cs, &graph);
    hipStreamDestroy(cs);
    if (rc) return rc;
    if (e != hipSuccess) return fail(LSAMD_ERR_HIP, "batch graph capture failed");
    e = hipGraphInstantiate(&s->bgraph[gi], graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(LSAMD_ERR_HIP, "batch graph instantiate failed");
    s->bgraph_chunk[gi] = chunk;
  }
  for (;;) {
    bool done = false;
    if (int rc = b_all_done(s, st, &done)) return rc;
    if (done) break;
    LSAMD_HIP(hipGraphLaunch(s->bgraph[gi], st));
  }
  return LSAMD_OK;
}

int ensure_batch_work(lsamd_fd_system* s) {
  if (int rc = ensure_work(s, false)) return rc;
  if (int rc = ensure_work(s, true)) return rc;
  if (s->nrows > 0 && !s->t64) {
    if (int rc = dalloc(s, &s->t64, s->nrows)) return rc;
    if (int rc = dalloc(s, &s->t32, s->nrows)) return rc;
  }
  return LSAMD_OK;
}

}  // namespace

// ------------------------------------------------------------ multigrid ----
namespace lsamd {

// device arrays of a sliced row set (Sell), row-dependent (pooled)
struct SellBuf {
  int64_t* off = nullptr;
  int32_t* width = nullptr;
  int32_t* col = nullptr;
  float* val = nullptr;
  const int32_t* node = nullptr;
  int64_t nrows = 0;
  int64_t nent = 0;  // entries with padding
};
inline Sell sell_dev(const SellBuf& b) { return Sell{b.off, b.width, b.col, b.val, b.node, b.nrows}; }

struct MgLevel {
  lsamd_fd_system* sys = nullptr;  // level 0: the system itself (not owned)
  bool owned = false;
  // coarse levels: data term as an assembled Gram matrix (node rows, FP32)
  int64_t* gptr = nullptr;
  int32_t* gcol = nullptr;
  float* gval = nullptr;
  double* gval64 = nullptr;
  GridDev g;
  int64_t n = 0;
  float *b = nullptr, *x = nullptr, *q = nullptr, *d = nullptr;
  float c0 = 0.f, c1[kMgDegMax] = {0.f}, c2[kMgDegMax] = {0.f};
  double lambda = 0.0;
  // smoother: tags (Jacobi scale, or -1 on Schwarz rows) and the summed block
  // inverses as CSR over all nodes (sptr == nullptr: point Jacobi only)
  float* sm = nullptr;
  int64_t* sptr = nullptr;
  int32_t* scol = nullptr;
  float* sval = nullptr;
  double* sval64 = nullptr;
  int64_t nblocks = 0;
  int32_t* cov = nullptr;  // nodes with CSR rows, listed on large levels (split smoother)
  int64_t ncov = 0;
  // the multi-rhs kernels' copies of the smoother rows (of the covered list
  // on split levels, else of every node, Jacobi nodes as one-entry rows) and
  // of the Gram rows, in 64-row slices; built on the first multi-rhs solve
  // after the rows change
  SellBuf ssl, gsl;
  // levels with a covered list: the covered rows sliced, their columns as
  // covered-list positions (the multi-rhs step gathers an interleaved copy
  // of r over the list), and each node's list position (-1: uncovered)
  SellBuf csl;
  int32_t* cpos = nullptr;
};

struct Mg {
  std::vector<MgLevel> lv;
  int deg = 4;            // Chebyshev degree of the coarse levels
  int deg0 = 2;           // Chebyshev degree of the fine level (<= deg)
  // coarse-level degree of the multi-rhs V-cycle: its fine level is HBM-bound
  // over R lanes, so coarse steps are cheap against the iterations they save
  // (at 128^3, 32 lanes: degree 4 101 iterations 757 ms per batch, 5 88 and
  // 713, 6 80 and 708, 7 75 and 726); the single-system and batched solves,
  // launch-bound, are fastest at 4
  int deg_mr = 6;
  bool schwarz = true;    // Schwarz blocks in the smoother (else point Jacobi)
  bool stale = true;      // data rows changed: row-dependent parts to rebuild
  float* cinv = nullptr;  // dense inverse of the coarsest operator
  int nc = 0;
  double* rp64 = nullptr;
  double* scal = nullptr;  // power-iteration scalars
  // memory: `mem` lives as long as the hierarchy (grids, level vectors);
  // `dmem` depends on the data rows and returns to `pool` when they change,
  // so that a rebuild with rows of the same sizes allocates nothing
  std::vector<void*> mem;
  std::vector<std::pair<void*, size_t>> dmem, pool;
  hipGraphExec_t graph = nullptr;
  int graph_chunk = 0;
  bool graph_stale = false;  // captured with addresses of an older row set
  bool sell_ready = false;   // the levels' sliced rows match the row set
  // level-0 data term as a Gram matrix (CSR, FP32 and FP64 values), kept
  // for the batched solve's node-wise operator (row-dependent, pooled)
  int64_t* g0ptr = nullptr;
  int32_t* g0col = nullptr;
  float* g0v32 = nullptr;
  double* g0v64 = nullptr;
  int32_t* g0list = nullptr;  // the nodes with a nonempty Gram row
  int64_t g0n = 0;
};

void mg_destroy(Mg* m) {
  if (!m) return;
  if (m->graph) hipGraphExecDestroy(m->graph);
  for (auto& L : m->lv)
    if (L.owned && L.sys) lsamd_fd_system_destroy(L.sys);
  for (void* p : m->mem) hipFree(p);
  for (auto& b : m->dmem) hipFree(b.first);
  for (auto& b : m->pool) hipFree(b.first);
  delete m;
}

// The data rows changed: return their memory to the pool and mark the
// captured iteration stale (it holds their addresses); the next solve
// re-captures it and updates the executable graph in place when the
// topology is unchanged, which costs far less than instantiating anew.
void mg_invalidate(Mg* m) {
  m->graph_stale = m->graph != nullptr;
  m->g0ptr = nullptr;
  m->g0col = nullptr;
  m->g0v32 = nullptr;
  m->g0v64 = nullptr;
  m->g0list = nullptr;
  m->g0n = 0;
  for (auto& b : m->dmem) m->pool.push_back(b);
  m->dmem.clear();
  for (MgLevel& L : m->lv) {
    L.gptr = nullptr;
    L.gcol = nullptr;
    L.gval = nullptr;
    L.gval64 = nullptr;
    L.sm = nullptr;
    L.sptr = nullptr;
    L.scol = nullptr;
    L.sval = nullptr;
    L.sval64 = nullptr;
    L.nblocks = 0;
    L.cov = nullptr;
    L.ncov = 0;
    L.ssl = SellBuf();
    L.gsl = SellBuf();
    L.csl = SellBuf();
    L.cpos = nullptr;
  }
  m->sell_ready = false;
  m->stale = true;
}

namespace {

template <typename P>
int mg_alloc(Mg* m, P** p, int64_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, std::max<int64_t>(count, 1) * sizeof(P)) != hipSuccess)
    return fail(LSAMD_ERR_ALLOC, "hipMalloc failed (multigrid)");
  m->mem.push_back(q);
  *p = (P*)q;
  return LSAMD_OK;
}

// row-dependent memory: the smallest pooled block that fits (and wastes at
// most half of itself), else a new one
void* mg_take(Mg* m, size_t bytes) {
  bytes = std::max<size_t>((bytes + 255) & ~(size_t)255, 256);
  int best = -1;
  for (size_t i = 0; i < m->pool.size(); ++i) {
    const size_t b = m->pool[i].second;
    if (b >= bytes && b <= 2 * bytes && (best < 0 || b < m->pool[best].second)) best = (int)i;
  }
  void* q = nullptr;
  size_t got = bytes;
  if (best >= 0) {
    q = m->pool[best].first;
    got = m->pool[best].second;
    m->pool.erase(m->pool.begin() + best);
  } else if (hipMalloc(&q, bytes) != hipSuccess) {
    return nullptr;
  }
  m->dmem.push_back({q, got});
  return q;
}
template <typename P>
int mg_dalloc(Mg* m, P** p, int64_t count) {
  void* q = mg_take(m, (size_t)std::max<int64_t>(count, 1) * sizeof(P));
  if (!q) return fail(LSAMD_ERR_ALLOC, "hipMalloc failed (multigrid)");
  *p = (P*)q;
  return LSAMD_OK;
}
// return one row-dependent block to the pool (no-op for nullptr)
void mg_give(Mg* m, const void* p) {
  if (!p) return;
  for (size_t i = 0; i < m->dmem.size(); ++i)
    if (m->dmem[i].first == p) {
      m->pool.push_back(m->dmem[i]);
      m->dmem.erase(m->dmem.begin() + i);
      return;
    }
}

// y = N x of a level, without the alpha reduction; `ctl` only gates on done
template <typename T>
void mg_apply(lsamd_fd_system* s, const T* x, T* y, const CgCtl* ctl, hipStream_t st) {
  const bool f32 = sizeof(T) == 4;
  RowsDev R = rows_dev(s, f32);
  T* t = f32 ? (T*)s->t32 : (T*)s->t64;
  const int nrb = s->nrows > 0 ? (int)ceil_div(s->nrows, 256) : 0;
  if (nrb > 0)
    hipLaunchKernelGGL(k_rows_t<T>, dim3((unsigned)nrb), dim3(256), 0, st, x, R, t, s->part_rows,
                       const_cast<CgCtl*>(ctl));
  launch_stencil<T>(s, x, y, ctl, st, s->zc[f32 ? 1 : 0]);
  if (nrb > 0)
    hipLaunchKernelGGL((k_rows_scatter<T, T>),
                       dim3((unsigned)std::max<int64_t>(1, ceil_div(s->ndn, 256))), dim3(256), 0,
                       st, R, (const T*)t, y, 1, const_cast<CgCtl*>(ctl), 0,
                       (const double*)nullptr, 0, (const double*)nullptr, 0);
}

// level apply in FP32: the fine level uses the data rows, coarse levels the
// assembled Gram matrix of their data term (rows there are many per node)
MgDev mg_dev(const MgLevel& L) {
  MgDev d;
  const lsamd_fd_system* s = L.sys;
  d.g = L.g;
  d.n = L.n;
  d.tab = s->regular ? s->tab32 : nullptr;
  d.w = make_regw<float>(s->reg);
  d.gptr = L.gptr;
  d.gcol = L.gcol;
  d.gval = L.gval;
  return d;
}

MgSm mg_sm(const MgLevel& L) {
  MgSm S;
  S.sm = L.sm ? L.sm : L.sys->minv32;
  S.sptr = L.sptr;
  S.scol = L.scol;
  S.sval = L.sval;
  return S;
}

// y = N x of a level in FP32: level 0 with its data rows, coarse levels in
// one fused launch (regulariser stencil + Gram row per node)
void mg_level_apply(const MgLevel& L, const float* x, float* y, const CgCtl* ctl,
                    hipStream_t st) {
  if (!L.owned) {
    mg_apply<float>(L.sys, x, y, ctl, st);
    return;
  }
  hipLaunchKernelGGL(k_mg_apply, dim3((unsigned)ceil_div(L.n, 256)), dim3(256), 0, st, mg_dev(L),
                     x, y, ctl);
}

// FP64 level apply (setup: power iteration and the dense coarsest operator)
int mg_level_apply64(const MgLevel& L, const double* x, double* y, hipStream_t st) {
  lsamd_fd_system* s = L.sys;
  if (!L.owned) return launch_apply<double>(s, x, y, s->t64, s->ctl_apply, st);
  launch_stencil<double>(s, x, y, s->ctl_apply, st, s->zc[0]);
  if (L.gptr)
    hipLaunchKernelGGL(k_gram_spmv<double>, dim3((unsigned)ceil_div(L.n, 256)), dim3(256), 0, st,
                       (const int64_t*)L.gptr, (const int32_t*)L.gcol, (const double*)L.gval64,
                       x, y, L.n, (const CgCtl*)s->ctl_apply);
  LSAMD_CHECK_LAUNCH("multigrid level apply");
  return LSAMD_OK;
}

lsamd_grid mg_coarse_grid(const lsamd_grid& f) {
  lsamd_grid c = f;
  for (int a = 0; a < 3; ++a) {
    const int64_t cells = f.nnodes[a] - 1;
    c.nnodes[a] = (cells + 1) / 2 + 1;
    c.step[a] = 2.0 * f.step[a];
  }
  return c;
}

int mg_prepare_level(lsamd_fd_system* sys, hipStream_t st) {
  if (int rc = ensure_work(sys, false)) return rc;
  if (int rc = ensure_work(sys, true)) return rc;
  if (int rc = ensure_tuned<double>(sys, sys->p64, sys->q64, st)) return rc;
  if (int rc = ensure_tuned<float>(sys, sys->p32, sys->q32, st)) return rc;
  return LSAMD_OK;
}

// lambda_max(D^-1 N) by power iteration (FP64), on the level's work vectors
int mg_lambda(Mg* m, const MgLevel& L, double* out, hipStream_t st) {
  lsamd_fd_system* sys = L.sys;
  const int64_t n = sys->n;
  const int vb = vec_blocks(n);
  double *v = sys->p64, *y = sys->q64, *w = sys->r64;
  const dim3 nb((unsigned)ceil_div(n, 256));
  hipLaunchKernelGGL(k_mg_start, nb, dim3(256), 0, st, v, n);
  for (int it = 0; it < 20; ++it) {
    if (int rc = mg_level_apply64(L, v, y, st)) return rc;
    hipLaunchKernelGGL(k_mg_power, dim3(vb), dim3(256), 0, st, (const double*)y,
                       (const double*)sys->minv64, (const float*)(L.sm ? L.sm : sys->minv32),
                       (const int64_t*)L.sptr, (const int32_t*)L.scol, (const double*)L.sval64,
                       (const double*)v, w, n, sys->part);
    hipLaunchKernelGGL(k_mg_norm_fin, dim3(1), dim3(256), 0, st, (const double*)sys->part, vb,
                       m->scal);
    hipLaunchKernelGGL(k_mg_scale, nb, dim3(256), 0, st, (const double*)w, v, n,
                       (const double*)m->scal);
  }
  LSAMD_CHECK_LAUNCH("multigrid power iteration");
  double sc[2] = {0.0, 0.0};
  LSAMD_HIP(hipMemcpyAsync(sc, m->scal, 16, hipMemcpyDeviceToHost, st));
  LSAMD_HIP(hipStreamSynchronize(st));
  *out = sc[0];
  return LSAMD_OK;
}

// dense inverse (Gauss-Jordan, partial pivoting) of the n x n matrix a
bool mg_invert(std::vector<double>& a, int n, std::vector<double>& inv) {
  inv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (fabs(a[(size_t)r * n + c]) > fabs(a[(size_t)piv * n + c])) piv = r;
    if (!(fabs(a[(size_t)piv * n + c]) > 0.0)) return false;
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[(size_t)c * n + k], a[(size_t)piv * n + k]);
        std::swap(inv[(size_t)c * n + k], inv[(size_t)piv * n + k]);
      }
    const double d = 1.0 / a[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      a[(size_t)c * n + k] *= d;
      inv[(size_t)c * n + k] *= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
      }
    }
  }
  return true;
}

// G = A^T W^2 A of a coarse level's (fixed-width) data rows: pairs, stable
// radix sort by (row, col), reduce by key in sorted order (deterministic)
// (row << 32 | col, value) pairs -> CSR over n rows with duplicates summed
// (radix sort + reduce-by-key); the outputs are level memory, the pair
// arrays are the caller's
struct CsrOut {
  int64_t* ptr = nullptr;
  int32_t* col = nullptr;
  float* v32 = nullptr;
  double* v64 = nullptr;
};
int pairs_csr(Mg* m, const uint64_t* key, const double* val, int64_t np, int64_t n, CsrOut* out,
              hipStream_t st) {
  if (np > 2000000000LL) return fail(LSAMD_ERR_ARG, "multigrid: too many matrix entries");
  std::vector<void*> tmp;
  auto talloc = [&](void** p, size_t bytes) {
    *p = mg_take(m, bytes);
    if (!*p) return false;
    tmp.push_back(*p);
    return true;
  };
  auto cleanup = [&]() {
    for (void* p : tmp) mg_give(m, p);
  };
  uint64_t *key_s = nullptr, *ukey = nullptr;
  double *val_s = nullptr, *uval = nullptr;
  int32_t* nrun = nullptr;
  if (!talloc((void**)&key_s, np * 8) || !talloc((void**)&val_s, np * 8) ||
      !talloc((void**)&ukey, np * 8) || !talloc((void**)&uval, np * 8) ||
      !talloc((void**)&nrun, 8)) {
    cleanup();
    return fail(LSAMD_ERR_ALLOC, "hipMalloc failed (multigrid CSR)");
  }
  int bits = 1;
  while (bits < 31 && ((int64_t)1 << bits) < n) ++bits;
  size_t tb1 = 0, tb2 = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, key, key_s, val, val_s, (int)np, 0, 32 + bits,
                                     st);
  hipcub::DeviceReduce::ReduceByKey(nullptr, tb2, key_s, ukey, val_s, uval, nrun, hipcub::Sum(),
                                    (int)np, st);
  void* work = nullptr;
  const size_t tb = std::max(tb1, tb2);
  if (!talloc(&work, tb)) {
    cleanup();
    return fail(LSAMD_ERR_ALLOC, "hipMalloc failed (multigrid CSR)");
  }
  size_t t1 = tb, t2 = tb;
  hipcub::DeviceRadixSort::SortPairs(work, t1, key, key_s, val, val_s, (int)np, 0, 32 + bits, st);
  hipcub::DeviceReduce::ReduceByKey(work, t2, key_s, ukey, val_s, uval, nrun, hipcub::Sum(),
                                    (int)np, st);
  int32_t nrun_h = 0;
  LSAMD_HIP(hipMemcpyAsync(&nrun_h, nrun, 4, hipMemcpyDeviceToHost, st));
  LSAMD_HIP(hipStreamSynchronize(st));
  const int64_t m_ = nrun_h;
  int rc = mg_dalloc(m, &out->ptr, n + 1);
  if (!rc) rc = mg_dalloc(m, &out->col, m_);
  if (!rc) rc = mg_dalloc(m, &out->v32, m_);
  if (!rc) rc = mg_dalloc(m, &out->v64, m_);
  if (rc) {
    cleanup();
    return rc;
  }
  hipLaunchKernelGGL(k_gram_ptr, dim3((unsigned)ceil_div(n + 1, 256)), dim3(256), 0, st,
                     (const uint64_t*)ukey, m_, n, out->ptr);
  hipLaunchKernelGGL(k_gram_cols, dim3((unsigned)ceil_div(std::max<int64_t>(m_, 1), 256)),
                     dim3(256), 0, st, (const uint64_t*)ukey, (const double*)uval, m_, out->col,
                     out->v32, out->v64);
  LSAMD_CHECK_LAUNCH("multigrid CSR");
  cleanup();  // stream-ordered: the pooled blocks are only reused by later work of st
  return LSAMD_OK;
}

// Gram matrix G = A^T W^2 A of fixed-width rows (wc columns each) on n nodes
int gram_csr(Mg* m, const int32_t* rcols, const double* rcoef, const double* w2, int64_t nr,
             int wc, int64_t n, CsrOut* out, hipStream_t st) {
  const int64_t np = nr * wc * wc;
  if (np > 2000000000LL) return fail(LSAMD_ERR_ARG, "multigrid: data term too large");
  uint64_t* key = nullptr;
  double* val = nullptr;
  if (int rc = mg_dalloc(m, &key, np)) return rc;
  if (int rc = mg_dalloc(m, &val, np)) return rc;
  hipLaunchKernelGGL(k_gram_pairs, dim3((unsigned)ceil_div(nr, 128)), dim3(128), 0, st, rcols,
                     rcoef, w2, nr, wc, key, val);
  const int rc = pairs_csr(m, key, val, np, n, out, st);
  mg_give(m, key);
  mg_give(m, val);
  return rc;
}

int mg_gram(Mg* m, MgLevel& L, const int32_t* rcols, const double* rcoef, const double* w2,
            int64_t nr, int wc, hipStream_t st) {
  lsamd_fd_system* cs = L.sys;
  CsrOut G;
  if (int rc = gram_csr(m, rcols, rcoef, w2, nr, wc, L.n, &G, st)) return rc;
  L.gptr = G.ptr;
  L.gcol = G.col;
  L.gval = G.v32;
  L.gval64 = G.v64;
  // coarse diagonal: regulariser (already in cs->diag) + Gram diagonal
  hipLaunchKernelGGL(k_gram_diag, dim3((unsigned)ceil_div(L.n, 256)), dim3(256), 0, st,
                     (const int64_t*)L.gptr, (const int32_t*)L.gcol, (const double*)L.gval64,
                     cs->diag, cs->minv64, cs->minv32, L.n);
  LSAMD_CHECK_LAUNCH("multigrid Gram diagonal");
  return LSAMD_OK;
}

// Schwarz blocks of one level: boundary-layer blocks of its grid and the
// 8-node blocks of `cells` (min-corner node ids); G is the level's data Gram
// matrix (nullable).  Builds L.sm / L.sptr / L.scol / L.sval / L.sval64.
int mg_schwarz(Mg* m, MgLevel& L, const std::vector<int64_t>& cells, const CsrOut& G,
               hipStream_t st) {
  const GridDev g = L.g;
  const int64_t nx = g.nx, ny = g.ny, nxy = nx * ny;
  struct Grp {
    int start, len;
    bool bnd;
  };
  auto groups = [](int n) {
    std::vector<Grp> v;
    if (n < 6) {  // the first and last three would overlap
      v.push_back({0, n, true});
      return v;
    }
    v.push_back({0, 3, true});
    for (int i = 3; i < n - 3; ++i) v.push_back({i, 1, false});
    v.push_back({n - 3, 3, true});
    return v;
  };
  const std::vector<Grp> gx = groups(g.nx), gy = groups(g.ny), gz = groups(g.nz);
  std::vector<Grp> gxb;
  for (const Grp& a : gx)
    if (a.bnd) gxb.push_back(a);
  std::vector<int64_t> boff(1, 0), moff(1, 0);
  std::vector<int32_t> nodes;
  auto add_block = [&](const Grp& a, const Grp& b, const Grp& c) {
    for (int k = c.start; k < c.start + c.len; ++k)
      for (int j = b.start; j < b.start + b.len; ++j)
        for (int i = a.start; i < a.start + a.len; ++i)
          nodes.push_back((int32_t)(i + nx * j + nxy * k));
    const int64_t sz = (int64_t)a.len * b.len * c.len;
    boff.push_back(boff.back() + sz);
    moff.push_back(moff.back() + sz * sz);
  };
  for (const Grp& c : gz)
    for (const Grp& b : gy) {
      const std::vector<Grp>& xs = (c.bnd || b.bnd) ? gx : gxb;
      for (const Grp& a : xs) add_block(a, b, c);
    }
  for (int64_t cell : cells) {
    for (int dk = 0; dk < 2; ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) nodes.push_back((int32_t)(cell + di + nx * dj + nxy * dk));
    boff.push_back(boff.back() + 8);
    moff.push_back(moff.back() + 64);
  }
  const int nb = (int)boff.size() - 1;
  L.nblocks = nb;
  if (nb == 0) return LSAMD_OK;
  const int64_t ntot = moff.back();
  std::vector<void*> tmp;
  auto talloc = [&](void** p, size_t bytes) {
    *p = mg_take(m, bytes);
    if (!*p) return false;
    tmp.push_back(*p);
    return true;
  };
  auto cleanup = [&]() {
    for (void* p : tmp) mg_give(m, p);
  };
  int64_t *dboff = nullptr, *dmoff = nullptr;
  int32_t* dnodes = nullptr;
  double *dense = nullptr, *val = nullptr;
  uint64_t* key = nullptr;
  int* bad = nullptr;
  if (!talloc((void**)&dboff, boff.size() * 8) || !talloc((void**)&dmoff, moff.size() * 8) ||
      !talloc((void**)&dnodes, nodes.size() * 4) || !talloc((void**)&dense, ntot * 8) ||
      !talloc((void**)&val, ntot * 8) || !talloc((void**)&key, ntot * 8) ||
      !talloc((void**)&bad, 4)) {
    cleanup();
    return fail(LSAMD_ERR_ALLOC, "hipMalloc failed (multigrid smoother blocks)");
  }
  LSAMD_HIP(hipMemcpyAsync(dboff, boff.data(), boff.size() * 8, hipMemcpyHostToDevice, st));
  LSAMD_HIP(hipMemcpyAsync(dmoff, moff.data(), moff.size() * 8, hipMemcpyHostToDevice, st));
  LSAMD_HIP(hipMemcpyAsync(dnodes, nodes.data(), nodes.size() * 4, hipMemcpyHostToDevice, st));
  LSAMD_HIP(hipMemsetAsync(bad, 0, 4, st));
  lsamd_fd_system* S = L.sys;
  hipLaunchKernelGGL(k_blk_fill, dim3((unsigned)ceil_div(ntot, 256)), dim3(256), 0, st,
                     (const int64_t*)dboff, (const int32_t*)dnodes, (const int64_t*)dmoff, nb,
                     ntot, g, make_regw<double>(S->reg), (const double*)S->scale2,
                     (const int64_t*)G.ptr, (const int32_t*)G.col, (const double*)G.v64, dense);
  hipLaunchKernelGGL(k_blk_inv, dim3((unsigned)nb), dim3(64), 0, st,
                     (const int64_t*)dboff, (const int64_t*)dmoff, nb, dense, bad);
  hipLaunchKernelGGL(k_blk_pairs, dim3((unsigned)ceil_div(ntot, 256)), dim3(256), 0, st,
                     (const int64_t*)dboff, (const int32_t*)dnodes, (const int64_t*)dmoff, nb,
                     ntot, (const double*)dense, key, val);
  LSAMD_CHECK_LAUNCH("multigrid smoother blocks");
  int hbad = 0;
  LSAMD_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, st));
  LSAMD_HIP(hipStreamSynchronize(st));
  if (hbad) {
    cleanup();
    return fail(LSAMD_ERR_ARG, "multigrid: a smoother block is not positive definite");
  }
  CsrOut M;
  if (int rc = pairs_csr(m, key, val, ntot, L.n, &M, st)) {
    cleanup();
    return rc;
  }
  L.sptr = M.ptr;
  L.scol = M.col;
  L.sval = M.v32;
  L.sval64 = M.v64;
  if (int rc = mg_dalloc(m, &L.sm, L.n)) {
    cleanup();
    return rc;
  }
  hipLaunchKernelGGL(k_blk_tags, dim3((unsigned)ceil_div(L.n, 256)), dim3(256), 0, st,
                     (const int64_t*)L.sptr, (const float*)S->minv32, L.sm, L.n);
  cleanup();
  LSAMD_CHECK_LAUNCH("multigrid smoother tags");
  if (L.n >= kCovListMin) {
    // compact list of the covered nodes (their tags are negative)
    uint8_t* flag = nullptr;
    int32_t *cov = nullptr, *nsel = nullptr;
    if (int rc = mg_dalloc(m, &flag, L.n)) return rc;
    if (int rc = mg_dalloc(m, &cov, L.n)) return rc;
    if (int rc = mg_dalloc(m, &nsel, 1)) return rc;
    hipLaunchKernelGGL(k_blk_flags, dim3((unsigned)ceil_div(L.n, 256)), dim3(256), 0, st,
                       (const float*)L.sm, flag, L.n);
    hipcub::CountingInputIterator<int32_t> it(0);
    size_t tb = 0;
    LSAMD_HIP(hipcub::DeviceSelect::Flagged(nullptr, tb, it, flag, cov, nsel, (int)L.n, st));
    void* work = mg_take(m, tb);
    if (!work) return fail(LSAMD_ERR_ALLOC, "hipMalloc failed (multigrid covered list)");
    LSAMD_HIP(hipcub::DeviceSelect::Flagged(work, tb, it, flag, cov, nsel, (int)L.n, st));
    int32_t h = 0;
    LSAMD_HIP(hipMemcpyAsync(&h, nsel, 4, hipMemcpyDeviceToHost, st));
    LSAMD_HIP(hipStreamSynchronize(st));
    mg_give(m, work);
    mg_give(m, flag);
    mg_give(m, nsel);
    L.cov = cov;
    L.ncov = h;
  }
  return LSAMD_OK;
}

// Sliced copy (Sell) of the CSR rows (ptr, col, val) of the nodes `node`
// (nullptr: nodes 0 .. nrows-1); with tags sm, a node whose tag is >= 0 gets
// the one-entry row (node, tag).
int sell_build(Mg* m, const int64_t* ptr, const int32_t* col, const float* val, const float* sm,
               const int32_t* node, int64_t nrows, SellBuf* out, hipStream_t st) {
  *out = SellBuf();
  if (nrows <= 0) return LSAMD_OK;
  const int64_t ns = ceil_div(nrows, 64);
  if (int rc = mg_dalloc(m, &out->width, ns)) return rc;
  if (int rc = mg_dalloc(m, &out->off, ns)) return rc;
  hipLaunchKernelGGL(k_sell_width, dim3((unsigned)ceil_div(ns, 256)), dim3(256), 0, st, ptr, sm,
                     node, nrows, out->width);
  LSAMD_CHECK_LAUNCH("sliced rows: widths");
  std::vector<int32_t> hw(ns);
  std::vector<int64_t> ho(ns);
  LSAMD_HIP(hipMemcpyAsync(hw.data(), out->width, ns * 4, hipMemcpyDeviceToHost, st));
  LSAMD_HIP(hipStreamSynchronize(st));
  int64_t tot = 0;
  for (int64_t i = 0; i < ns; ++i) {
    ho[i] = tot;
    tot += 64 * (int64_t)hw[i];
  }
  if (tot >= ((int64_t)1 << 40)) return fail(LSAMD_ERR_ARG, "sliced rows: too many entries");
  if (int rc = mg_dalloc(m, &out->col, tot)) return rc;
  if (int rc = mg_dalloc(m, &out->val, tot)) return rc;
  LSAMD_HIP(hipMemcpyAsync(out->off, ho.data(), ns * 8, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_sell_fill, dim3((unsigned)ceil_div(ns * 64, 256)), dim3(256), 0, st, ptr,
                     col, val, sm, node, nrows, (const int64_t*)out->off,
                     (const int32_t*)out->width, out->col, out->val);
  LSAMD_CHECK_LAUNCH("sliced rows: fill");
  LSAMD_HIP(hipStreamSynchronize(st));  // ho is read by the copy
  out->node = node;
  out->nrows = nrows;
  out->nent = tot;
  return LSAMD_OK;
}

// cells (min-corner node ids) of the level-0 data rows: every run of 8
// entries of a row whose nodes lie within one cell gives that cell
int mg_data_cells(lsamd_fd_system* s, std::vector<int64_t>& cells, int* wmax, hipStream_t st) {
  cells.clear();
  *wmax = 0;
  if (s->nrows == 0) return LSAMD_OK;
  const GridDev g = s->gd;
  std::vector<int64_t> rp(s->nrows + 1);
  std::vector<int32_t> cols(s->nnz);
  LSAMD_HIP(hipMemcpyAsync(rp.data(), s->row_ptr, rp.size() * 8, hipMemcpyDeviceToHost, st));
  LSAMD_HIP(hipMemcpyAsync(cols.data(), s->cols, cols.size() * 4, hipMemcpyDeviceToHost, st));
  LSAMD_HIP(hipStreamSynchronize(st));
  for (int64_t r = 0; r < s->nrows; ++r) *wmax = std::max<int>(*wmax, (int)(rp[r + 1] - rp[r]));
  if (g.nx < 2 || g.ny < 2 || g.nz < 2) return LSAMD_OK;
  const int64_t nx = g.nx, nxy = (int64_t)g.nx * g.ny;
  for (int64_t r = 0; r < s->nrows; ++r)
    for (int64_t e0 = rp[r]; e0 < rp[r + 1]; e0 += 8) {
      const int64_t e1 = std::min<int64_t>(e0 + 8, rp[r + 1]);
      int lo[3] = {1 << 30, 1 << 30, 1 << 30}, hi[3] = {-1, -1, -1};
      for (int64_t e = e0; e < e1; ++e) {
        const int64_t id = cols[e];
        const int c[3] = {(int)(id % nx), (int)((id / nx) % g.ny), (int)(id / nxy)};
        for (int a = 0; a < 3; ++a) {
          lo[a] = std::min(lo[a], c[a]);
          hi[a] = std::max(hi[a], c[a]);
        }
      }
      if (hi[0] - lo[0] > 1 || hi[1] - lo[1] > 1 || hi[2] - lo[2] > 1) continue;
      const int ci = std::min(lo[0], g.nx - 2), cj = std::min(lo[1], g.ny - 2),
                ck = std::min(lo[2], g.nz - 2);
      cells.push_back(ci + nx * cj + nxy * ck);
    }
  std::sort(cells.begin(), cells.end());
  cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
  return LSAMD_OK;
}

// cells of the next coarser level: each fine cell lies in coarse cell c/2
void mg_coarse_cells(const GridDev& gf, const GridDev& gc, std::vector<int64_t>& cells) {
  const int64_t fnx = gf.nx, fnxy = (int64_t)gf.nx * gf.ny;
  const int64_t cnx = gc.nx, cnxy = (int64_t)gc.nx * gc.ny;
  for (int64_t& c : cells) {
    const int i = (int)(c % fnx), j = (int)((c / fnx) % gf.ny), k = (int)(c / fnxy);
    const int ci = std::min(i >> 1, gc.nx - 2), cj = std::min(j >> 1, gc.ny - 2),
              ck = std::min(k >> 1, gc.nz - 2);
    c = ci + cnx * cj + cnxy * ck;
  }
  std::sort(cells.begin(), cells.end());
  cells.erase(std::unique(cells.begin(), cells.end()), cells.end());
}
int mg_level_vectors(Mg* m, MgLevel& L) {
  if (int rc = mg_alloc(m, &L.b, L.n)) return rc;
  if (int rc = mg_alloc(m, &L.x, L.n)) return rc;
  if (int rc = mg_alloc(m, &L.q, L.n)) return rc;
  if (int rc = mg_alloc(m, &L.d, L.n)) return rc;
  for (float* v : {L.b, L.x, L.q, L.d}) LSAMD_HIP(legacy_memset(v, 0, L.n * 4));
  return LSAMD_OK;
}

// LSAMD_MG_TIMING=1: per-phase wall time of the multigrid setup on stderr
struct MgTimer {
  hipStream_t st;
  bool on;
  double t0 = 0.0;
  MgTimer(hipStream_t s) : st(s), on(getenv("LSAMD_MG_TIMING") != nullptr) { tick(nullptr); }
  void tick(const char* what) {
    if (!on) return;
    hipStreamSynchronize(st);
    const double t =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
            .count();
    if (what) fprintf(stderr, "[multigrid setup] %-24s %8.2f ms\n", what, t - t0);
    t0 = t;
  }
};

// Row-independent part of the hierarchy: the coarse grids as FD systems
// (re-discretised regulariser, tuned stencils, work vectors) and the level
// vectors.  Built once per system.
int mg_build_static(lsamd_fd_system* s, Mg* m, hipStream_t st) {
  if (int rc = mg_prepare_level(s, st)) return rc;
  if (int rc = mg_alloc(m, &m->scal, 2)) return rc;
  const char* sm_env = getenv("LSAMD_MG_SMOOTHER");
  m->schwarz = !(sm_env && strcmp(sm_env, "jacobi") == 0);
  // degree 4 on the coarse levels, 2 on the fine one: at 128^3 96 outer
  // iterations against 145 with degree 2 everywhere, and the coarse steps are
  // cheap; fine-level steps cost HBM traffic
  if (const char* dg = getenv("LSAMD_MG_DEGREE")) m->deg = std::max(1, std::min(kMgDegMax, atoi(dg)));
  if (const char* dg = getenv("LSAMD_MG_DEGREE0")) m->deg0 = std::max(1, atoi(dg));
  m->deg0 = std::min(m->deg0, m->deg);
  if (const char* dg = getenv("LSAMD_MR_DEGREE"))
    m->deg_mr = std::max(1, std::min(kMgDegMax, atoi(dg)));
  MgLevel L0;
  L0.sys = s;
  L0.g = s->gd;
  L0.n = s->n;
  if (int rc = mg_level_vectors(m, L0)) return rc;
  m->lv.push_back(L0);
  lsamd_grid gprev = s->grid;
  for (int l = 1;; ++l) {
    if (gprev.nnodes[0] <= kMgCoarse && gprev.nnodes[1] <= kMgCoarse &&
        gprev.nnodes[2] <= kMgCoarse)
      break;
    const lsamd_grid gc = mg_coarse_grid(gprev);
    lsamd_fd_reg rc_reg = s->reg;
    const double sc = ldexp(1.0, -l);
    for (double* f : {&rc_reg.w2_xx, &rc_reg.w2_yy, &rc_reg.w2_zz, &rc_reg.w2_xy, &rc_reg.w2_xz,
                      &rc_reg.w2_yz, &rc_reg.w2_bx, &rc_reg.w2_by, &rc_reg.w2_bz, &rc_reg.ridge2})
      *f *= sc;
    lsamd_fd_system* cs = nullptr;
    if (int rc = lsamd_fd_system_create(&gc, &rc_reg, nullptr, &cs)) return rc;
    MgLevel L;
    L.sys = cs;
    L.owned = true;
    L.g = cs->gd;
    L.n = cs->n;
    m->lv.push_back(L);
    if (int rc = mg_level_vectors(m, m->lv.back())) return rc;
    if (int rc = mg_prepare_level(cs, st)) return rc;
    gprev = gc;
  }
  const int64_t nc = m->lv.back().n;
  if (nc > 4096) return fail(LSAMD_ERR_ARG, "multigrid: coarsest level too large");
  m->nc = (int)nc;
  if (int rc = mg_alloc(m, &m->cinv, nc * nc)) return rc;
  if (int rc = mg_alloc(m, &m->rp64, s->n)) return rc;
  LSAMD_HIP(legacy_memset(m->rp64, 0, s->n * 8));
  return LSAMD_OK;
}

// Dense FP64 operator of the coarsest level from its entries (regulariser
// rows and Gram matrix), inverted on the host, stored in FP32.
int mg_coarsest(Mg* m, const CsrOut& G, hipStream_t st) {
  MgLevel& C = m->lv.back();
  const int nc = m->nc;
  std::vector<int64_t> boff = {0, nc}, moff = {0, (int64_t)nc * nc};
  std::vector<int32_t> nodes(nc);
  for (int i = 0; i < nc; ++i) nodes[i] = i;
  int64_t *dboff = nullptr, *dmoff = nullptr;
  int32_t* dnodes = nullptr;
  double* dense = nullptr;
  if (int rc = mg_dalloc(m, &dboff, 2)) return rc;
  if (int rc = mg_dalloc(m, &dmoff, 2)) return rc;
  if (int rc = mg_dalloc(m, &dnodes, nc)) return rc;
  if (int rc = mg_dalloc(m, &dense, (int64_t)nc * nc)) return rc;
  LSAMD_HIP(hipMemcpyAsync(dboff, boff.data(), 16, hipMemcpyHostToDevice, st));
  LSAMD_HIP(hipMemcpyAsync(dmoff, moff.data(), 16, hipMemcpyHostToDevice, st));
  LSAMD_HIP(hipMemcpyAsync(dnodes, nodes.data(), nc * 4, hipMemcpyHostToDevice, st));
  lsamd_fd_system* S = C.sys;
  hipLaunchKernelGGL(k_blk_fill, dim3((unsigned)ceil_div((int64_t)nc * nc, 256)), dim3(256), 0,
                     st, (const int64_t*)dboff, (const int32_t*)dnodes, (const int64_t*)dmoff, 1,
                     (int64_t)nc * nc, C.g, make_regw<double>(S->reg), (const double*)S->scale2,
                     (const int64_t*)G.ptr, (const int32_t*)G.col, (const double*)G.v64, dense);
  LSAMD_CHECK_LAUNCH("multigrid coarsest operator");
  std::vector<double> a((size_t)nc * nc), inv;
  LSAMD_HIP(hipMemcpyAsync(a.data(), dense, a.size() * 8, hipMemcpyDeviceToHost, st));
  LSAMD_HIP(hipStreamSynchronize(st));
  for (void* p : {(void*)dboff, (void*)dmoff, (void*)dnodes, (void*)dense}) mg_give(m, p);
  if (!mg_invert(a, nc, inv)) return fail(LSAMD_ERR_ARG, "multigrid: singular coarsest operator");
  std::vector<float> invf(inv.begin(), inv.end());
  LSAMD_HIP(hipMemcpyAsync(m->cinv, invf.data(), invf.size() * 4, hipMemcpyHostToDevice, st));
  LSAMD_HIP(hipStreamSynchronize(st));
  return LSAMD_OK;
}

// Row-dependent part: coarse data rows (A P per level) and their Gram
// matrices, the coarse diagonals, the Schwarz blocks, the coarsest inverse
// and the smoother polynomials.  Rebuilt after every set_rows.
int mg_build_rows(lsamd_fd_system* s, Mg* m, hipStream_t st) {
  MgTimer tm(st);
  const bool schwarz = m->schwarz;
  const bool single = m->lv.size() == 1;  // level 0 is the coarsest
  // level 0 keeps its data term as rows; a Gram matrix of them is built here
  // for the entries of its Schwarz blocks (or of its dense inverse) and kept
  // for the batched solve
  std::vector<std::vector<int64_t>> cells(1);
  CsrOut G0;
  {
    int wmax = 0;
    if (int rc = mg_data_cells(s, cells[0], &wmax, st)) return rc;
    if (s->nrows > 0 && wmax > 0) {
      int32_t* fc = nullptr;
      double* fcoef = nullptr;
      if (int rc = mg_dalloc(m, &fc, s->nrows * wmax)) return rc;
      if (int rc = mg_dalloc(m, &fcoef, s->nrows * wmax)) return rc;
      hipLaunchKernelGGL(k_rows_fixed, dim3((unsigned)ceil_div(s->nrows, 128)), dim3(128), 0, st,
                         (const int64_t*)s->row_ptr, (const int32_t*)s->cols,
                         (const double*)s->coef64, s->nrows, wmax, fc, fcoef);
      LSAMD_CHECK_LAUNCH("multigrid fixed-width rows");
      const int rc = gram_csr(m, fc, fcoef, s->w2_64, s->nrows, wmax, s->n, &G0, st);
      mg_give(m, fc);
      mg_give(m, fcoef);
      if (rc) return rc;
    }
    tm.tick("level-0 cells + gram");
  }
  if (schwarz && !single) {
    if (int rc = mg_schwarz(m, m->lv[0], cells[0], G0, st)) return rc;
    tm.tick("level-0 schwarz");
  }
  if (single) {
    if (int rc = mg_coarsest(m, G0, st)) return rc;
  }
  m->g0ptr = G0.ptr;
  m->g0col = G0.col;
  m->g0v32 = G0.v32;
  m->g0v64 = G0.v64;
  if (G0.ptr) {
    // compact list of the nodes with Gram entries (the batched solve adds
    // the level-0 data term over these only)
    uint8_t* flag = nullptr;
    int32_t *list = nullptr, *nsel = nullptr;
    if (int rc = mg_dalloc(m, &flag, s->n)) return rc;
    if (int rc = mg_dalloc(m, &list, s->n)) return rc;
    if (int rc = mg_dalloc(m, &nsel, 1)) return rc;
    hipLaunchKernelGGL(k_rows_nonempty, dim3((unsigned)ceil_div(s->n, 256)), dim3(256), 0, st,
                       (const int64_t*)G0.ptr, flag, s->n);
    hipcub::CountingInputIterator<int32_t> it(0);
    size_t tb = 0;
    LSAMD_HIP(hipcub::DeviceSelect::Flagged(nullptr, tb, it, flag, list, nsel, (int)s->n, st));
    void* work = mg_take(m, tb);
    if (!work) return fail(LSAMD_ERR_ALLOC, "hipMalloc failed (multigrid Gram node list)");
    LSAMD_HIP(hipcub::DeviceSelect::Flagged(work, tb, it, flag, list, nsel, (int)s->n, st));
    int32_t h = 0;
    LSAMD_HIP(hipMemcpyAsync(&h, nsel, 4, hipMemcpyDeviceToHost, st));
    LSAMD_HIP(hipStreamSynchronize(st));
    mg_give(m, work);
    mg_give(m, flag);
    mg_give(m, nsel);
    m->g0list = list;
    m->g0n = h;
  }
  // data rows of the previous level: CSR (level 0: the system's rows)
  const int64_t nr = s->nrows;
  const int64_t* prow = s->row_ptr;
  const int32_t* pcols = s->cols;
  const double* pcoef = s->coef64;
  const double* w2 = s->w2_64;
  int64_t* own_row = nullptr;  // restricted rows of the previous level (pooled)
  int32_t* own_cols = nullptr;
  double* own_coef = nullptr;
  int32_t* cnt = nullptr;
  if (nr > 0)
    if (int rc = mg_dalloc(m, &cnt, nr)) return rc;
  std::vector<int32_t> hc(std::max<int64_t>(nr, 1));
  for (size_t l = 1; l < m->lv.size(); ++l) {
    MgLevel& L = m->lv[l];
    lsamd_fd_system* cs = L.sys;
    const GridDev gf = m->lv[l - 1].g, gcd = L.g;
    // regulariser-only diagonal of the coarse system, then its Gram diagonal
    launch_reg_diag(cs, st);
    if (nr > 0) {
      hipLaunchKernelGGL(k_mg_rows, dim3((unsigned)ceil_div(nr, 128)), dim3(128), 0, st, prow,
                         pcols, pcoef, nr, gf, gcd, 1, 0, cnt, (int32_t*)nullptr,
                         (double*)nullptr);
      LSAMD_HIP(hipMemcpyAsync(hc.data(), cnt, nr * 4, hipMemcpyDeviceToHost, st));
      LSAMD_HIP(hipStreamSynchronize(st));
      int wc = 1;
      for (int64_t r = 0; r < nr; ++r) wc = std::max(wc, (int)hc[r]);
      if (wc > 64) return fail(LSAMD_ERR_ARG, "multigrid: a data row spans too many cells");
      int32_t* ccols = nullptr;
      double* ccoef = nullptr;
      int64_t* rptr = nullptr;
      if (int rc = mg_dalloc(m, &ccols, nr * wc)) return rc;
      if (int rc = mg_dalloc(m, &ccoef, nr * wc)) return rc;
      if (int rc = mg_dalloc(m, &rptr, nr + 1)) return rc;
      hipLaunchKernelGGL(k_iota_rows, dim3((unsigned)ceil_div(nr + 1, 256)), dim3(256), 0, st,
                         rptr, nr + 1, (int64_t)wc);
      hipLaunchKernelGGL(k_mg_rows, dim3((unsigned)ceil_div(nr, 128)), dim3(128), 0, st, prow,
                         pcols, pcoef, nr, gf, gcd, 0, wc, (int32_t*)nullptr, ccols, ccoef);
      LSAMD_CHECK_LAUNCH("multigrid rows");
      mg_give(m, own_row);
      mg_give(m, own_cols);
      mg_give(m, own_coef);
      own_row = rptr;
      own_cols = ccols;
      own_coef = ccoef;
      tm.tick("coarse rows");
      if (int rc = mg_gram(m, L, ccols, ccoef, w2, nr, wc, st)) return rc;
      tm.tick("coarse gram");
      prow = rptr;
      pcols = ccols;
      pcoef = ccoef;
    } else {
      hipLaunchKernelGGL(k_minv, dim3((unsigned)ceil_div(L.n, 256)), dim3(256), 0, st, cs->diag,
                         cs->minv64, cs->minv32, L.n);
    }
    if (schwarz) {
      cells.push_back(cells.back());
      mg_coarse_cells(gf, gcd, cells.back());
    }
  }
  mg_give(m, own_row);
  mg_give(m, own_cols);
  mg_give(m, own_coef);
  mg_give(m, cnt);
  // Schwarz blocks of the intermediate levels (the coarsest is solved densely)
  if (schwarz)
    for (size_t l = 1; l + 1 < m->lv.size(); ++l) {
      MgLevel& L = m->lv[l];
      CsrOut G;
      G.ptr = L.gptr;
      G.col = L.gcol;
      G.v32 = L.gval;
      G.v64 = L.gval64;
      if (int rc = mg_schwarz(m, L, cells[l], G, st)) return rc;
      tm.tick("coarse schwarz");
    }
  if (!single) {
    MgLevel& C = m->lv.back();
    CsrOut G;
    G.ptr = C.gptr;
    G.col = C.gcol;
    G.v32 = C.gval;
    G.v64 = C.gval64;
    if (int rc = mg_coarsest(m, G, st)) return rc;
  }
  tm.tick("coarsest inverse");
  // smoother coefficients of the other levels
  for (size_t l = 0; l + 1 < m->lv.size(); ++l) {
    MgLevel& L = m->lv[l];
    if (int rc = mg_lambda(m, L, &L.lambda, st)) return rc;
    tm.tick("lambda");
    const double lmax = 1.15 * L.lambda, lmin = L.lambda / 30.0;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    L.c0 = (float)(1.0 / theta);
    // the Chebyshev recursion does not depend on the final degree: every
    // coefficient up to the largest degree, each smoother takes its prefix
    for (int k = 1; k < kMgDegMax; ++k) {
      const double rn = 1.0 / (2.0 * sigma - rho);
      L.c1[k] = (float)(rn * rho);
      L.c2[k] = (float)(2.0 * rn / delta);
      rho = rn;
    }
  }
  m->stale = false;
  return LSAMD_OK;
}

void mg_vcycle(Mg* m, int l, const CgCtl* ctl, hipStream_t st) {
  MgLevel& L = m->lv[l];
  if (l == (int)m->lv.size() - 1) {
    hipLaunchKernelGGL(k_mg_dense, dim3((unsigned)m->nc), dim3(64), 0, st,
                       (const float*)m->cinv, (const float*)L.b, L.x, m->nc, ctl);
    return;
  }
  const dim3 nb((unsigned)ceil_div(L.n, 256));
  const MgSm S = mg_sm(L);
  const int split = L.cov && L.n >= kSplitMin ? 1 : 0;
  const dim3 ncb((unsigned)std::max<int64_t>(1, ceil_div(L.ncov, 256)));
  auto smooth = [&](int from_zero) {
    hipLaunchKernelGGL(k_mg_cheb_first, nb, dim3(256), 0, st, (const float*)L.b,
                       (const float*)L.q, S, L.x, L.d, L.n, L.c0, from_zero, split, ctl);
    const int deg = l == 0 ? m->deg0 : m->deg;
    if (split && L.ncov > 0)
      hipLaunchKernelGGL(k_mg_cheb_cov, ncb, dim3(256), 0, st, (const int32_t*)L.cov, L.ncov,
                         (const float*)L.b, (const float*)L.q, S, L.x, L.d, from_zero ? 0 : 1,
                         L.c0, 0.f, 0.f, 0, ctl);
    for (int k = 1; k < deg; ++k) {
      const int last = k == deg - 1;
      mg_level_apply(L, L.x, L.q, ctl, st);
      hipLaunchKernelGGL(k_mg_cheb_step, nb, dim3(256), 0, st, (const float*)L.b,
                         (const float*)L.q, S, L.x, L.d, L.n, L.c1[k], L.c2[k], split, last, ctl);
      if (split && L.ncov > 0)
        hipLaunchKernelGGL(k_mg_cheb_cov, ncb, dim3(256), 0, st, (const int32_t*)L.cov, L.ncov,
                           (const float*)L.b, (const float*)L.q, S, L.x, L.d, 2, 0.f, L.c1[k],
                           L.c2[k], last, ctl);
    }
  };
  smooth(1);
  mg_level_apply(L, L.x, L.q, ctl, st);
  MgLevel& C = m->lv[l + 1];
  hipLaunchKernelGGL(k_mg_restrict, dim3((unsigned)ceil_div(C.n, 256)), dim3(256), 0, st,
                     (const float*)L.b, (const float*)L.q, L.g, C.b, C.g, ctl);
  mg_vcycle(m, l + 1, ctl, st);
  hipLaunchKernelGGL(k_mg_prolong, nb, dim3(256), 0, st, (const float*)C.x, C.g, L.x, L.g, ctl);
  mg_level_apply(L, L.x, L.q, ctl, st);
  smooth(0);
}

int launch_fcg_iteration(lsamd_fd_system* s, hipStream_t st) {
  Mg* m = s->mg;
  const int64_t n = s->n;
  const int vb = vec_blocks(n);
  MgLevel& L0 = m->lv[0];
  // L0.b holds (float) r: written by the previous update (or the set-up)
  mg_vcycle(m, 0, s->ctl, st);
  hipLaunchKernelGGL(k_fcg_dots, dim3(vb), dim3(256), 0, st, (const float*)L0.x,
                     (const double*)s->r64, (const double*)m->rp64, n, s->part,
                     (const CgCtl*)s->ctl);
  hipLaunchKernelGGL(k_fcg_dots_fin, dim3(1), dim3(256), 0, st, (const double*)s->part, vb,
                     s->ctl);
  hipLaunchKernelGGL(k_fcg_dir, dim3(vb), dim3(256), 0, st, (const float*)L0.x, s->p64, n,
                     (const CgCtl*)s->ctl);
  if (int rc = launch_apply<double>(s, s->p64, s->q64, s->t64, s->ctl, st)) return rc;
  hipLaunchKernelGGL(k_fcg_upd, dim3(vb), dim3(256), 0, st, s->x64, s->r64, m->rp64,
                     (const double*)s->p64, (const double*)s->q64, (const double*)s->minv64,
                     L0.b, n, s->part, (const CgCtl*)s->ctl);
  hipLaunchKernelGGL(k_fcg_upd_fin, dim3(1), dim3(256), 0, st, (const double*)s->part, vb,
                     s->ctl);
  LSAMD_CHECK_LAUNCH("multigrid PCG iteration");
  return LSAMD_OK;
}

int run_fcg(lsamd_fd_system* s, int chunk, hipStream_t st) {
  Mg* m = s->mg;
  if (m->graph && m->graph_chunk != chunk) {
    hipGraphExecDestroy(m->graph);
    m->graph = nullptr;
  }
  if (!m->graph || m->graph_stale) {
    std::lock_guard<std::recursive_mutex> capture_lock(g_legacy_mu);
    hipStream_t cs;
    LSAMD_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t graph;
    LSAMD_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    int rc = LSAMD_OK;
    for (int i = 0; i < chunk && rc == LSAMD_OK; ++i) rc = launch_fcg_iteration(s, cs);
    hipError_t e = hipStreamEndCapture(cs, &graph);
    hipStreamDestroy(cs);
    if (rc) return rc;
    if (e != hipSuccess) return fail(LSAMD_ERR_HIP, "multigrid graph capture failed");
    bool updated = false;
    if (m->graph) {
      hipGraphNode_t err_node = nullptr;
      hipGraphExecUpdateResult res = hipGraphExecUpdateError;
      updated = hipGraphExecUpdate(m->graph, graph, &err_node, &res) == hipSuccess &&
                res == hipGraphExecUpdateSuccess;
      if (!updated) {
        hipGetLastError();  // a refused update is not an error: instantiate instead
        hipGraphExecDestroy(m->graph);
        m->graph = nullptr;
      }
    }
    if (!updated) e = hipGraphInstantiate(&m->graph, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(LSAMD_ERR_HIP, "multigrid graph instantiate failed");
    m->graph_chunk = chunk;
    m->graph_stale = false;
  }
  for (;;) {
    LSAMD_HIP(hipMemcpyAsync(s->ctl_host, s->ctl, sizeof(CgCtl), hipMemcpyDeviceToHost, st));
    LSAMD_HIP(hipStreamSynchronize(st));
    if (s->ctl_host->done) break;
    LSAMD_HIP(hipGraphLaunch(m->graph, st));
  }
  return LSAMD_OK;
}

}  // namespace

// hierarchy of the system, built or brought up to date with its rows
int mg_ensure(lsamd_fd_system* s, hipStream_t st) {
  if (!s->mg) {
    Mg* m = new Mg();
    if (int rc = mg_build_static(s, m, st)) {
      mg_destroy(m);
      return rc;
    }
    s->mg = m;
  }
  if (s->mg->stale) {
    if (int rc = mg_build_rows(s, s->mg, st)) {
      mg_destroy(s->mg);
      s->mg = nullptr;
      return rc;
    }
  }
  return LSAMD_OK;
}

int mg_solve(lsamd_fd_system* s, const lsamd_solve_opts* o, const double* b, double* c,
             lsamd_solve_info* inf, hipStream_t st) {
  if (s->op != 0) return fail(LSAMD_ERR_ARG, "multigrid is available for FD systems only");
  if (int rc = mg_ensure(s, st)) return rc;
  const int64_t n = s->n;
  const int64_t maxiter64 = o->maxiter > 0 ? o->maxiter : 10 * n;
  const int maxiter = (int)std::min<int64_t>(maxiter64, 2000000000LL);
  const int chunk = std::max(1, std::min(o->check_every > 0 ? o->check_every : 8, 8));
  if (int rc = launch_init<double>(s, b, s->minv64, s->x64, s->r64, o->rtol, o->atol, maxiter, st))
    return rc;
  hipLaunchKernelGGL(k_mg_to_f32, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st,
                     (const double*)s->r64, s->mg->lv[0].b, n, (const CgCtl*)s->ctl);
  if (int rc = run_fcg(s, chunk, st)) return rc;
  hipLaunchKernelGGL(k_copy_d, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, s->x64, c, n);
  LSAMD_HIP(hipStreamSynchronize(st));
  inf->iterations = s->ctl_host->iter;
  inf->outer = 0;
  inf->converged = s->ctl_host->converged;
  inf->residual = s->ctl_host->iter > 0 ? sqrt(s->ctl_host->res2) : sqrt(s->ctl_host->rho);
  inf->threshold = s->ctl_host->thr;
  return LSAMD_OK;
}

// ------------------------------------------- batched multigrid solve ----
// Multigrid-preconditioned flexible CG of many FD systems in one launch
// sequence per iteration (kernels in bmg.hpp).  Each system's hierarchy is
// built or brought up to date by the single-system code; the descriptors,
// block tables and per-system control live in one temporary allocation.
int bmg_solve(lsamd_fd_system* const* ss, int nsys, const lsamd_solve_opts* o,
              double* const* c, lsamd_solve_info* infos, hipStream_t st) {
  for (int i = 0; i < nsys; ++i) {
    lsamd_fd_system* s = ss[i];
    if (!s || s->op != 0 || !c[i])
      return fail(LSAMD_ERR_ARG, "batched multigrid: every system must be an FD system with an "
                                 "output");
    if (int rc = ensure_work(s, false)) return rc;
    if (int rc = mg_ensure(s, st)) return rc;
  }
  const int deg = ss[0]->mg->deg, deg0 = ss[0]->mg->deg0;
  int nl = 0;
  int64_t nmax = 0;
  for (int i = 0; i < nsys; ++i) {
    const Mg* m = ss[i]->mg;
    if (m->deg != deg || m->deg0 != deg0)
      return fail(LSAMD_ERR_ARG, "batched multigrid: hierarchies with different smoothers");
    if (m->nc > kMgCoarse * kMgCoarse * kMgCoarse)
      return fail(LSAMD_ERR_ARG, "batched multigrid: coarsest level too large");
    nl = std::max(nl, (int)m->lv.size());
    nmax = std::max(nmax, ss[i]->n);
  }
  // host descriptors and tables; a system's blocks are contiguous
  std::vector<std::vector<BmgLevel>> hl(nl, std::vector<BmgLevel>(nsys));
  std::vector<std::vector<int2>> all(nl), sm(nl);
  std::vector<std::vector<int>> dense(nl);
  std::vector<BmgSys> hs(nsys);
  for (int i = 0; i < nsys; ++i) {
    lsamd_fd_system* s = ss[i];
    const Mg* m = s->mg;
    const int last = (int)m->lv.size() - 1;
    for (int l = 0; l <= last; ++l) {
      const MgLevel& V = m->lv[l];
      BmgLevel& d = hl[l][i];
      memset(&d, 0, sizeof(d));
      d.A = mg_dev(V);
      if (l == 0) {
        d.A.gptr = m->g0ptr;
        d.A.gcol = m->g0col;
        d.A.gval = m->g0v32;
        d.g0list = m->g0list;
        d.g0n = m->g0list ? m->g0n : 0;
      }
      d.S = mg_sm(V);
      d.b = V.b;
      d.x = V.x;
      d.q = V.q;
      d.d = V.d;
      d.c0 = V.c0;
      for (int k = 0; k < kMgDegMax; ++k) {
        d.c1[k] = V.c1[k];
        d.c2[k] = V.c2[k];
      }
      if (l == last) {
        d.cinv = m->cinv;
        d.nc = m->nc;
        dense[l].push_back(i);
      }
      const int nb = (int)ceil_div(V.n, 256);
      if (l == 0) hs[i].b0 = (int)all[0].size();
      for (int b = 0; b < nb; ++b) {
        all[l].push_back(make_int2(i, b));
        if (l < last) sm[l].push_back(make_int2(i, b));
      }
      if (l == 0) hs[i].b1 = (int)all[0].size();
    }
    BmgSys& y = hs[i];
    y.g = s->gd;
    y.n = s->n;
    y.tab = s->regular ? s->tab64 : nullptr;
    y.w = make_regw<double>(s->reg);
    y.gptr = m->g0ptr;
    y.gcol = m->g0col;
    y.gval = m->g0v64;
    y.x = s->x64;
    y.r = s->r64;
    y.rp = m->rp64;
    y.p = s->p64;
    y.q = s->q64;
    y.minv = s->minv64;
    y.rhs = s->rhs;
    y.out = c[i];
    y.g0list = m->g0list;
    y.g0n = m->g0list ? m->g0n : 0;
    y.gb0 = y.b0;  // node-block partials (the plane-streaming form resets these)
    y.gb1 = y.b1;
  }
  // level 0 by the plane-streaming stencil when every system shares the
  // class table (one regulariser) and is large enough for z-chunks of 16 on
  // 32 x 16 tiles, and level 0 is not its coarsest; the work items of every
  // system in one table, the descriptors in the pointer form (x, y = the
  // system's own level-0 x, q in FP32 and p, q in FP64)
  constexpr int kB0Zc = 16, kB0Tx = 32;
  bool tiles0 = true;
  for (int i = 0; i < nsys && tiles0; ++i) {
    const lsamd_fd_system* s = ss[i];
    tiles0 = s->regular && s->tab32 && s->tab64 && s->mg->lv.size() > 1 &&
             memcmp(&s->reg, &ss[0]->reg, sizeof(lsamd_fd_reg)) == 0 && s->gd.nx >= 16 &&
             s->gd.ny >= 16 && s->gd.nz - 6 >= kB0Zc;
  }
  std::vector<int4> ht;
  std::vector<BSys> hb32, hb64;
  std::vector<int2> hgl;  // blocks of 256 entries of every system's Gram node list
  if (tiles0) {
    hb32.resize(nsys);
    hb64.resize(nsys);
    for (int i = 0; i < nsys; ++i) {
      lsamd_fd_system* s = ss[i];
      hs[i].gb0 = (int)hgl.size();
      for (int64_t b = 0; b < ceil_div(hs[i].g0n, 256); ++b) hgl.push_back(make_int2(i, (int)b));
      hs[i].gb1 = (int)hgl.size();
      hs[i].t0 = (int)ht.size();
      stencil_tiles(i, s->gd, kB0Zc, ht, kB0Tx);
      hs[i].t1 = (int)ht.size();
      for (BSys* b : {&hb32[i], &hb64[i]}) {
        memset(b, 0, sizeof(BSys));
        b->g = s->gd;
        b->n = s->n;
      }
      hb32[i].xp = s->mg->lv[0].x;
      hb32[i].yp = s->mg->lv[0].q;
      hb64[i].xp = s->p64;
      hb64[i].yp = s->q64;
    }
  }
  // one device allocation: descriptors, tables, partials, control, flag
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t bytes = up(nsys * sizeof(BmgSys)) + up(nsys * sizeof(CgCtl)) + up(sizeof(int)) +
                 up(2 * all[0].size() * sizeof(double)) + up(ht.size() * sizeof(int4)) +
                 up(ht.size() * sizeof(double)) + 2 * up(hb32.size() * sizeof(BSys)) +
                 up(hgl.size() * sizeof(int2));
  for (int l = 0; l < nl; ++l)
    bytes += up(nsys * sizeof(BmgLevel)) + up(all[l].size() * sizeof(int2)) +
             up(sm[l].size() * sizeof(int2)) + up(dense[l].size() * sizeof(int));
  unsigned char* base = nullptr;
  if (hipMalloc((void**)&base, bytes) != hipSuccess)
    return fail(LSAMD_ERR_ALLOC, "batched multigrid: hipMalloc failed");
  int* hflag = nullptr;
  if (hipHostMalloc((void**)&hflag, sizeof(int), hipHostMallocDefault) != hipSuccess) {
    hipFree(base);
    return fail(LSAMD_ERR_ALLOC, "batched multigrid: hipHostMalloc failed");
  }
  unsigned char* cur = base;
  auto take = [&](size_t b) {
    unsigned char* p = cur;
    cur += up(b);
    return p;
  };
  hipGraphExec_t exec = nullptr;
  auto finish = [&](int rc) {
    if (exec) hipGraphExecDestroy(exec);
    hipStreamSynchronize(st);
    hipFree(base);
    hipHostFree(hflag);
    return rc;
  };
  BmgSys* dsys = (BmgSys*)take(nsys * sizeof(BmgSys));
  CgCtl* ctl = (CgCtl*)take(nsys * sizeof(CgCtl));
  int* flag = (int*)take(sizeof(int));
  double* part = (double*)take(2 * all[0].size() * sizeof(double));
  int4* dtiles = (int4*)take(ht.size() * sizeof(int4));
  double* tpart = (double*)take(ht.size() * sizeof(double));
  BSys* db32 = (BSys*)take(hb32.size() * sizeof(BSys));
  BSys* db64 = (BSys*)take(hb64.size() * sizeof(BSys));
  int2* dgl = (int2*)take(hgl.size() * sizeof(int2));
  std::vector<BmgLevel*> dlv(nl);
  std::vector<int2*> dall(nl), dsm(nl);
  std::vector<int*> ddense(nl);
  for (int l = 0; l < nl; ++l) {
    dlv[l] = (BmgLevel*)take(nsys * sizeof(BmgLevel));
    dall[l] = (int2*)take(all[l].size() * sizeof(int2));
    dsm[l] = (int2*)take(sm[l].size() * sizeof(int2));
    ddense[l] = (int*)take(dense[l].size() * sizeof(int));
  }
  auto h2d = [&](void* d, const void* h, size_t b) {
    return b ? hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, st) : hipSuccess;
  };
  hipError_t e = h2d(dsys, hs.data(), nsys * sizeof(BmgSys));
  if (e == hipSuccess) e = h2d(dtiles, ht.data(), ht.size() * sizeof(int4));
  if (e == hipSuccess) e = h2d(db32, hb32.data(), hb32.size() * sizeof(BSys));
  if (e == hipSuccess) e = h2d(db64, hb64.data(), hb64.size() * sizeof(BSys));
  if (e == hipSuccess) e = h2d(dgl, hgl.data(), hgl.size() * sizeof(int2));
  for (int l = 0; l < nl && e == hipSuccess; ++l) {
    e = h2d(dlv[l], hl[l].data(), nsys * sizeof(BmgLevel));
    if (e == hipSuccess) e = h2d(dall[l], all[l].data(), all[l].size() * sizeof(int2));
    if (e == hipSuccess) e = h2d(dsm[l], sm[l].data(), sm[l].size() * sizeof(int2));
    if (e == hipSuccess) e = h2d(ddense[l], dense[l].data(), dense[l].size() * sizeof(int));
  }
  if (e != hipSuccess) return finish(fail(LSAMD_ERR_HIP, "batched multigrid: upload failed"));
  const int64_t maxiter64 = o->maxiter > 0 ? o->maxiter : 10 * nmax;
  const int maxiter = (int)std::min<int64_t>(maxiter64, 2000000000LL);
  const int chunk = std::max(1, std::min(o->check_every > 0 ? o->check_every : 8, 8));
  const dim3 b256(256);
  auto g = [](size_t nb) { return dim3((unsigned)nb); };
  const unsigned n0 = (unsigned)all[0].size();
  hipLaunchKernelGGL(k_bmg_init, g(n0), b256, 0, st, (const BmgSys*)dsys, dlv[0],
                     (const int2*)dall[0], part);
  hipLaunchKernelGGL(k_bmg_init_fin, g(nsys), b256, 0, st, (const BmgSys*)dsys,
                     (const double*)part, o->rtol, o->atol, maxiter, ctl);
  hipLaunchKernelGGL(k_bmg_all_done, dim3(1), b256, 0, st, (const CgCtl*)ctl, nsys, flag);
  LSAMD_CHECK_LAUNCH("batched multigrid init");
  // one iteration for every system
  auto cheb = [&](int l, int mode, int k, int last, hipStream_t cs) {
    hipLaunchKernelGGL(k_bmg_cheb, g(sm[l].size()), b256, 0, cs, (const BmgLevel*)dlv[l],
                       (const int2*)dsm[l], mode, k, last, (const CgCtl*)ctl);
  };
  const unsigned nt0 = (unsigned)ht.size();
  auto apply = [&](int l, hipStream_t cs) {
    if (l == 0 && tiles0) {
      hipLaunchKernelGGL((k_b_tiles<float, kB0Zc, false, kB0Tx, true>), dim3(nt0), dim3(NT), 0,
                         cs, (const float*)nullptr, (float*)nullptr, (const BSys*)db32,
                         (const int4*)dtiles, (const float*)ss[0]->tab32, ss[0]->ku32, tpart,
                         (const CgCtl*)ctl);
      if (!hgl.empty())
        hipLaunchKernelGGL(k_bmg_gram0, g(hgl.size()), b256, 0, cs, (const BmgLevel*)dlv[0],
                           (const int2*)dgl, (const CgCtl*)ctl);
      return;
    }
    hipLaunchKernelGGL(k_bmg_apply, g(sm[l].size()), b256, 0, cs, (const BmgLevel*)dlv[l],
                       (const int2*)dsm[l], (const CgCtl*)ctl);
  };
  auto smooth = [&](int l, int from_zero, hipStream_t cs) {
    const int dg = l == 0 ? deg0 : deg;
    cheb(l, from_zero ? 0 : 1, 0, 0, cs);
    for (int k = 1; k < dg; ++k) {
      apply(l, cs);
      cheb(l, 2, k, k == dg - 1, cs);
    }
  };
  auto iteration = [&](hipStream_t cs) {
    for (int l = 0; l < nl; ++l) {
      if (!sm[l].empty()) {
        smooth(l, 1, cs);
        apply(l, cs);
        hipLaunchKernelGGL(k_bmg_restrict, g(all[l + 1].size()), b256, 0, cs,
                           (const BmgLevel*)dlv[l], (const BmgLevel*)dlv[l + 1],
                           (const int2*)dall[l + 1], (const CgCtl*)ctl);
      }
      if (!dense[l].empty())
        hipLaunchKernelGGL(k_bmg_dense, dim3(kMgCoarse * kMgCoarse * kMgCoarse,
                                             (unsigned)dense[l].size()),
                           dim3(64), 0, cs, (const BmgLevel*)dlv[l], (const int*)ddense[l],
                           (const CgCtl*)ctl);
    }
    for (int l = nl - 2; l >= 0; --l) {
      if (sm[l].empty()) continue;
      hipLaunchKernelGGL(k_bmg_prolong, g(sm[l].size()), b256, 0, cs, (const BmgLevel*)dlv[l],
                         (const BmgLevel*)dlv[l + 1], (const int2*)dsm[l], (const CgCtl*)ctl);
      apply(l, cs);
      smooth(l, 0, cs);
    }
    hipLaunchKernelGGL(k_bmg_dots, g(n0), b256, 0, cs, (const BmgSys*)dsys,
                       (const BmgLevel*)dlv[0], (const int2*)dall[0], part, (const CgCtl*)ctl);
    hipLaunchKernelGGL(k_bmg_dots_fin, g(nsys), b256, 0, cs, (const BmgSys*)dsys,
                       (const double*)part, ctl);
    hipLaunchKernelGGL(k_bmg_dir, g(n0), b256, 0, cs, (const BmgSys*)dsys,
                       (const BmgLevel*)dlv[0], (const int2*)dall[0], (const CgCtl*)ctl);
    if (tiles0) {
      hipLaunchKernelGGL((k_b_tiles<double, kB0Zc, false, kB0Tx, true>), dim3(nt0), dim3(NT), 0,
                         cs, (const double*)nullptr, (double*)nullptr, (const BSys*)db64,
                         (const int4*)dtiles, (const double*)ss[0]->tab64, ss[0]->ku64, tpart,
                         (const CgCtl*)ctl);
      if (!hgl.empty())
        hipLaunchKernelGGL(k_bmg_gram64, g(hgl.size()), b256, 0, cs, (const BmgSys*)dsys,
                           (const int2*)dgl, part, (const CgCtl*)ctl);
    } else {
      hipLaunchKernelGGL(k_bmg_apply64, g(n0), b256, 0, cs, (const BmgSys*)dsys,
                         (const int2*)dall[0], part, (const CgCtl*)ctl);
    }
    hipLaunchKernelGGL(k_bmg_alpha, g(nsys), b256, 0, cs, (const BmgSys*)dsys,
                       (const double*)part, (const double*)tpart, ctl);
    hipLaunchKernelGGL(k_bmg_upd, g(n0), b256, 0, cs, (const BmgSys*)dsys,
                       (const BmgLevel*)dlv[0], (const int2*)dall[0], part, (const CgCtl*)ctl);
    hipLaunchKernelGGL(k_bmg_upd_fin, g(nsys), b256, 0, cs, (const BmgSys*)dsys,
                       (const double*)part, ctl);
    hipLaunchKernelGGL(k_bmg_all_done, dim3(1), b256, 0, cs, (const CgCtl*)ctl, nsys, flag);
  };
  {
    std::lock_guard<std::recursive_mutex> capture_lock(g_legacy_mu);
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
      return finish(fail(LSAMD_ERR_HIP, "batched multigrid: stream creation failed"));
    hipGraph_t graph;
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      for (int it = 0; it < chunk; ++it) iteration(cs);
      e = hipStreamEndCapture(cs, &graph);
    }
    hipStreamDestroy(cs);
    if (e != hipSuccess)
      return finish(fail(LSAMD_ERR_HIP, "batched multigrid: graph capture failed"));
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) {
      exec = nullptr;
      return finish(fail(LSAMD_ERR_HIP, "batched multigrid: graph instantiate failed"));
    }
  }
  for (;;) {
    e = hipMemcpyAsync(hflag, flag, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return finish(fail(LSAMD_ERR_HIP, "batched multigrid: poll failed"));
    if (*hflag) break;
    if (hipGraphLaunch(exec, st) != hipSuccess)
      return finish(fail(LSAMD_ERR_HIP, "batched multigrid: graph launch failed"));
  }
  hipLaunchKernelGGL(k_bmg_out, g(n0), b256, 0, st, (const BmgSys*)dsys, (const int2*)dall[0]);
  std::vector<CgCtl> hc(nsys);
  e = hipMemcpyAsync(hc.data(), ctl, nsys * sizeof(CgCtl), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess || hipGetLastError() != hipSuccess)
    return finish(fail(LSAMD_ERR_HIP, "batched multigrid: solve failed"));
  for (int i = 0; i < nsys; ++i) {
    infos[i].iterations = hc[i].iter;
    infos[i].outer = 0;
    infos[i].converged = hc[i].converged;
    infos[i].residual = hc[i].iter > 0 ? sqrt(hc[i].res2) : sqrt(hc[i].rho);
    infos[i].threshold = hc[i].thr;
  }
  return finish(LSAMD_OK);
}

// ------------------------------------------------- multi-rhs multigrid ----
struct MrWork {
  int R = 0;
  std::vector<float*> b, x, q, d;  // per level, R lane vectors one after the other
  // outer flexible CG: x in FP64; the search direction p, N p and the
  // residual (b[0], the V-cycle's input) in FP32, the residual replaced by
  // the FP64 rhs - N x at the start of every captured chunk (q64 = N x)
  double *rhs = nullptr, *x64 = nullptr, *q64 = nullptr;
  float *p32 = nullptr, *q32 = nullptr;
  float* t32 = nullptr;
  double* t64 = nullptr;
  double* upart = nullptr;  // per-block partials of the row part of p.Np, per lane
  int64_t t_rows = -1;
  double* part = nullptr;
  MrCtl* ctl = nullptr;
  int* all_done = nullptr;
  MrCtl* host = nullptr;  // pinned: R controls and the all-done flag
  // level-0 stencil over the R lanes as R systems of the plane-streaming
  // kernel: system table, work items per precision (FP64, FP32), zero
  // controls (no lane skipped), per-work-item partials of x.Rx (the FP64
  // ones give the stencil part of p.Np)
  BSys* sysR = nullptr;
  int4* tilesR[2] = {nullptr, nullptr};
  int ntR[2] = {0, 0};
  int zcR[2] = {0, 0};  // stencil form per precision (kMrForm code)
  CgCtl* ctlR = nullptr;
  double* partR = nullptr;
  // coarse levels on regular grids (index = level; entries of level 0 and of
  // irregular levels unused): the same kernel over the R lanes with the
  // level's class table, the Gram rows added for all lanes after it
  std::vector<BSys*> lsys;
  std::vector<int4*> ltiles;
  std::vector<int> lnt, lzc;  // work items and stencil form (kMrForm code) per level
  // per level with a covered list: interleaved r over the list and the
  // lane-major M r of the listed nodes (R floats per covered node each),
  // capacity in floats
  std::vector<float*> rint, zcov;
  std::vector<int64_t> rint_cap;
  std::vector<void*> mem;
  hipGraphExec_t graph = nullptr;
  int graph_chunk = 0;
};

void mr_invalidate(MrWork* w) {
  if (w && w->graph) {
    hipGraphExecDestroy(w->graph);
    w->graph = nullptr;
  }
}

void mr_destroy(MrWork* w) {
  if (!w) return;
  mr_invalidate(w);
  for (void* p : w->mem) hipFree(p);
  if (w->host) hipHostFree(w->host);
  delete w;
}

namespace {

template <typename P>
int mr_alloc(MrWork* w, P** p, int64_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, (size_t)std::max<int64_t>(count, 1) * sizeof(P)) != hipSuccess)
    return fail(LSAMD_ERR_ALLOC, "hipMalloc failed (multi-rhs multigrid)");
  w->mem.push_back(q);
  *p = (P*)q;
  return LSAMD_OK;
}

// stencil forms of the multi-rhs levels: z-chunk 8/16/32 on 64 x 8 tiles,
// or the same chunks on 32 x 16 tiles (code 2000 + chunk)
constexpr int kMrNForm = 6;
constexpr int kMrForm[kMrNForm] = {8, 16, 32, 2008, 2016, 2032};
inline int mr_form_zc(int form) { return form % 1000; }
inline int mr_form_txt(int form) { return form >= 2000 ? 32 : 64; }

template <typename T>
void mr_form_launch(int form, unsigned nt, hipStream_t st, const T* x, T* y, const BSys* sys,
                    const int4* tiles, const T* tab, const Coef25<T>& ku, double* part,
                    const CgCtl* ctl) {
  const dim3 g(nt), b(NT);
  switch (form) {
    case 8:
      hipLaunchKernelGGL((k_b_tiles<T, 8>), g, b, 0, st, x, y, sys, tiles, tab, ku, part, ctl);
      break;
    case 16:
      hipLaunchKernelGGL((k_b_tiles<T, 16>), g, b, 0, st, x, y, sys, tiles, tab, ku, part, ctl);
      break;
    case 32:
      hipLaunchKernelGGL((k_b_tiles<T, 32>), g, b, 0, st, x, y, sys, tiles, tab, ku, part, ctl);
      break;
    case 2008:
      hipLaunchKernelGGL((k_b_tiles<T, 8, false, 32>), g, b, 0, st, x, y, sys, tiles, tab, ku,
                         part, ctl);
      break;
    case 2016:
      hipLaunchKernelGGL((k_b_tiles<T, 16, false, 32>), g, b, 0, st, x, y, sys, tiles, tab, ku,
                         part, ctl);
      break;
    default:
      hipLaunchKernelGGL((k_b_tiles<T, 32, false, 32>), g, b, 0, st, x, y, sys, tiles, tab, ku,
                         part, ctl);
      break;
  }
}

// lane vectors, controls and the level-0 stencil plan of R lanes for the
// system's hierarchy
template <typename T>
void mr_tiles(lsamd_fd_system* s, MrWork* w, const T* x, T* y, hipStream_t st,
              const CgCtl* ctl = nullptr);
void mr_level_tiles(lsamd_fd_system* s, MrWork* w, int l, const float* x, float* y,
                    hipStream_t st);

int mr_build(lsamd_fd_system* s, MrWork* w, int R) {
  const Mg* m = s->mg;
  const int64_t n = s->n;
  w->R = R;
  for (const MgLevel& L : m->lv) {
    float *b, *x, *q, *d;
    for (float** v : {&b, &x, &q, &d}) {
      if (int rc = mr_alloc(w, v, L.n * R)) return rc;
      LSAMD_HIP(legacy_memset(*v, 0, L.n * R * 4));
    }
    w->b.push_back(b);
    w->x.push_back(x);
    w->q.push_back(q);
    w->d.push_back(d);
  }
  if (int rc = injected("multi_rhs_work")) return rc;
  for (double** v : {&w->rhs, &w->x64, &w->q64}) {
    if (int rc = mr_alloc(w, v, n * R)) return rc;
    LSAMD_HIP(legacy_memset(*v, 0, n * R * 8));
  }
  for (float** v : {&w->p32, &w->q32}) {
    if (int rc = mr_alloc(w, v, n * R)) return rc;
    LSAMD_HIP(legacy_memset(*v, 0, n * R * 4));
  }
  if (int rc = mr_alloc(w, &w->part, (int64_t)vec_blocks(n) * 2 * R)) return rc;
  if (int rc = mr_alloc(w, &w->ctl, R)) return rc;
  if (int rc = mr_alloc(w, &w->all_done, 1)) return rc;
  if (hipHostMalloc((void**)&w->host, (R + 1) * sizeof(MrCtl), hipHostMallocDefault) !=
      hipSuccess) {
    w->host = nullptr;
    return fail(LSAMD_ERR_ALLOC, "hipHostMalloc failed (multi-rhs multigrid)");
  }
  // R-system plan of the level-0 stencil
  std::vector<BSys> hs(R);
  memset(hs.data(), 0, R * sizeof(BSys));
  for (int r = 0; r < R; ++r) {
    hs[r].g = s->gd;
    hs[r].n = n;
    hs[r].off = (int64_t)r * n;
  }
  if (int rc = mr_alloc(w, &w->sysR, R)) return rc;
  LSAMD_HIP(legacy_memcpy(w->sysR, hs.data(), R * sizeof(BSys), hipMemcpyHostToDevice));
  // work-item tables of every stencil form; the R lanes make level 0
  // HBM-bound (unlike one cache-resident system), so the form is timed here
  // per precision rather than taken from the single-system tuning, and each
  // coarse level times its own (their axis lengths 2^k + 1 fit the two tile
  // widths differently)
  const char* pin = getenv("LSAMD_STENCIL_ZC");
  const int pinned = pin ? atoi(pin) : 0;
  auto eligible = [&](int form, const GridDev& g) {
    if (pinned == 8 || pinned == 16 || pinned == 32) return mr_form_zc(form) == pinned;
    return mr_form_zc(form) == 8 || g.nz - 6 >= mr_form_zc(form);
  };
  int4* tc[kMrNForm] = {};
  int ntc[kMrNForm] = {};
  int64_t maxt = 1;
  for (int c = 0; c < kMrNForm; ++c) {
    if (!eligible(kMrForm[c], s->gd)) continue;
    std::vector<int4> t;
    for (int r = 0; r < R; ++r) stencil_tiles(r, s->gd, mr_form_zc(kMrForm[c]), t, mr_form_txt(kMrForm[c]));
    if (int rc = mr_alloc(w, &tc[c], (int64_t)t.size())) return rc;
    LSAMD_HIP(legacy_memcpy(tc[c], t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice));
    ntc[c] = (int)t.size();
    maxt = std::max<int64_t>(maxt, (int64_t)t.size());
  }
  // coarse-level tables of every eligible form (the coarsest level is solved
  // densely); the best is kept after timing
  const size_t nl = m->lv.size();
  w->lsys.assign(nl, nullptr);
  w->ltiles.assign(nl, nullptr);
  w->lnt.assign(nl, 0);
  w->lzc.assign(nl, 0);
  std::vector<std::vector<int4*>> lt(nl, std::vector<int4*>(kMrNForm, nullptr));
  std::vector<std::vector<int>> lnt(nl, std::vector<int>(kMrNForm, 0));
  for (size_t l = 1; l + 1 < nl; ++l) {
    const MgLevel& L = m->lv[l];
    if (!L.sys->regular) continue;
    std::vector<BSys> ls(R);
    memset(ls.data(), 0, R * sizeof(BSys));
    for (int r = 0; r < R; ++r) {
      ls[r].g = L.g;
      ls[r].n = L.n;
      ls[r].off = (int64_t)r * L.n;
    }
    if (int rc = mr_alloc(w, &w->lsys[l], R)) return rc;
    LSAMD_HIP(legacy_memcpy(w->lsys[l], ls.data(), R * sizeof(BSys), hipMemcpyHostToDevice));
    for (int c = 0; c < kMrNForm; ++c) {
      if (!eligible(kMrForm[c], L.g)) continue;
      std::vector<int4> t;
      for (int r = 0; r < R; ++r) stencil_tiles(r, L.g, mr_form_zc(kMrForm[c]), t, mr_form_txt(kMrForm[c]));
      if (int rc = mr_alloc(w, &lt[l][c], (int64_t)t.size())) return rc;
      LSAMD_HIP(legacy_memcpy(lt[l][c], t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice));
      lnt[l][c] = (int)t.size();
      maxt = std::max<int64_t>(maxt, (int64_t)t.size());
    }
  }
  if (int rc = mr_alloc(w, &w->ctlR, R)) return rc;
  LSAMD_HIP(legacy_memset(w->ctlR, 0, R * sizeof(CgCtl)));
  if (int rc = mr_alloc(w, &w->partR, maxt)) return rc;
  hipEvent_t e0, e1;
  LSAMD_HIP(hipEventCreate(&e0));
  LSAMD_HIP(hipEventCreate(&e1));
  // fastest of the forms with tables: run() launches the candidate set up
  // by use(c)
  auto tune = [&](int4* const* tabs, auto use, auto run) -> int {
    float best = 1e30f;
    int bc = -1;
    for (int round = 0; round < 2; ++round)
      for (int c = 0; c < kMrNForm; ++c) {
        if (!tabs[c]) continue;
        use(c);
        run();
        LSAMD_HIP(hipEventRecord(e0, 0));
        for (int k = 0; k < 3; ++k) run();
        LSAMD_HIP(hipEventRecord(e1, 0));
        LSAMD_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        LSAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
          best = ms;
          bc = c;
        }
      }
    return bc;
  };
  for (int p = 0; p < 2; ++p) {
    auto use = [&](int c) {
      w->tilesR[p] = tc[c];
      w->ntR[p] = ntc[c];
      w->zcR[p] = kMrForm[c];
    };
    const int bc = tune(tc, use, [&]() {
      if (p)
        mr_tiles<float>(s, w, (const float*)w->x[0], w->q[0], 0);
      else
        mr_tiles<double>(s, w, (const double*)w->x64, w->q64, 0);
    });
    if (bc < 0) return fail(LSAMD_ERR_ARG, "multi-rhs: no stencil form fits the grid");
    use(bc);
  }
  for (size_t l = 1; l + 1 < nl; ++l) {
    if (!w->lsys[l]) continue;
    auto use = [&](int c) {
      w->ltiles[l] = lt[l][c];
      w->lnt[l] = lnt[l][c];
      w->lzc[l] = kMrForm[c];
    };
    const int bc = tune(lt[l].data(), use,
                        [&]() { mr_level_tiles(s, w, (int)l, w->x[l], w->q[l], 0); });
    if (bc < 0) return fail(LSAMD_ERR_ARG, "multi-rhs: no stencil form fits a coarse grid");
    use(bc);
  }
  if (getenv("LSAMD_MG_TIMING")) {
    fprintf(stderr, "[multi-rhs stencil] R %d level 0 forms fp64 %d fp32 %d;", R, w->zcR[0],
            w->zcR[1]);
    for (size_t l = 1; l + 1 < nl; ++l) fprintf(stderr, " level %zu %d", l, w->lzc[l]);
    fprintf(stderr, "\n");
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  LSAMD_CHECK_LAUNCH("multi-rhs stencil tuning");
  return LSAMD_OK;
}

// per-lane row buffers, sized by the system's current rows
int mr_rows(lsamd_fd_system* s, MrWork* w) {
  if (w->t_rows == s->nrows) return LSAMD_OK;
  const int R = w->R;
  mr_invalidate(w);
  for (void* p : {(void*)w->t32, (void*)w->t64, (void*)w->upart}) {
    if (!p) continue;
    for (size_t i = 0; i < w->mem.size(); ++i)
      if (w->mem[i] == p) {
        hipFree(p);
        w->mem.erase(w->mem.begin() + i);
        break;
      }
  }
  w->t32 = nullptr;
  w->t64 = nullptr;
  w->upart = nullptr;
  w->t_rows = -1;
  if (int rc = mr_alloc(w, &w->t32, std::max<int64_t>(s->nrows, 1) * R)) return rc;
  if (int rc = mr_alloc(w, &w->t64, std::max<int64_t>(s->nrows, 1) * R)) return rc;
  if (int rc = mr_alloc(w, &w->upart, std::max<int64_t>(ceil_div(s->nrows, 256), 1) * R))
    return rc;
  w->t_rows = s->nrows;
  return LSAMD_OK;
}

// work of R lanes (R = 4, 8, 16 or 32), one per width, kept on the handle
// only when complete: a failed build or row resize frees it
int mr_ensure(lsamd_fd_system* s, int R, MrWork** out) {
  const int slot = R == 4 ? 0 : R == 8 ? 1 : R == 16 ? 2 : R == 32 ? 3 : 4;
  MrWork* w = s->mr[slot];
  if (!w) {
    w = new MrWork();
    if (int rc = mr_build(s, w, R)) {
      mr_destroy(w);
      return rc;
    }
    s->mr[slot] = w;
  }
  if (int rc = mr_rows(s, w)) {
    mr_destroy(w);
    s->mr[slot] = nullptr;
    return rc;
  }
  *out = w;
  return LSAMD_OK;
}

// level-0 regulariser over the R lanes with the plane-streaming kernel (ctl:
// per-lane control blocks, the group's own by default)
template <typename T>
void mr_tiles(lsamd_fd_system* s, MrWork* w, const T* x, T* y, hipStream_t st,
              const CgCtl* ctl) {
  const int p = sizeof(T) == 4 ? 1 : 0;
  if (!ctl) ctl = (const CgCtl*)w->ctlR;
  const T* tab = p ? (const T*)s->tab32 : (const T*)s->tab64;
  Coef25<T> ku;
  for (int q = 0; q < kTaps; ++q) ku.k[q] = p ? (T)s->ku32.k[q] : (T)s->ku64.k[q];
  mr_form_launch<T>(w->zcR[p], (unsigned)w->ntR[p], st, x, y, (const BSys*)w->sysR,
                    (const int4*)w->tilesR[p], tab, ku, w->partR, ctl);
}

// regulariser of coarse level l over the R lanes (its tuned form)
void mr_level_tiles(lsamd_fd_system* s, MrWork* w, int l, const float* x, float* y,
                    hipStream_t st) {
  const lsamd_fd_system* ls = s->mg->lv[l].sys;
  mr_form_launch<float>(w->lzc[l], (unsigned)w->lnt[l], st, x, y, (const BSys*)w->lsys[l],
                        (const int4*)w->ltiles[l], (const float*)ls->tab32, ls->ku32, w->partR,
                        (const CgCtl*)w->ctlR);
}

// coarse levels of at least this many nodes add their Gram term four lanes
// per thread, smaller ones one lane per thread
constexpr int64_t kMrGramGroup = 65536;

// sliced smoother and Gram rows of every level (see MgLevel::ssl, gsl)
int mr_sell_ensure(lsamd_fd_system* s, hipStream_t st) {
  Mg* m = s->mg;
  if (m->sell_ready) return LSAMD_OK;
  const bool verbose = getenv("LSAMD_MG_TIMING") != nullptr;
  for (size_t l = 0; l + 1 < m->lv.size(); ++l) {
    MgLevel& L = m->lv[l];
    // levels with a covered list: the covered rows over list positions;
    // the others: every node's row
    if (L.sptr && !L.cov)
      if (int rc = sell_build(m, L.sptr, L.scol, L.sval, L.sm, nullptr, L.n, &L.ssl, st)) return rc;
    if (L.sptr && L.cov && L.ncov > 0) {
      if (int rc = sell_build(m, L.sptr, L.scol, L.sval, nullptr, L.cov, L.ncov, &L.csl, st))
        return rc;
      int32_t* pos = nullptr;
      int* bad = nullptr;
      if (int rc = mg_dalloc(m, &pos, L.n)) return rc;
      if (int rc = mg_dalloc(m, &bad, 1)) return rc;
      LSAMD_HIP(hipMemsetAsync(pos, 0xff, L.n * 4, st));
      LSAMD_HIP(hipMemsetAsync(bad, 0, 4, st));
      hipLaunchKernelGGL(k_cov_pos, dim3((unsigned)ceil_div(L.ncov, 256)), dim3(256), 0, st,
                         (const int32_t*)L.cov, L.ncov, pos);
      hipLaunchKernelGGL(k_sell_to_pos, dim3((unsigned)ceil_div(L.csl.nent, 256)), dim3(256), 0,
                         st, L.csl.col, L.csl.nent, (const int32_t*)pos, bad);
      LSAMD_CHECK_LAUNCH("covered rows over list positions");
      int hbad = 0;
      LSAMD_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, st));
      LSAMD_HIP(hipStreamSynchronize(st));
      mg_give(m, bad);
      if (hbad) return fail(LSAMD_ERR_ARG, "multigrid: a covered row reaches an uncovered node");
      L.cpos = pos;
    }
    if (l > 0 && L.gptr)
      if (int rc = sell_build(m, L.gptr, L.gcol, L.gval, nullptr, nullptr, L.n, &L.gsl, st))
        return rc;
    if (verbose)
      fprintf(stderr, "[multigrid sliced rows] level %zu: n %lld, smoother rows %lld (%lld entries), "
              "covered rows %lld (%lld entries), gram rows %lld (%lld entries)\n", l,
              (long long)L.n, (long long)L.ssl.nrows, (long long)L.ssl.nent,
              (long long)L.csl.nrows, (long long)L.csl.nent, (long long)L.gsl.nrows,
              (long long)L.gsl.nent);
  }
  m->sell_ready = true;
  return LSAMD_OK;
}

// interleaved r buffers of the levels with covered rows, sized for the
// current covered lists (a larger list reallocates and drops the captured
// iteration, which holds the old addresses)
int mr_rint_ensure(lsamd_fd_system* s, MrWork* w) {
  const Mg* m = s->mg;
  w->rint.resize(m->lv.size(), nullptr);
  w->zcov.resize(m->lv.size(), nullptr);
  w->rint_cap.resize(m->lv.size(), 0);
  for (size_t l = 0; l < m->lv.size(); ++l) {
    const int64_t need = m->lv[l].csl.nrows * w->R;
    if (need == 0 || w->rint_cap[l] >= need) continue;
    mr_invalidate(w);
    for (float** v : {&w->rint[l], &w->zcov[l]}) {
      if (*v) {
        for (size_t i = 0; i < w->mem.size(); ++i)
          if (w->mem[i] == (void*)*v) {
            hipFree(*v);
            w->mem.erase(w->mem.begin() + i);
            break;
          }
      }
      *v = nullptr;
    }
    w->rint_cap[l] = 0;
    if (int rc = mr_alloc(w, &w->rint[l], need)) return rc;
    if (int rc = mr_alloc(w, &w->zcov[l], need)) return rc;
    w->rint_cap[l] = need;
  }
  return LSAMD_OK;
}

// lanes per thread of a sliced-row kernel over `rows` rows (environment
// override: 1, 2, 4 or 8); at most R
int mr_sell_lanes(const char* env, int64_t rows, int R, int64_t big) {
  const char* e = getenv(env);
  int k = e ? atoi(e) : (rows >= big ? 4 : 1);
  if (k != 1 && k != 2 && k != 4 && k != 8) k = 1;
  return std::min(k, R);
}

template <int R>
void mr_apply(lsamd_fd_system* s, MrWork* w, int l, const float* x, float* y, hipStream_t st) {
  const Mg* m = s->mg;
  const MgLevel& L = m->lv[l];
  const int* ad = w->all_done;
  const dim3 nb((unsigned)ceil_div(L.n, 256), R), b256(256);
  if (l == 0) {
    RowsDev Rw = rows_dev(s, true);
    if (s->nrows > 0)
      hipLaunchKernelGGL(k_mr_rows_t<float>, dim3((unsigned)ceil_div(s->nrows, 256), R), b256, 0,
                         st, x, L.n, Rw, w->t32, R, ad);
    mr_tiles<float>(s, w, x, y, st);
    if (s->nrows > 0)
      hipLaunchKernelGGL(k_mr_rows_scatter<float>,
                         dim3((unsigned)std::max<int64_t>(1, ceil_div(s->ndn, 256)), R), b256, 0,
                         st, Rw, (const float*)w->t32, y, L.n, R, ad);
    return;
  }
  if (l < (int)w->ltiles.size() && w->ltiles[l]) {
    mr_level_tiles(s, w, l, x, y, st);
    if (L.gsl.nrows > 0) {
      const Sell G = sell_dev(L.gsl);
      const int lg = mr_sell_lanes("LSAMD_MR_GRAM_LG", L.gsl.nrows, R, 32768);
      const int xcd = 1;  // XCD-contiguous workgroups (measured faster)
      const dim3 gg((unsigned)ceil_div(L.gsl.nrows, 256), R / lg);
      switch (lg) {
        case 1:
          hipLaunchKernelGGL(k_mr_gram_sell<1>, gg, b256, 0, st, G, L.n, x, y, ad, xcd);
          break;
        case 2:
          hipLaunchKernelGGL(k_mr_gram_sell<2>, gg, b256, 0, st, G, L.n, x, y, ad, xcd);
          break;
        case 4:
          hipLaunchKernelGGL(k_mr_gram_sell<4>, gg, b256, 0, st, G, L.n, x, y, ad, xcd);
          break;
        default:
          hipLaunchKernelGGL(k_mr_gram_sell<8>, gg, b256, 0, st, G, L.n, x, y, ad, xcd);
          break;
      }
    } else if (L.gptr) {
      if (L.n >= kMrGramGroup)
        hipLaunchKernelGGL(k_mr_gram_add<4>, dim3((unsigned)ceil_div(L.n, 256), R / 4), b256, 0,
                           st, (const int64_t*)L.gptr, (const int32_t*)L.gcol,
                           (const float*)L.gval, L.n, x, y, ad);
      else
        hipLaunchKernelGGL(k_mr_gram_add<1>, dim3((unsigned)ceil_div(L.n, 256), R), b256, 0, st,
                           (const int64_t*)L.gptr, (const int32_t*)L.gcol, (const float*)L.gval,
                           L.n, x, y, ad);
    }
    return;
  }
  if (L.n >= kMrNodeThreads)
    hipLaunchKernelGGL(k_mr_apply_coarse_all<R>, dim3((unsigned)ceil_div(L.n, 256)), b256, 0, st,
                       mg_dev(L), x, y, ad);
  else
    hipLaunchKernelGGL(k_mr_apply_coarse, nb, b256, 0, st, mg_dev(L), x, y, R, ad);
}

// lanes per thread of the covered-node smoothing step (LSAMD_MR_COV_LANES:
// 1, 2, 4 or 8, at most R; default 2, the fastest at 128^3 with 16 lanes)
int mr_cov_lanes(int R) {
  static const int v = [] {
    const char* e = getenv("LSAMD_MR_COV_LANES");
    const int k = e ? atoi(e) : 2;
    return k == 1 || k == 2 || k == 4 || k == 8 ? k : 2;
  }();
  return std::min(v, R);
}

template <int R>
void mr_vcycle(lsamd_fd_system* s, MrWork* w, int l, hipStream_t st) {
  const Mg* m = s->mg;
  const MgLevel& L = m->lv[l];
  const int* ad = w->all_done;
  const dim3 b256(256);
  if (l == (int)m->lv.size() - 1) {
    hipLaunchKernelGGL(k_mr_dense, dim3((unsigned)ceil_div((int64_t)m->nc * R, 256)), b256, 0, st,
                       (const float*)m->cinv, (const float*)w->b[l], w->x[l], m->nc, R, ad);
    return;
  }
  const dim3 nb((unsigned)ceil_div(L.n, 256), R);
  const MgSm S = mg_sm(L);
  const int split = L.cov ? 1 : 0;
  // XCD-contiguous workgroups where measured faster: the sliced smoother
  // and the restriction; the covered-node step and the prolongation keep the
  // hardware order
  const int xcd = 1;
  const dim3 ncb((unsigned)std::max<int64_t>(1, ceil_div(L.ncov, 256)), R);
  // the smoother rows of every node in slices (levels without a covered list)
  auto step_sell = [&](int mode, float c1, float c2) {
    const Sell P = sell_dev(L.ssl);
    const int lg = mr_sell_lanes("LSAMD_MR_SM_LG", L.ssl.nrows, R, ((int64_t)1) << 40);
    const dim3 gg((unsigned)ceil_div(L.ssl.nrows, 256), R / lg);
    const float *bl = w->b[l], *ql = w->q[l];
    float *xl = w->x[l], *dl = w->d[l];
    switch (lg) {
      case 1:
        hipLaunchKernelGGL(k_mr_cheb_sell<1>, gg, b256, 0, st, P, bl, ql, xl, dl, L.n, mode, L.c0,
                           c1, c2, ad, xcd);
        break;
      case 2:
        hipLaunchKernelGGL(k_mr_cheb_sell<2>, gg, b256, 0, st, P, bl, ql, xl, dl, L.n, mode, L.c0,
                           c1, c2, ad, xcd);
        break;
      case 4:
        hipLaunchKernelGGL(k_mr_cheb_sell<4>, gg, b256, 0, st, P, bl, ql, xl, dl, L.n, mode, L.c0,
                           c1, c2, ad, xcd);
        break;
      default:
        hipLaunchKernelGGL(k_mr_cheb_sell<8>, gg, b256, 0, st, P, bl, ql, xl, dl, L.n, mode, L.c0,
                           c1, c2, ad, xcd);
        break;
    }
  };
  auto step = [&](int mode, float c1, float c2) {
    if (!split && L.ssl.nrows > 0) {
      step_sell(mode, c1, c2);
      return;
    }
    if (split && L.csl.nrows > 0 && l < (int)w->rint.size() && w->rint[l]) {
      // covered rows from the interleaved r, then one pass over every node
      hipLaunchKernelGGL(k_mr_cov_pack<R>, dim3((unsigned)ceil_div(L.ncov, 64)), b256, 0, st,
                         (const int32_t*)L.cov, L.ncov, (const float*)w->b[l],
                         (const float*)w->q[l], L.n, (mode & 3) == 0 ? 1 : 0, w->rint[l], ad);
      constexpr int LG = R < 16 ? R : 16;
      hipLaunchKernelGGL((k_mr_cov_z<R, LG>), dim3((unsigned)ceil_div(L.csl.nrows, 256), R / LG),
                         b256, 0, st, sell_dev(L.csl), (const float*)w->rint[l], w->zcov[l], ad);
      constexpr int LA = 4;  // lanes per thread of the all-node pass
      hipLaunchKernelGGL(k_mr_cheb_all<LA>, dim3((unsigned)ceil_div(L.n, 256), R / LA), b256, 0, st,
                         (const float*)w->b[l], (const float*)w->q[l], (const float*)L.sm,
                         (const int32_t*)L.cpos, (const float*)w->zcov[l], L.ncov, w->x[l],
                         w->d[l], L.n, mode, L.c0, c1, c2, ad);
      return;
    }
    hipLaunchKernelGGL(k_mr_cheb, nb, b256, 0, st, (const float*)w->b[l], (const float*)w->q[l], S,
                       w->x[l], w->d[l], L.n, split, mode, L.c0, c1, c2, ad);
    if (split && L.ncov > 0) {
      const int lg = mr_cov_lanes(R);
      const dim3 ncg((unsigned)std::max<int64_t>(1, ceil_div(L.ncov, 256)), R / lg);
      const int32_t* cv = (const int32_t*)L.cov;
      const float *bl = w->b[l], *ql = w->q[l];
      switch (lg) {
        case 1:
          hipLaunchKernelGGL(k_mr_cheb_cov, ncb, b256, 0, st, cv, L.ncov, bl, ql, S, w->x[l],
                             w->d[l], L.n, mode, L.c0, c1, c2, ad);
          break;
        case 2:
          hipLaunchKernelGGL(k_mr_cheb_cov_lg<2>, ncg, b256, 0, st, cv, L.ncov, bl, ql, S,
                             w->x[l], w->d[l], L.n, mode, L.c0, c1, c2, ad, 0);
          break;
        case 8:
          hipLaunchKernelGGL(k_mr_cheb_cov_lg<8>, ncg, b256, 0, st, cv, L.ncov, bl, ql, S,
                             w->x[l], w->d[l], L.n, mode, L.c0, c1, c2, ad, 0);
          break;
        default:
          hipLaunchKernelGGL(k_mr_cheb_cov_lg<4>, ncg, b256, 0, st, cv, L.ncov, bl, ql, S,
                             w->x[l], w->d[l], L.n, mode, L.c0, c1, c2, ad, 0);
          break;
      }
    }
  };
  auto smooth = [&](int from_zero) {
    const int deg = l == 0 ? m->deg0 : m->deg_mr;
    step((from_zero ? 0 : 1) | (deg == 1 ? kMrLast : 0), 0.f, 0.f);
    for (int k = 1; k < deg; ++k) {
      mr_apply<R>(s, w, l, w->x[l], w->q[l], st);
      step(2 | (k == deg - 1 ? kMrLast : 0), L.c1[k], L.c2[k]);
    }
  };
  smooth(1);
  mr_apply<R>(s, w, l, w->x[l], w->q[l], st);
  const MgLevel& C = m->lv[l + 1];
  constexpr int TG = 4;  // lanes per thread of the transfers
  static_assert(R % TG == 0, "lane groups");
  if (C.n < kMrNodeThreads) {
    hipLaunchKernelGGL(k_mr_restrict_lg<TG>, dim3((unsigned)ceil_div(C.n, 256), R / TG), b256, 0,
                       st, (const float*)w->b[l], (const float*)w->q[l], L.g, w->b[l + 1], C.g,
                       ad, xcd);
  } else {
    // z-streamed restriction on large levels (small ones lack the work
    // items): 256-position bands of the coarse plane, kRz coarse planes per
    // work item
    constexpr int kRz = 4;
    const int64_t ncp = (int64_t)C.g.nx * C.g.ny;
    const int nband = (int)ceil_div(ncp, 256);
    const int nkc = (int)ceil_div(C.g.nz, kRz);
    hipLaunchKernelGGL((k_mr_restrict_z<TG, kRz>), dim3((unsigned)(nband * nkc), R / TG), b256, 0,
                       st, (const float*)w->b[l], (const float*)w->q[l], L.g, w->b[l + 1], C.g,
                       nband, ad);
  }
  mr_vcycle<R>(s, w, l + 1, st);
  if (L.n < kMrNodeThreads) {
    hipLaunchKernelGGL(k_mr_prolong_lg<TG>, dim3((unsigned)ceil_div(L.n, 256), R / TG), b256, 0,
                       st, (const float*)w->x[l + 1], C.g, w->x[l], L.g, ad, 0);
  } else {
    // z-streamed prolongation on large levels: 256-position bands of the
    // fine plane, kPz fine planes per work item
    constexpr int kPz = 8;
    const int64_t fnxy = (int64_t)L.g.nx * L.g.ny;
    const int nband = (int)ceil_div(fnxy, 256);
    const int nkc = (int)ceil_div(L.g.nz, kPz);
    hipLaunchKernelGGL((k_mr_prolong_z<TG, kPz>), dim3((unsigned)(nband * nkc), R / TG), b256, 0,
                       st, (const float*)w->x[l + 1], C.g, w->x[l], L.g, nband, ad);
  }
  mr_apply<R>(s, w, l, w->x[l], w->q[l], st);
  smooth(0);
}

// r = (float) (rhs - N x) with N x applied in FP64, and the stopping test
// on it (see k_mr_resid)
template <int R>
void mr_replace_residual(lsamd_fd_system* s, MrWork* w, hipStream_t st) {
  const int64_t n = s->n;
  const int vb = vec_blocks(n);
  const int* ad = w->all_done;
  const dim3 b256(256), fin(mr_fin_threads(R));
  RowsDev Rw = rows_dev(s, false);
  if (s->nrows > 0)
    hipLaunchKernelGGL(k_mr_rows_t<double>, dim3((unsigned)ceil_div(s->nrows, 256), R), b256, 0,
                       st, (const double*)w->x64, n, Rw, w->t64, R, ad);
  mr_tiles<double>(s, w, (const double*)w->x64, w->q64, st);
  if (s->nrows > 0)
    hipLaunchKernelGGL(k_mr_rows_scatter<double>,
                       dim3((unsigned)std::max<int64_t>(1, ceil_div(s->ndn, 256)), R), b256, 0, st,
                       Rw, (const double*)w->t64, w->q64, n, R, ad);
  const dim3 vbg(vb, R / mr_outer_lanes(R));
  hipLaunchKernelGGL(k_mr_resid<R>, vbg, b256, 0, st, (const double*)w->rhs,
                     (const double*)w->q64, (const double*)s->minv64, w->b[0], n, w->part, ad);
  hipLaunchKernelGGL(k_mr_resid_fin<R>, dim3(1), fin, 0, st, (const double*)w->part, vb, w->ctl,
                     w->all_done);
}

template <int R>
void mr_iteration(lsamd_fd_system* s, MrWork* w, hipStream_t st) {
  const int64_t n = s->n;
  const int vb = vec_blocks(n);
  const int* ad = w->all_done;
  const dim3 b256(256), fin(mr_fin_threads(R));
  mr_vcycle<R>(s, w, 0, st);  // from w->b[0] = r
  const dim3 vbg(vb, R / mr_outer_lanes(R));  // lane groups of the vector kernels
  hipLaunchKernelGGL(k_mr_dots<R>, vbg, b256, 0, st, (const float*)w->x[0],
                     (const float*)w->b[0], (const float*)w->q32, n, w->part, ad);
  hipLaunchKernelGGL(k_mr_dots_fin<R>, dim3(1), fin, 0, st, (const double*)w->part, vb, w->ctl,
                     ad);
  hipLaunchKernelGGL(k_mr_dir<R>, vbg, b256, 0, st, (const float*)w->x[0], w->p32, n,
                     (const MrCtl*)w->ctl, ad);
  // q = N p in FP32; p.q from the stencil's per-work-item partials and the
  // rows' per-block partials (both summed in FP64)
  RowsDev Rw = rows_dev(s, true);
  const int nbu = s->nrows > 0 ? (int)ceil_div(s->nrows, 256) : 0;
  if (s->nrows > 0)
    hipLaunchKernelGGL(k_mr_rows_tu<float>, dim3((unsigned)nbu, R), b256, 0, st,
                       (const float*)w->p32, n, Rw, w->t32, w->upart, ad);
  mr_tiles<float>(s, w, (const float*)w->p32, w->q32, st);
  if (s->nrows > 0)
    hipLaunchKernelGGL(k_mr_rows_scatter<float>,
                       dim3((unsigned)std::max<int64_t>(1, ceil_div(s->ndn, 256)), R), b256, 0, st,
                       Rw, (const float*)w->t32, w->q32, n, R, ad);
  hipLaunchKernelGGL(k_mr_alpha_st<R>, dim3(1), fin, 0, st, (const double*)w->partR,
                     w->ntR[1] / R, (const double*)w->upart, nbu, w->ctl, ad);
  hipLaunchKernelGGL(k_mr_upd<R>, vbg, b256, 0, st, w->x64, (const float*)w->p32,
                     (const float*)w->q32, (const double*)s->minv64, w->b[0], n, w->part,
                     (const MrCtl*)w->ctl, ad);
  hipLaunchKernelGGL(k_mr_upd_fin<R>, dim3(1), fin, 0, st, (const double*)w->part, vb, w->ctl,
                     w->all_done);
}

template <int R>
int mr_solve_group(lsamd_fd_system* s, MrWork* w, const lsamd_solve_opts* o, int nr,
                   const double* rhs, double* c, lsamd_solve_info* infos, hipStream_t st) {
  const int64_t n = s->n;
  const int vb = vec_blocks(n);
  const int64_t maxiter64 = o->maxiter > 0 ? o->maxiter : 10 * n;
  const int maxiter = (int)std::min<int64_t>(maxiter64, 2000000000LL);
  const int chunk = std::max(1, std::min(o->check_every > 0 ? o->check_every : 8, 8));
  hipLaunchKernelGGL(k_mr_pack, dim3((unsigned)ceil_div(n * R, 256)), dim3(256), 0, st, rhs, nr,
                     R, n, w->rhs);
  hipLaunchKernelGGL(k_mr_init<R>, dim3(vb), dim3(256), 0, st, (const double*)w->rhs,
                     (const double*)s->minv64, w->x64, w->b[0], n, w->part);
  hipLaunchKernelGGL(k_mr_init_fin<R>, dim3(1), dim3(mr_fin_threads(R)), 0, st, (const double*)w->part, vb,
                     nr, o->rtol, o->atol, maxiter, w->ctl, w->all_done);
  LSAMD_CHECK_LAUNCH("multi-rhs init");
  if (w->graph && w->graph_chunk != chunk) mr_invalidate(w);
  if (!w->graph) {
    std::lock_guard<std::recursive_mutex> capture_lock(g_legacy_mu);
    hipStream_t cs;
    LSAMD_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t graph;
    LSAMD_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    // each chunk starts from the true residual (at the first chunk: the rhs)
    mr_replace_residual<R>(s, w, cs);
    for (int i = 0; i < chunk; ++i) mr_iteration<R>(s, w, cs);
    hipError_t e = hipStreamEndCapture(cs, &graph);
    hipStreamDestroy(cs);
    if (e != hipSuccess) return fail(LSAMD_ERR_HIP, "multi-rhs graph capture failed");
    e = hipGraphInstantiate(&w->graph, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(LSAMD_ERR_HIP, "multi-rhs graph instantiate failed");
    w->graph_chunk = chunk;
  }
  int* hdone = (int*)(w->host + R);
  for (;;) {
    LSAMD_HIP(hipMemcpyAsync(hdone, w->all_done, sizeof(int), hipMemcpyDeviceToHost, st));
    LSAMD_HIP(hipStreamSynchronize(st));
    if (*hdone) break;
    LSAMD_HIP(hipGraphLaunch(w->graph, st));
  }
  LSAMD_HIP(hipMemcpyAsync(c, w->x64, (size_t)n * nr * 8, hipMemcpyDeviceToDevice, st));
  LSAMD_HIP(hipMemcpyAsync(w->host, w->ctl, R * sizeof(MrCtl), hipMemcpyDeviceToHost, st));
  LSAMD_HIP(hipStreamSynchronize(st));
  for (int r = 0; r < nr; ++r) {
    const MrCtl& h = w->host[r];
    infos[r].iterations = h.iter;
    infos[r].outer = 0;
    infos[r].converged = h.converged;
    infos[r].pad = 0;
    infos[r].residual = sqrt(h.res2);
    infos[r].threshold = h.thr;
  }
  return LSAMD_OK;
}

}  // namespace

// Several right-hand sides of the system's normal equations; multigrid with
// all lanes per launch when every level has a class table, else one
// multigrid solve per right-hand side.
int mr_solve(lsamd_fd_system* s, const lsamd_solve_opts* o, int nrhs, const double* rhs,
             double* c, lsamd_solve_info* infos, hipStream_t st) {
  if (s->op != 0) return fail(LSAMD_ERR_ARG, "multi-rhs multigrid: FD systems only");
  if (int rc = mg_ensure(s, st)) return rc;
  if (int rc = mr_sell_ensure(s, st)) return rc;
  const Mg* m = s->mg;
  bool lanes = s->regular && s->tab32 && s->tab64;
  for (size_t l = 1; l + 1 < m->lv.size(); ++l) lanes = lanes && m->lv[l].sys->regular;
  const int64_t n = s->n;
  if (!lanes) {
    for (int r = 0; r < nrhs; ++r)
      if (int rc = mg_solve(s, o, rhs + (int64_t)r * n, c + (int64_t)r * n, &infos[r], st))
        return rc;
    return LSAMD_OK;
  }
  for (int r0 = 0; r0 < nrhs;) {
    const int left = nrhs - r0;
    // lanes per group: 32 from 24 left, 64 from 48 left (an idle lane costs
    // as much as a busy one)
    const int R = left <= 4 ? 4 : (left <= 8 ? 8 : (left < 24 ? 16 : (left < 48 ? 32 : 64)));
    const int nr = std::min(left, R);
    MrWork* w = nullptr;
    if (int rc = mr_ensure(s, R, &w)) return rc;
    if (int rc = mr_rint_ensure(s, w)) return rc;
    const double* rr = rhs + (int64_t)r0 * n;
    double* cc = c + (int64_t)r0 * n;
    const int rc = R == 4    ? mr_solve_group<4>(s, w, o, nr, rr, cc, infos + r0, st)
                   : R == 8  ? mr_solve_group<8>(s, w, o, nr, rr, cc, infos + r0, st)
                   : R == 16 ? mr_solve_group<16>(s, w, o, nr, rr, cc, infos + r0, st)
                   : R == 32 ? mr_solve_group<32>(s, w, o, nr, rr, cc, infos + r0, st)
                             : mr_solve_group<64>(s, w, o, nr, rr, cc, infos + r0, st);
    if (rc) return rc;
    r0 += nr;
  }
  return LSAMD_OK;
}


// ------------------------------------------------------------------ LSMR ----
struct Ls {
  double *v = nullptr, *h = nullptr, *hbar = nullptr, *x = nullptr, *xs = nullptr, *y = nullptr;
  double *ureg = nullptr, *urid = nullptr, *ud = nullptr, *tr = nullptr;
  double *pa = nullptr, *pb = nullptr, *pv = nullptr, *px = nullptr;  // partials
  int na = 0, nb = 0, nv = 0;
  LsCtl* ctl = nullptr;
  LsCtl* host = nullptr;
  std::vector<void*> mem;
  hipGraphExec_t graph = nullptr;
  int graph_chunk = 0;
  const double* graph_scale = nullptr;  // column scaling the graph was captured with
};

void ls_destroy(Ls* l) {
  if (!l) return;
  if (l->graph) hipGraphExecDestroy(l->graph);
  for (void* p : l->mem) hipFree(p);
  if (l->host) hipHostFree(l->host);
  delete l;
}

namespace {

template <typename P>
int ls_alloc(Ls* l, P** p, int64_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, std::max<int64_t>(count, 1) * sizeof(P)) != hipSuccess)
    return fail(LSAMD_ERR_ALLOC, "hipMalloc failed (LSMR)");
  l->mem.push_back(q);
  *p = (P*)q;
  return LSAMD_OK;
}

int ls_build(lsamd_fd_system* s, Ls* l) {
  const int64_t n = s->n, m = std::max<int64_t>(s->nrows, 1);
  for (double** p : {&l->v, &l->h, &l->hbar, &l->x, &l->xs, &l->y, &l->urid})
    if (int rc = ls_alloc(l, p, n)) return rc;
  if (int rc = ls_alloc(l, &l->ureg, 6 * n)) return rc;
  if (int rc = ls_alloc(l, &l->ud, m)) return rc;
  if (int rc = ls_alloc(l, &l->tr, m)) return rc;
  l->na = (int)ceil_div(n, 256);
  l->nb = (int)ceil_div(m, 256);
  l->nv = vec_blocks(n);
  if (int rc = ls_alloc(l, &l->pa, l->na)) return rc;
  if (int rc = ls_alloc(l, &l->pb, l->nb)) return rc;
  if (int rc = ls_alloc(l, &l->pv, l->nv)) return rc;
  if (int rc = ls_alloc(l, &l->px, l->nv)) return rc;
  if (int rc = ls_alloc(l, &l->ctl, 1)) return rc;
  if (hipHostMalloc((void**)&l->host, sizeof(LsCtl), hipHostMallocDefault) != hipSuccess)
    return fail(LSAMD_ERR_ALLOC, "hipHostMalloc failed (LSMR)");
  for (double* p : {l->v, l->h, l->hbar, l->x, l->xs, l->y, l->urid})
    LSAMD_HIP(legacy_memset(p, 0, n * 8));
  LSAMD_HIP(legacy_memset(l->ureg, 0, 6 * n * 8));
  LSAMD_HIP(legacy_memset(l->ud, 0, m * 8));
  LSAMD_HIP(legacy_memset(l->tr, 0, m * 8));
  return LSAMD_OK;
}

// u <- A S xs - fwd_coef u ; beta ; v <- S A^T u / beta - beta v ; partials
// of |v|^2.  init: u = b instead of the product, v = S A^T u / beta.
// sc: squared column scaling (diag(A^T A)^-1, or ones)
void ls_bidiag(lsamd_fd_system* s, Ls* l, const double* sc, int init, hipStream_t st) {
  const int64_t n = s->n;
  const RegW<double> w = make_regw<double>(s->reg);
  RowsDev R = rows_dev(s, false);
  const dim3 nb((unsigned)l->na), b256(256);
  if (!init)
    hipLaunchKernelGGL(k_ls_scale, nb, b256, 0, st, (const double*)l->v,
                       sc, l->xs, n, (const LsCtl*)l->ctl);
  hipLaunchKernelGGL(k_ls_reg_fwd, nb, b256, 0, st, (const double*)l->xs, s->gd, w,
                     (const double*)s->scale2, l->ureg, l->urid, n, init, l->pa,
                     (const LsCtl*)l->ctl);
  int nbp = 0;
  if (s->nrows > 0) {
    nbp = l->nb;
    hipLaunchKernelGGL(k_ls_data_fwd, dim3((unsigned)l->nb), b256, 0, st, (const double*)l->xs,
                       R, (const double*)s->w, (const double*)s->b, l->ud, init, l->pb,
                       (const LsCtl*)l->ctl);
  }
  hipLaunchKernelGGL(k_ls_beta, dim3(1), b256, 0, st, (const double*)l->pa, l->na,
                     (const double*)l->pb, nbp, l->ctl, init);
  hipLaunchKernelGGL(k_ls_reg_t, nb, b256, 0, st, (const double*)l->ureg, (const double*)l->urid,
                     s->gd, w, (const double*)s->scale2, l->y, n, (const LsCtl*)l->ctl);
  if (s->nrows > 0) {
    hipLaunchKernelGGL(k_ls_data_t, dim3((unsigned)l->nb), b256, 0, st, (const double*)l->ud,
                       (const double*)s->w, l->tr, s->nrows, (const LsCtl*)l->ctl);
    hipLaunchKernelGGL((k_rows_scatter<double, double>),
                       dim3((unsigned)std::max<int64_t>(1, ceil_div(s->ndn, 256))), b256, 0, st, R,
                       (const double*)l->tr, l->y, 1, (CgCtl*)nullptr, 0, (const double*)nullptr,
                       0, (const double*)nullptr, 0);
  }
  hipLaunchKernelGGL(k_ls_vupd, dim3((unsigned)l->nv), b256, 0, st, (const double*)l->y,
                     sc, l->v, n, init, l->pv, (const LsCtl*)l->ctl);
  hipLaunchKernelGGL(k_ls_step, dim3(1), b256, 0, st, (const double*)l->pv, l->nv, l->ctl, init);
  hipLaunchKernelGGL(k_ls_xupd, dim3((unsigned)l->nv), b256, 0, st, l->v, l->h, l->hbar, l->x, n,
                     init, l->px, (const LsCtl*)l->ctl);
  if (!init)
    hipLaunchKernelGGL(k_ls_test, dim3(1), b256, 0, st, (const double*)l->px, l->nv, l->ctl);
}

__global__ void k_ls_setup(LsCtl* c, double atol, double btol, double ctol, int maxiter) {
  LsCtl z = {};
  z.atol = atol;
  z.btol = btol;
  z.ctol = ctol;
  z.maxiter = maxiter;
  *c = z;
}

}  // namespace

int ls_solve(lsamd_fd_system* s, const lsamd_lsmr_opts* o, double* c, lsamd_lsmr_info* inf,
             hipStream_t st) {
  if (s->op != 0) return fail(LSAMD_ERR_ARG, "LSMR is available for FD systems only");
  if (!s->ls) {
    Ls* l = new Ls();
    if (int rc = ls_build(s, l)) {
      ls_destroy(l);
      return rc;
    }
    s->ls = l;
  }
  Ls* l = s->ls;
  const int64_t n = s->n;
  const int64_t maxiter64 = o->maxiter > 0 ? o->maxiter : n;
  const int maxiter = (int)std::min<int64_t>(maxiter64, 2000000000LL);
  const int chunk = o->check_every > 0 ? o->check_every : 64;
  const double ctol = o->conlim > 0 ? 1.0 / o->conlim : 0.0;
  hipLaunchKernelGGL(k_ls_setup, dim3(1), dim3(1), 0, st, l->ctl, o->atol, o->btol, ctol, maxiter);
  const double* sc = s->minv64;
  if (!o->scaling) {
    if (!s->ones64) {
      if (int rc = dalloc(s, &s->ones64, n)) return rc;
      if (int rc = dalloc(s, &s->ones32, n)) return rc;
      hipLaunchKernelGGL(k_fill_one, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st,
                         s->ones64, s->ones32, n);
    }
    sc = s->ones64;
  }
  hipLaunchKernelGGL(k_zero_d, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, l->x, n);
  ls_bidiag(s, l, sc, 1, st);
  LSAMD_CHECK_LAUNCH("lsmr init");
  if (l->graph && (l->graph_chunk != chunk || l->graph_scale != sc)) {
    hipGraphExecDestroy(l->graph);
    l->graph = nullptr;
  }
  if (!l->graph) {
    std::lock_guard<std::recursive_mutex> capture_lock(g_legacy_mu);
    hipStream_t cs;
    LSAMD_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t graph;
    LSAMD_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < chunk; ++i) ls_bidiag(s, l, sc, 0, cs);
    hipError_t e = hipStreamEndCapture(cs, &graph);
    hipStreamDestroy(cs);
    if (e != hipSuccess) return fail(LSAMD_ERR_HIP, "graph capture failed (LSMR)");
    e = hipGraphInstantiate(&l->graph, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(LSAMD_ERR_HIP, "graph instantiate failed (LSMR)");
    l->graph_chunk = chunk;
    l->graph_scale = sc;
  }
  for (;;) {
    LSAMD_HIP(hipMemcpyAsync(l->host, l->ctl, sizeof(LsCtl), hipMemcpyDeviceToHost, st));
    LSAMD_HIP(hipStreamSynchronize(st));
    if (l->host->done) break;
    LSAMD_HIP(hipGraphLaunch(l->graph, st));
  }
  hipLaunchKernelGGL(k_ls_out, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st,
                     (const double*)l->x, sc, c, n);
  LSAMD_CHECK_LAUNCH("lsmr out");
  LSAMD_HIP(hipStreamSynchronize(st));
  const LsCtl* h = l->host;
  inf->iterations = h->itn;
  inf->istop = h->istop;
  inf->normr = h->normr;
  inf->normar = h->normar;
  inf->normA = h->normA;
  inf->condA = h->condA;
  inf->normx = h->normx;
  return LSAMD_OK;
}

// Scratch of lsamd_fd_evaluate_adjoint_noalloc: the per-entry arrays of up to
// LSAMD_NOALLOC_MAX_POINTS points (24 bytes per entry, 8 entries per point:
// 12 MiB) and the radix sort's temporary storage.  One copy per GPU, held
// until the process ends.  Host code never passes the name to a kernel: it
// takes the device address from hipGetSymbolAddress, per device.
constexpr size_t kScratchBytes = (size_t)24 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
unsigned char* g_scratch_addr[kMaxDevices] = {};
}  // namespace

// Device address of the current device's scratch, looked up once per device.
int scratch_base(unsigned char** out) {
  int dev = 0;
  LSAMD_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(LSAMD_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(LSAMD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(e));
    g_scratch_addr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_addr[dev];
  return LSAMD_OK;
}

// Bytes of caller storage eval_adjoint_into needs for n points on grid g:
// the per-entry arrays (8 entries per point: two int32 keys, two int32 ids,
// a double) at 256-byte alignment, and the radix sort's temporary storage.
size_t eval_adjoint_bytes(const lsamd_grid* g, int64_t n) {
  if (n <= 0) return 0;
  const int64_t ne = 8 * n;
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const int64_t nn = g->nnodes[0] * g->nnodes[1] * g->nnodes[2];
  int bits = 1;
  while (bits < 31 && ((int64_t)1 << bits) <= nn) ++bits;
  size_t tb = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (int32_t*)nullptr, (int32_t*)nullptr,
                                     (int32_t*)nullptr, (int32_t*)nullptr, (int)ne, 0, bits,
                                     (hipStream_t)0);
  return 4 * up(ne * 4) + up(ne * 8) + up(tb);
}

// out = E^T gv (the adjoint of trilinear evaluation) with every temporary in
// the caller's storage [base, base + room): no allocation, no synchronisation.
// Every buffer taken from the storage is written in full before it is read
// (the entries by k_eval_entries, the sorted copies and the sort's own
// storage by the radix sort), so earlier contents are never seen.
int eval_adjoint_into(const lsamd_grid* g, const double* pts, int64_t n, const double* gv,
                      double* out, unsigned char* base, size_t room, hipStream_t st,
                      const char* fn) {
  if (int rc = check_grid(g)) return rc;
  if (n < 0 || !out || (n > 0 && (!pts || !gv)))
    return fail(LSAMD_ERR_ARG, std::string(fn) + ": bad arguments");
  const int64_t nn = g->nnodes[0] * g->nnodes[1] * g->nnodes[2];
  if (nn > 2147483647LL) return fail(LSAMD_ERR_ARG, "grid exceeds int32 node ids");
  if (8 * n >= 2147483647LL) return fail(LSAMD_ERR_ARG, std::string(fn) + ": too many points");
  if (eval_adjoint_bytes(g, n) > room)
    return fail(LSAMD_ERR_ARG, std::string(fn) + ": workspace smaller than " +
                                   std::to_string(eval_adjoint_bytes(g, n)) + " bytes");
  hipLaunchKernelGGL(k_zero_d, dim3((unsigned)ceil_div(nn, 256)), dim3(256), 0, st, out, nn);
  LSAMD_CHECK_LAUNCH("evaluate_adjoint zero");
  if (n == 0) return LSAMD_OK;
  const int64_t ne = 8 * n;
  auto take = [&](size_t bytes) {
    unsigned char* p = base;
    base += (bytes + 255) & ~(size_t)255;
    return p;
  };
  int32_t* key = (int32_t*)take(ne * 4);
  int32_t* key_s = (int32_t*)take(ne * 4);
  int32_t* idx = (int32_t*)take(ne * 4);
  int32_t* idx_s = (int32_t*)take(ne * 4);
  double* val = (double*)take(ne * 8);
  int bits = 1;
  while (bits < 31 && ((int64_t)1 << bits) <= nn) ++bits;
  size_t tb = 0;
  if (hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key, key_s, idx, idx_s, (int)ne, 0, bits,
                                         st) != hipSuccess)
    return fail(LSAMD_ERR_HIP, std::string(fn) + ": sort size query failed");
  launch_eval_entries(to_dev(*g), pts, gv, n, key, idx, val, st);
  if (hipcub::DeviceRadixSort::SortPairs(base, tb, key, key_s, idx, idx_s, (int)ne, 0, bits,
                                         st) != hipSuccess)
    return fail(LSAMD_ERR_HIP, std::string(fn) + ": sort failed");
  hipLaunchKernelGGL(k_eval_segsum, dim3((unsigned)ceil_div(ne, 256)), dim3(256), 0, st, key_s,
                     idx_s, val, ne, out);
  LSAMD_CHECK_LAUNCH(fn);
  return LSAMD_OK;
}

}  // namespace lsamd

// lsamd_fd_system_time_multi for a group of R lanes
template <int R>
int time_multi(lsamd_fd_system* s, MrWork* w, int32_t kernel, int32_t iters, double* avg_ms,
               double* bytes, hipStream_t st) {
  if (!s->tzero) {  // no lane done, all-done flag clear (room for the widest group)
    if (int rc = dmalloc(s, &s->tzero, 64 * sizeof(CgCtl))) return rc;
    LSAMD_HIP(legacy_memset(s->tzero, 0, 64 * sizeof(CgCtl)));
  }
  const int* zflag = (const int*)s->tzero;
  const MgLevel& L = s->mg->lv[0];
  const int64_t n = s->n;
  const int split = L.cov ? 1 : 0;
  // the all-node step (covered nodes read M r from the list) where the
  // level has one, else the split step over the uncovered nodes
  const bool all = split && L.cpos && w->zcov.size() > 0 && w->zcov[0];
  const int64_t nu = split ? n - L.ncov : n;  // uncovered nodes
  auto launch = [&]() {
    if (kernel == 0 && all)
      hipLaunchKernelGGL(k_mr_cheb_all<4>, dim3((unsigned)ceil_div(n, 256), R / 4), dim3(256), 0,
                         st, (const float*)w->b[0], (const float*)w->q[0], (const float*)L.sm,
                         (const int32_t*)L.cpos, (const float*)w->zcov[0], L.ncov, w->x[0],
                         w->d[0], n, 2, L.c0, L.c1[1], L.c2[1], zflag);
    else if (kernel == 0)
      hipLaunchKernelGGL(k_mr_cheb, dim3((unsigned)ceil_div(n, 256), R), dim3(256), 0, st,
                         (const float*)w->b[0], (const float*)w->q[0], mg_sm(L), w->x[0],
                         w->d[0], n, split, 2, L.c0, L.c1[1], L.c2[1], zflag);
    else if (kernel == 1)
      hipLaunchKernelGGL(k_mr_upd<R>, dim3(vec_blocks(n), R / mr_outer_lanes(R)), dim3(256), 0,
                         st, w->x64, (const float*)w->p32, (const float*)w->q32,
                         (const double*)s->minv64, w->b[0], n, w->part, (const MrCtl*)w->ctl,
                         zflag);
    else
      mr_tiles<float>(s, w, (const float*)w->x[0], w->q[0], st, (const CgCtl*)s->tzero);
  };
  // algorithmic bytes: every input and output element once
  // (all-node step: 24 B per uncovered node-lane, 20 B per covered one (M r,
  // d, x in; d, x out), the tags and the covered nodes' list positions)
  *bytes = kernel == 0 && all ? 24.0 * R * (double)nu + 20.0 * R * (double)L.ncov +
                                    4.0 * (double)n + 4.0 * (double)L.ncov
           : kernel == 0      ? 24.0 * R * (double)nu + 4.0 * (double)n
           : kernel == 1 ? 32.0 * R * (double)n + 8.0 * (double)n
                         : 8.0 * R * (double)n;
  hipEvent_t e0, e1;
  LSAMD_HIP(hipEventCreate(&e0));
  LSAMD_HIP(hipEventCreate(&e1));
  launch();
  LSAMD_HIP(hipEventRecord(e0, st));
  for (int it = 0; it < iters; ++it) launch();
  LSAMD_HIP(hipEventRecord(e1, st));
  LSAMD_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  LSAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  LSAMD_CHECK_LAUNCH("time_multi");
  *avg_ms = ms / iters;
  return LSAMD_OK;
}

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
    LSAMD_HIP(legacy_memcpy(s->scale2, scale2, s->n * sizeof(double), hipMemcpyDefault));
  }
  s->regular = scale2 == nullptr && s->gd.nx >= 7 && s->gd.ny >= 7 && s->gd.nz >= 7;
  if (s->regular) {
    s->nbx = (int)ceil_div(s->gd.nx, TX);
    s->nby = (int)ceil_div(s->gd.ny, TY);
    BSys one;
    memset(&one, 0, sizeof(one));
    one.g = s->gd;
    one.n = s->n;
    if (int rc = dalloc(s, &s->sys1, 1)) return bail(rc);
    LSAMD_HIP(legacy_memcpy(s->sys1, &one, sizeof(BSys), hipMemcpyHostToDevice));
    for (int c = 0; c < 6; ++c) {
      std::vector<int4> t;
      stencil_tiles(0, s->gd, kZcCand[c % 3], t, c < 3 ? 64 : 32);
      s->nstile[c] = (int)t.size();
      if (int rc = dalloc(s, &s->stiles[c], t.size())) return bail(rc);
      LSAMD_HIP(legacy_memcpy(s->stiles[c], t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    std::vector<double> tab(kClasses * kTaps);
    build_class_table(s->reg, tab.data());
    std::vector<float> tabf(tab.begin(), tab.end());
    if (int rc = dalloc(s, &s->tab64, kClasses * kTaps)) return bail(rc);
    if (int rc = dalloc(s, &s->tab32, kClasses * kTaps)) return bail(rc);
    LSAMD_HIP(legacy_memcpy(s->tab64, tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
    LSAMD_HIP(legacy_memcpy(s->tab32, tabf.data(), tabf.size() * 4, hipMemcpyHostToDevice));
    const int bulk = ((3 * 7 + 3) * 7 + 3) * kTaps;
    for (int q = 0; q < kTaps; ++q) {
      s->ku64.k[q] = tab[bulk + q];
      s->ku32.k[q] = (float)tab[bulk + q];
    }
  }
  if (int rc = init_common(s)) return bail(rc);
  *out = s;
  return LSAMD_OK;
}

int lsamd_p1_system_create(const lsamd_grid* g, double cgw, double ridge2,
                           lsamd_fd_system** out) {
  if (int rc = injected("p1_system_create")) return rc;
  if (!out) return fail(LSAMD_ERR_ARG, "lsamd_p1_system_create: NULL argument");
  *out = nullptr;
  if (int rc = check_uniform_grid(g, "lsamd_p1_system_create")) return rc;
  if (!(cgw >= 0.0) || !(ridge2 >= 0.0))
    return fail(LSAMD_ERR_ARG, "lsamd_p1_system_create: cgw and ridge2 must be >= 0");
  lsamd_fd_system* s = new lsamd_fd_system();
  s->grid = *g;
  s->gd = to_dev(*g);
  memset(&s->reg, 0, sizeof(s->reg));
  s->n = g->nnodes[0] * g->nnodes[1] * g->nnodes[2];
  s->op = 1;
  s->p1_ridge2 = ridge2;
  auto bail = [&](int rc) {
    lsamd_fd_system_destroy(s);
    return rc;
  };
  std::vector<double> tab(343 * kP1Taps);
  build_p1_table(*g, cgw * cgw, tab.data(), s->p1nv);
  const int64_t nt = (int64_t)s->p1nv[0] * s->p1nv[1] * s->p1nv[2] * kP1Taps;
  std::vector<float> tabf(tab.begin(), tab.begin() + nt);
  if (int rc = dalloc(s, &s->p1tab64, nt)) return bail(rc);
  if (int rc = dalloc(s, &s->p1tab32, nt)) return bail(rc);
  LSAMD_HIP(legacy_memcpy(s->p1tab64, tab.data(), nt * 8, hipMemcpyHostToDevice));
  LSAMD_HIP(legacy_memcpy(s->p1tab32, tabf.data(), nt * 4, hipMemcpyHostToDevice));
  if (int rc = init_common(s)) return bail(rc);
  *out = s;
  return LSAMD_OK;
}

int lsamd_fd_system_destroy(lsamd_fd_system* s) {
  if (!s) return LSAMD_OK;
  for (int i = 0; i < 2; ++i) {
    if (s->graph[i]) hipGraphExecDestroy(s->graph[i]);
    if (s->bgraph[i]) hipGraphExecDestroy(s->bgraph[i]);
  }
  if (s->mg) mg_destroy(s->mg);
  if (s->ls) ls_destroy(s->ls);
  for (lsamd::MrWork* w : s->mr) mr_destroy(w);
  for (void* p : s->owned) hipFree(p);
  if (s->ctl_host) hipHostFree(s->ctl_host);
  if (s->bctl_host) hipHostFree(s->bctl_host);
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
  if (s->mg) mg_invalidate(s->mg);
  for (lsamd::MrWork* w : s->mr) mr_invalidate(w);
  if (s->ls) {
    ls_destroy(s->ls);
    s->ls = nullptr;
  }
  for (int i = 0; i < 2; ++i) {
    if (s->graph[i]) {
      hipGraphExecDestroy(s->graph[i]);
      s->graph[i] = nullptr;
    }
    if (s->bgraph[i]) {
      hipGraphExecDestroy(s->bgraph[i]);
      s->bgraph[i] = nullptr;
    }
  }
  // the previous row set's buffers stay in s->rpool for this one
  s->u_rows = nullptr;
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
  launch_reg_diag(s, st);
  hipLaunchKernelGGL(k_zero_d, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, s->rhs, n);
  if (nrows > 0) {
    if (int rc = rbuf(s, kRsRowPtr, &s->row_ptr, nrows + 1)) return rc;
    if (int rc = rbuf(s, kRsCols, &s->cols, nnz)) return rc;
    if (int rc = rbuf(s, kRsCoef64, &s->coef64, nnz)) return rc;
    if (int rc = rbuf(s, kRsCoef32, &s->coef32, nnz)) return rc;
    if (int rc = rbuf(s, kRsB, &s->b, nrows)) return rc;
    if (int rc = rbuf(s, kRsW, &s->w, nrows)) return rc;
    if (int rc = rbuf(s, kRsW2_64, &s->w2_64, nrows)) return rc;
    if (int rc = rbuf(s, kRsW2_32, &s->w2_32, nrows)) return rc;
    if (int rc = rbuf(s, kRsT64, &s->t64, nrows)) return rc;
    if (int rc = rbuf(s, kRsT32, &s->t32, nrows)) return rc;
    if (int rc = rbuf(s, kRsPartRows, &s->part_rows, ceil_div(nrows, 256))) return rc;
    if (int rc = rbuf(s, kRsURows, &s->u_rows, nrows)) return rc;
    LSAMD_HIP(hipMemcpyAsync(s->row_ptr, row_ptr, (nrows + 1) * 8, hipMemcpyDefault, st));
    LSAMD_HIP(hipMemcpyAsync(s->cols, cols, nnz * 4, hipMemcpyDefault, st));
    LSAMD_HIP(hipMemcpyAsync(s->coef64, coef, nnz * 8, hipMemcpyDefault, st));
    LSAMD_HIP(hipMemcpyAsync(s->b, b, nrows * 8, hipMemcpyDefault, st));
    LSAMD_HIP(hipMemcpyAsync(s->w, w, nrows * 8, hipMemcpyDefault, st));
    {
      int* bad = nullptr;
      if (int rc = rbuf(s, kRsBad, &bad, 1)) return rc;
      LSAMD_HIP(hipMemsetAsync(bad, 0, sizeof(int), st));
      hipLaunchKernelGGL(k_check_cols, dim3((unsigned)ceil_div(nnz, 256)), dim3(256), 0, st,
                         (const int32_t*)s->cols, nnz, (int64_t)0, s->n, bad);
      int hb = 0;
      LSAMD_HIP(hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, st));
      LSAMD_HIP(hipStreamSynchronize(st));
      if (hb) {
        s->nrows = s->nnz = 0;
        return fail(LSAMD_ERR_ARG, "lsamd_fd_system_set_rows: a column id is outside the grid");
      }
    }
    const int64_t m = std::max(nrows, nnz);
    hipLaunchKernelGGL(k_w2, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, st, s->w, s->w2_64,
                       s->w2_32, s->coef64, s->coef32, nrows, nnz);
    // transpose: stable radix sort of entries by column
    int32_t *erow = nullptr, *keys_out = nullptr, *vals_in = nullptr;
    int32_t* uniq = nullptr;
    int32_t* counts = nullptr;
    int32_t* nrun = nullptr;
    if (int rc = rbuf(s, kRsErow, &erow, nnz)) return rc;
    if (int rc = rbuf(s, kRsKeys, &keys_out, nnz)) return rc;
    if (int rc = rbuf(s, kRsVals, &vals_in, nnz)) return rc;
    if (int rc = rbuf(s, kRsSegEntry, &s->seg_entry, nnz)) return rc;
    if (int rc = rbuf(s, kRsSegRow, &s->seg_row, nnz)) return rc;
    if (int rc = rbuf(s, kRsUniq, &uniq, nnz)) return rc;
    if (int rc = rbuf(s, kRsCounts, &counts, nnz)) return rc;
    if (int rc = rbuf(s, kRsNrun, &nrun, 1)) return rc;
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
    if (int rc = rbuf(s, kRsTmp, (unsigned char**)&tmp, (int64_t)tbytes)) return rc;
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
    if (int rc = rbuf(s, kRsSegPtr, &s->seg_ptr, nr + 1)) return rc;
    if (int rc = rbuf(s, kRsDnode, &s->dnode, nr)) return rc;
    // segment offsets (int32 scan, then widened)
    int32_t* off32 = nullptr;
    if (int rc = rbuf(s, kRsOff32, &off32, nr + 1)) return rc;
    tb = tbytes;
    LSAMD_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, counts, off32, (int)nr, st));
    std::vector<int32_t> hoff(nr + 1);
    LSAMD_HIP(hipMemcpyAsync(hoff.data(), off32, nr * 4, hipMemcpyDeviceToHost, st));
    LSAMD_HIP(hipStreamSynchronize(st));
    hoff[nr] = (int32_t)nnz;
    std::vector<int64_t> hoff64(hoff.begin(), hoff.end());
    LSAMD_HIP(legacy_memcpy(s->seg_ptr, hoff64.data(), (nr + 1) * 8, hipMemcpyHostToDevice));
    LSAMD_HIP(hipMemcpyAsync(s->dnode, uniq, nr * 4, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_gather_i32, dim3((unsigned)ceil_div(nnz, 256)), dim3(256), 0, st, erow,
                       s->seg_entry, s->seg_row, nnz);
    LSAMD_CHECK_LAUNCH("transpose");
    LSAMD_HIP(hipStreamSynchronize(st));
    // rhs = A^T W^2 b, data diagonal
    RowsDev R = rows_dev(s, false);
    hipLaunchKernelGGL(k_w2b, dim3((unsigned)ceil_div(nrows, 256)), dim3(256), 0, st, s->w2_64,
                       s->b, s->t64, nrows);
    hipLaunchKernelGGL((k_rows_scatter<double, double>), dim3((unsigned)ceil_div(s->ndn, 256)),
                       dim3(256), 0, st, R, (const double*)s->t64, s->rhs, 1, (CgCtl*)nullptr, 0,
                       (const double*)nullptr, 0, (const double*)nullptr, 0);
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
  if (precision != LSAMD_FP64 && precision != LSAMD_FP32)
    return fail(LSAMD_ERR_ARG, "unknown precision");
  if (int rc = ensure_work(s, precision == LSAMD_FP32)) return rc;
  if (precision == LSAMD_FP64) {
    if (int rc = ensure_tuned<double>(s, s->p64, s->q64, st)) return rc;
  } else {
    if (int rc = ensure_tuned<float>(s, s->p32, s->q32, st)) return rc;
  }
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
  // preconditioner: Jacobi (= the reference's column scaling) or identity
  const double* minv64 = s->minv64;
  const float* minv32 = s->minv32;
  if (o->jacobi == 2) {
    lsamd_solve_info inf;
    memset(&inf, 0, sizeof(inf));
    if (int rc = mg_solve(s, o, rhs ? rhs : s->rhs, c, &inf, (hipStream_t)stream)) return rc;
    if (info) *info = inf;
    return LSAMD_OK;
  }
  if (!o->jacobi) {
    if (!s->ones64) {
      if (int rc = dalloc(s, &s->ones64, n)) return rc;
      if (int rc = dalloc(s, &s->ones32, n)) return rc;
      hipLaunchKernelGGL(k_fill_one, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st,
                         s->ones64, s->ones32, n);
      LSAMD_CHECK_LAUNCH("k_fill_one");
    }
    minv64 = s->ones64;
    minv32 = s->ones32;
  }
  if (s->minv_graph[0] != minv64 || s->minv_graph[1] != (const void*)minv32) {
    for (int i = 0; i < 2; ++i)
      if (s->graph[i]) {
        hipGraphExecDestroy(s->graph[i]);
        s->graph[i] = nullptr;
      }
    s->minv_graph[0] = minv64;
    s->minv_graph[1] = minv32;
  }
  const int64_t maxiter64 = o->maxiter > 0 ? o->maxiter : 10 * n;
  const int maxiter = (int)std::min<int64_t>(maxiter64, 2000000000LL);
  const int chunk = o->check_every > 0 ? o->check_every : 64;
  const int vb = vec_blocks(n);
  lsamd_solve_info inf;
  memset(&inf, 0, sizeof(inf));
  if (o->precision == LSAMD_FP64) {
    if (int rc = ensure_work(s, false)) return rc;
    if (int rc = ensure_tuned<double>(s, s->p64, s->q64, st)) return rc;
    if (int rc = launch_init<double>(s, b, minv64, s->x64, s->r64, o->rtol, o->atol, maxiter, st))
      return rc;
    if (int rc = run_pcg<double>(s, chunk, s->x64, s->r64, s->p64, s->q64, s->t64, minv64, st))
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
    if (int rc = ensure_tuned<double>(s, s->p64, s->q64, st)) return rc;
    if (int rc = ensure_tuned<float>(s, s->p32, s->q32, st)) return rc;
    const double inner = o->inner_rtol > 0 ? o->inner_rtol : 1e-4;
    const int64_t max_outer = o->max_outer > 0 ? o->max_outer : 50;
    hipLaunchKernelGGL(k_zero_d, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, c, n);
    // threshold from |S rhs| (same test as the FP64 form)
    if (int rc = launch_init<double>(s, b, minv64, s->x64, s->r64, o->rtol, o->atol, maxiter, st))
      return rc;
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
      if (int rc = launch_init<float>(s, s->r64, minv32, s->x32, s->r32, irtol, 0.0, left, st))
        return rc;
      if (int rc = run_pcg<float>(s, chunk, s->x32, s->r32, s->p32, s->q32, s->t32, minv32, st))
        return rc;
      total += s->ctl_host->iter;
      ++outer;
      hipLaunchKernelGGL(k_axpy_f32_to_f64, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st,
                         s->x32, c, n);
      // r64 = rhs - N c in FP64
      if (int rc = launch_apply<double>(s, c, s->q64, s->t64, s->ctl_apply, st)) return rc;
      hipLaunchKernelGGL(k_residual, dim3(vb), dim3(VB), 0, st, b, s->q64, minv64, s->r64, n,
                         s->part);
      hipLaunchKernelGGL(k_sum_fin, dim3(1), dim3(VB), 0, st, (const double*)s->part, vb,
                         s->ctl_apply);
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

int lsamd_fd_system_lsmr(lsamd_fd_system* s, const lsamd_lsmr_opts* o, double* c,
                         lsamd_lsmr_info* info, void* stream) {
  if (int rc = injected("lsmr")) return rc;
  if (!s || !o || !c) return fail(LSAMD_ERR_ARG, "lsamd_fd_system_lsmr: NULL argument");
  if (!(o->atol >= 0.0) || !(o->btol >= 0.0))
    return fail(LSAMD_ERR_ARG, "atol/btol must be non-negative");
  lsamd_lsmr_info inf;
  memset(&inf, 0, sizeof(inf));
  if (int rc = ls_solve(s, o, c, &inf, (hipStream_t)stream)) return rc;
  if (info) *info = inf;
  return LSAMD_OK;
}

int lsamd_fd_system_solve_multi(lsamd_fd_system* s, const lsamd_solve_opts* o, int32_t nrhs,
                                const double* rhs, double* c, lsamd_solve_info* infos,
                                void* stream) {
  if (int rc = injected("solve")) return rc;
  if (!s || !o || !rhs || !c || !infos || nrhs < 0)
    return fail(LSAMD_ERR_ARG, "lsamd_fd_system_solve_multi: bad argument");
  if (!(o->rtol >= 0.0) || !(o->atol >= 0.0))
    return fail(LSAMD_ERR_ARG, "rtol/atol must be non-negative");
  if (o->jacobi != 2)
    return fail(LSAMD_ERR_ARG, "lsamd_fd_system_solve_multi: multigrid only (jacobi = 2)");
  if (nrhs == 0) return LSAMD_OK;
  return mr_solve(s, o, nrhs, rhs, c, infos, (hipStream_t)stream);
}

int lsamd_fd_system_rhs_rows(lsamd_fd_system* s, const double* b, double* rhs, void* stream) {
  if (!s || !b || !rhs) return fail(LSAMD_ERR_ARG, "lsamd_fd_system_rhs_rows: NULL argument");
  if (s->op == 2) return fail(LSAMD_ERR_ARG, "lsamd_fd_system_rhs_rows: not for batches");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = s->n;
  hipLaunchKernelGGL(k_zero_d, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, rhs, n);
  if (s->nrows > 0) {
    RowsDev R = rows_dev(s, false);
    hipLaunchKernelGGL(k_w2b, dim3((unsigned)ceil_div(s->nrows, 256)), dim3(256), 0, st,
                       (const double*)s->w2_64, b, s->t64, s->nrows);
    hipLaunchKernelGGL((k_rows_scatter<double, double>),
                       dim3((unsigned)std::max<int64_t>(1, ceil_div(s->ndn, 256))), dim3(256), 0,
                       st, R, (const double*)s->t64, rhs, 1, (CgCtl*)nullptr, 0,
                       (const double*)nullptr, 0, (const double*)nullptr, 0);
  }
  LSAMD_CHECK_LAUNCH("rhs_rows");
  return LSAMD_OK;
}

int lsamd_fd_system_time_multi(lsamd_fd_system* s, int32_t kernel, int32_t iters,
                               double* avg_ms, double* bytes, void* stream) {
  if (!s || !avg_ms || !bytes || iters < 1 || kernel < 0 || kernel > 2)
    return fail(LSAMD_ERR_ARG, "time_multi: bad arguments");
  if (!s->mg || !s->regular)
    return fail(LSAMD_ERR_ARG, "time_multi: needs a 32- or 64-lane solve_multi on this handle");
  hipStream_t st = (hipStream_t)stream;
  // the widest group this handle has built
  if (MrWork* w = s->mr[4]; w && w->R == 64)
    return time_multi<64>(s, w, kernel, iters, avg_ms, bytes, st);
  if (MrWork* w = s->mr[3]; w && w->R == 32)
    return time_multi<32>(s, w, kernel, iters, avg_ms, bytes, st);
  return fail(LSAMD_ERR_ARG, "time_multi: needs a 32- or 64-lane solve_multi on this handle");
}


int lsamd_fd_system_time_stencil(lsamd_fd_system* s, int32_t precision, int32_t iters,
                                 double* avg_ms, void* stream) {
  if (!s || !avg_ms || iters < 1) return fail(LSAMD_ERR_ARG, "time_stencil: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const bool f32 = precision == LSAMD_FP32;
  if (int rc = ensure_work(s, f32)) return rc;
  if (f32) {
    if (int rc = ensure_tuned<float>(s, s->p32, s->q32, st)) return rc;
  } else {
    if (int rc = ensure_tuned<double>(s, s->p64, s->q64, st)) return rc;
  }
  hipEvent_t e0, e1;
  LSAMD_HIP(hipEventCreate(&e0));
  LSAMD_HIP(hipEventCreate(&e1));
  // one warm launch, then `iters` back-to-back launches between one event pair
  if (f32)
    launch_stencil<float>(s, s->p32, s->q32, s->ctl_apply, st, s->zc[1]);
  else
    launch_stencil<double>(s, s->p64, s->q64, s->ctl_apply, st, s->zc[0]);
  LSAMD_HIP(hipEventRecord(e0, st));
  for (int it = 0; it < iters; ++it) {
    if (f32)
      launch_stencil<float>(s, s->p32, s->q32, s->ctl_apply, st, s->zc[1]);
    else
      launch_stencil<double>(s, s->p64, s->q64, s->ctl_apply, st, s->zc[0]);
  }
  LSAMD_HIP(hipEventRecord(e1, st));
  LSAMD_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  LSAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  LSAMD_CHECK_LAUNCH("time_stencil");
  *avg_ms = (double)ms / iters;
  return LSAMD_OK;
}

int lsamd_fd_evaluate_adjoint(const lsamd_grid* g, const double* pts, int64_t n, const double* gv,
                              double* out, void* stream) {
  // out = E^T gv with E the trilinear evaluation matrix of the points (outside
  // points contribute nothing).  Deterministic: entries are sorted by node and
  // summed in point order.  One temporary allocation, freed before return.
  if (int rc = injected("evaluate_adjoint")) return rc;
  if (int rc = check_grid(g)) return rc;
  if (n < 0) return fail(LSAMD_ERR_ARG, "lsamd_fd_evaluate_adjoint: bad arguments");
  const size_t bytes = eval_adjoint_bytes(g, n);
  void* ws = nullptr;
  if (bytes > 0 && hipMalloc(&ws, bytes) != hipSuccess)
    return fail(LSAMD_ERR_ALLOC, "evaluate_adjoint: hipMalloc failed");
  hipStream_t st = (hipStream_t)stream;
  const int rc = eval_adjoint_into(g, pts, n, gv, out, (unsigned char*)ws, bytes, st,
                                   "lsamd_fd_evaluate_adjoint");
  if (ws) {
    hipStreamSynchronize(st);
    hipFree(ws);
  }
  return rc;
}

int lsamd_fd_evaluate_adjoint_noalloc(const lsamd_grid* g, const double* pts, int64_t n,
                                      const double* gv, double* out, void* stream) {
  // lsamd_fd_evaluate_adjoint with its per-entry arrays and sort storage
  // carved from the device's scratch array instead of allocated: no
  // hipMalloc/hipFree and no stream synchronisation.
  if (int rc = injected("evaluate_adjoint_noalloc")) return rc;
  if (n > LSAMD_NOALLOC_MAX_POINTS)
    return fail(LSAMD_ERR_ARG, "lsamd_fd_evaluate_adjoint_noalloc: " + std::to_string(n) +
                                   " points exceed the scratch capacity of " +
                                   std::to_string(LSAMD_NOALLOC_MAX_POINTS));
  unsigned char* base = nullptr;
  if (int rc = scratch_base(&base)) return rc;
  return eval_adjoint_into(g, pts, n, gv, out, base, kScratchBytes, (hipStream_t)stream,
                           "lsamd_fd_evaluate_adjoint_noalloc");
}

int lsamd_fd_system_evaluate_adjoint(lsamd_fd_system* s, const double* pts, int64_t n,
                                     const double* gv, double* out, void* stream) {
  // lsamd_fd_evaluate_adjoint on the handle's grid with its temporaries in a
  // buffer cached on the handle (grown when a call needs more): repeated
  // calls (an autograd backward pass each) stop allocating and never
  // synchronise.
  if (int rc = injected("system_evaluate_adjoint")) return rc;
  if (!s) return fail(LSAMD_ERR_ARG, "lsamd_fd_system_evaluate_adjoint: NULL system");
  if (n < 0) return fail(LSAMD_ERR_ARG, "lsamd_fd_system_evaluate_adjoint: n < 0");
  unsigned char* ws = nullptr;
  const size_t bytes = eval_adjoint_bytes(&s->grid, n);
  if (bytes > 0)
    if (int rc = rbuf(s, kRsAdjoint, &ws, (int64_t)bytes)) return rc;
  return eval_adjoint_into(&s->grid, pts, n, gv, out, ws,
                           bytes > 0 ? s->rpool[kRsAdjoint].second : 0, (hipStream_t)stream,
                           "lsamd_fd_system_evaluate_adjoint");
}

int lsamd_fd_system_mg_build(lsamd_fd_system* s, void* stream) {
  if (int rc = injected("mg_build")) return rc;
  if (!s || s->op != 0) return fail(LSAMD_ERR_ARG, "lsamd_fd_system_mg_build: not an FD system");
  if (int rc = ensure_work(s, false)) return rc;
  return mg_ensure(s, (hipStream_t)stream);
}

int lsamd_fd_systems_solve_mg(lsamd_fd_system* const* systems, int32_t nsys,
                              const lsamd_solve_opts* o, double* const* c,
                              lsamd_solve_info* infos, void* stream) {
  if (int rc = injected("systems_solve_mg")) return rc;
  if (!systems || nsys < 1 || !o || !c || !infos)
    return fail(LSAMD_ERR_ARG, "lsamd_fd_systems_solve_mg: bad arguments");
  if (!(o->rtol >= 0.0) || !(o->atol >= 0.0))
    return fail(LSAMD_ERR_ARG, "rtol/atol must be non-negative");
  for (int i = 0; i < nsys; ++i)
    for (int j = 0; j < i; ++j)
      if (systems[i] == systems[j])
        return fail(LSAMD_ERR_ARG, "lsamd_fd_systems_solve_mg: a system appears twice");
  return bmg_solve(systems, nsys, o, c, infos, (hipStream_t)stream);
}

int lsamd_fd_system_stencil_info(lsamd_fd_system* s, int32_t precision, int32_t* zc,
                                 float* tune_us) {
  if (!s || !zc) return fail(LSAMD_ERR_ARG, "stencil_info: NULL argument");
  const int pi = precision == LSAMD_FP32 ? 1 : 0;
  *zc = s->regular ? s->zc[pi] : 0;
  if (tune_us)
    for (int c = 0; c < kNStCand; ++c) tune_us[c] = s->tune_us[pi][c];
  return LSAMD_OK;
}

int lsamd_fd_nearest_d2(const lsamd_grid* g, const double* pts, int64_t m, double* d2,
                        void* stream) {
  if (int rc = injected("nearest_d2")) return rc;
  if (int rc = check_grid(g)) return rc;
  if (m < 0 || !d2 || (m > 0 && !pts)) return fail(LSAMD_ERR_ARG, "lsamd_fd_nearest_d2: bad args");
  const int64_t n = g->nnodes[0] * g->nnodes[1] * g->nnodes[2];
  hipLaunchKernelGGL(k_nearest_d2, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, to_dev(*g), pts, m, d2);
  LSAMD_CHECK_LAUNCH("k_nearest_d2");
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


// ------------------------------------------------------------- batch ----
int lsamd_batch_create(int32_t nsys, const lsamd_grid* grids, const lsamd_fd_reg* reg,
                       lsamd_fd_system** out) {
  if (int rc = injected("batch_create")) return rc;
  if (!out || !grids || !reg) return fail(LSAMD_ERR_ARG, "lsamd_batch_create: NULL argument");
  *out = nullptr;
  if (nsys < 1) return fail(LSAMD_ERR_ARG, "lsamd_batch_create: nsys < 1");
  std::vector<BSys> hs(nsys);
  int64_t off = 0;
  int64_t nb = 0;
  for (int i = 0; i < nsys; ++i) {
    if (int rc = check_uniform_grid(&grids[i], "lsamd_batch_create")) return rc;
    for (int a = 0; a < 3; ++a)
      if (grids[i].nnodes[a] < 7)
        return fail(LSAMD_ERR_ARG, "lsamd_batch_create: every grid needs >= 7 nodes per axis");
    const int64_t n = grids[i].nnodes[0] * grids[i].nnodes[1] * grids[i].nnodes[2];
    hs[i].g = to_dev(grids[i]);
    hs[i].off = off;
    hs[i].n = n;
    hs[i].r0 = hs[i].r1 = 0;
    hs[i].t0 = hs[i].t1 = 0;
    hs[i].b0 = (int)nb;
    nb += ceil_div(n, kBNodes);
    hs[i].b1 = (int)nb;
    off += n;
    if (nb > 2000000000LL || off > ((int64_t)1 << 31) - 1)
      return fail(LSAMD_ERR_ARG, "lsamd_batch_create: ensemble too large");
  }
  lsamd_fd_system* s = new lsamd_fd_system();
  s->grid = grids[0];
  s->gd = to_dev(grids[0]);
  s->reg = *reg;
  s->n = off;
  s->op = 2;
  s->nsys = nsys;
  s->nblk = (int)nb;
  auto bail = [&](int rc) {
    lsamd_fd_system_destroy(s);
    return rc;
  };
  std::vector<int> bsys_tab(nb);
  for (int i = 0; i < nsys; ++i)
    for (int b = hs[i].b0; b < hs[i].b1; ++b) bsys_tab[b] = i;
  // stencil tiles: per system, bulk z-chunks of every xy tile (uniform-
  // coefficient tiles flagged) then the two z-boundary parts
  std::vector<int4> tiles[4];
  for (int f = 0; f < 4; ++f) {
    s->trange[f].resize(nsys);
    for (int i = 0; i < nsys; ++i) {
      const int a = (int)tiles[f].size();
      stencil_tiles(i, hs[i].g, (f & 2) ? 2 * kBZc : kBZc, tiles[f], (f & 1) ? 32 : 64);
      s->trange[f][i] = make_int2(a, (int)tiles[f].size());
      if (tiles[f].size() > 2000000000ULL) {
        lsamd_fd_system_destroy(s);
        return fail(LSAMD_ERR_ARG, "lsamd_batch_create: ensemble too large");
      }
    }
  }
  for (int i = 0; i < nsys; ++i) {  // form 0 until the batch is tuned
    hs[i].t0 = s->trange[0][i].x;
    hs[i].t1 = s->trange[0][i].y;
  }
  if (int rc = dalloc(s, &s->bsys, nsys)) return bail(rc);
  if (int rc = dalloc(s, &s->blk_sys, nb)) return bail(rc);
  if (int rc = dalloc(s, &s->bctl, nsys)) return bail(rc);
  if (int rc = dalloc(s, &s->bctl_apply, nsys)) return bail(rc);
  if (int rc = dalloc(s, &s->bctl64, nsys)) return bail(rc);
  if (int rc = dalloc(s, &s->odone, nsys)) return bail(rc);
  if (int rc = dalloc(s, &s->bres, nsys)) return bail(rc);
  s->bsys_h = hs;  // with the tile ranges
  for (int f = 0; f < 4; ++f) {
    s->nbt[f] = (int)tiles[f].size();
    if (int rc = dalloc(s, &s->btab[f], tiles[f].size())) return bail(rc);
    LSAMD_HIP(legacy_memcpy(s->btab[f], tiles[f].data(), tiles[f].size() * sizeof(int4),
                            hipMemcpyHostToDevice));
  }
  LSAMD_HIP(legacy_memcpy(s->bsys, hs.data(), nsys * sizeof(BSys), hipMemcpyHostToDevice));
  LSAMD_HIP(legacy_memcpy(s->blk_sys, bsys_tab.data(), nb * sizeof(int), hipMemcpyHostToDevice));
  LSAMD_HIP(legacy_memset(s->bctl, 0, nsys * sizeof(CgCtl)));
  LSAMD_HIP(legacy_memset(s->bctl_apply, 0, nsys * sizeof(CgCtl)));
  LSAMD_HIP(legacy_memset(s->bctl64, 0, nsys * sizeof(CgCtl)));
  if (hipHostMalloc((void**)&s->bctl_host, nsys * sizeof(CgCtl), hipHostMallocDefault) !=
      hipSuccess)
    return bail(fail(LSAMD_ERR_ALLOC, "hipHostMalloc failed"));
  std::vector<double> tab(kClasses * kTaps);
  build_class_table(s->reg, tab.data());
  std::vector<float> tabf(tab.begin(), tab.end());
  if (int rc = dalloc(s, &s->tab64, kClasses * kTaps)) return bail(rc);
  if (int rc = dalloc(s, &s->tab32, kClasses * kTaps)) return bail(rc);
  LSAMD_HIP(legacy_memcpy(s->tab64, tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
  LSAMD_HIP(legacy_memcpy(s->tab32, tabf.data(), tabf.size() * 4, hipMemcpyHostToDevice));
  const int bulk = ((3 * 7 + 3) * 7 + 3) * kTaps;
  for (int q = 0; q < kTaps; ++q) {
    s->bku64.k[q] = tab[bulk + q];
    s->bku32.k[q] = (float)tab[bulk + q];
  }
  if (int rc = init_common(s)) return bail(rc);
  *out = s;
  return LSAMD_OK;
}

int lsamd_batch_offsets(lsamd_fd_system* s, int64_t* node_off) {
  if (!s || !node_off || s->op != 2) return fail(LSAMD_ERR_ARG, "lsamd_batch_offsets: not a batch");
  for (int i = 0; i < s->nsys; ++i) node_off[i] = s->bsys_h[i].off;
  node_off[s->nsys] = s->n;
  return LSAMD_OK;
}

int lsamd_batch_set_rows(lsamd_fd_system* s, const int64_t* sys_rows, int64_t nrows, int64_t nnz,
                         const int64_t* row_ptr, const int32_t* cols, const double* coef,
                         const double* b, const double* w, void* stream) {
  if (int rc = injected("batch_set_rows")) return rc;
  if (!s || s->op != 2 || !sys_rows) return fail(LSAMD_ERR_ARG, "lsamd_batch_set_rows: not a batch");
  if (sys_rows[0] != 0 || sys_rows[s->nsys] != nrows)
    return fail(LSAMD_ERR_ARG, "lsamd_batch_set_rows: sys_rows must run from 0 to nrows");
  for (int i = 0; i < s->nsys; ++i)
    if (sys_rows[i + 1] < sys_rows[i])
      return fail(LSAMD_ERR_ARG, "lsamd_batch_set_rows: sys_rows must be non-decreasing");
  if (int rc = lsamd_fd_system_set_rows(s, nrows, nnz, row_ptr, cols, coef, b, w, stream))
    return rc;
  for (int i = 0; i < s->nsys; ++i) {
    s->bsys_h[i].r0 = sys_rows[i];
    s->bsys_h[i].r1 = sys_rows[i + 1];
  }
  LSAMD_HIP(legacy_memcpy(s->bsys, s->bsys_h.data(), s->nsys * sizeof(BSys), hipMemcpyHostToDevice));
  if (nrows > 0) {
    hipStream_t st = (hipStream_t)stream;
    int* bad = nullptr;
    if (int rc = dalloc(s, &bad, 1)) return rc;
    LSAMD_HIP(hipMemsetAsync(bad, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_b_check_rows, dim3((unsigned)s->nsys), dim3(256), 0, st,
                       rows_dev(s, false), (const BSys*)s->bsys, bad);
    int hb = 0;
    LSAMD_HIP(hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, st));
    LSAMD_HIP(hipStreamSynchronize(st));
    dfree(s, bad);
    if (hb) {
      lsamd_fd_system_set_rows(s, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, stream);
      for (auto& q : s->bsys_h) q.r0 = q.r1 = 0;
      LSAMD_HIP(legacy_memcpy(s->bsys, s->bsys_h.data(), s->nsys * sizeof(BSys),
                          hipMemcpyHostToDevice));
      return fail(LSAMD_ERR_ARG, "lsamd_batch_set_rows: a row addresses another system's nodes");
    }
  }
  return LSAMD_OK;
}

int lsamd_batch_solve(lsamd_fd_system* s, const lsamd_solve_opts* o, const double* rhs, double* c,
                      lsamd_solve_info* info, void* stream) {
  if (int rc = injected("batch_solve")) return rc;
  if (!s || !o || !c || s->op != 2) return fail(LSAMD_ERR_ARG, "lsamd_batch_solve: bad argument");
  if (!(o->rtol >= 0.0) || !(o->atol >= 0.0))
    return fail(LSAMD_ERR_ARG, "rtol/atol must be non-negative");
  if (!o->jacobi) return fail(LSAMD_ERR_ARG, "lsamd_batch_solve: only the Jacobi form is batched");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = s->n;
  const int ns = s->nsys;
  const double* b = rhs ? rhs : s->rhs;
  if (int rc = ensure_batch_work(s)) return rc;
  if (int rc = ensure_batch_tuned(s, st)) return rc;
  int64_t nmax = 0;
  for (const BSys& q : s->bsys_h) nmax = std::max(nmax, q.n);
  const int64_t maxiter64 = o->maxiter > 0 ? o->maxiter : 10 * nmax;
  const int maxiter = (int)std::min<int64_t>(maxiter64, 2000000000LL);
  const int chunk = o->check_every > 0 ? o->check_every : 64;
  std::vector<lsamd_solve_info> inf(ns);
  memset(inf.data(), 0, ns * sizeof(lsamd_solve_info));
  if (o->precision == LSAMD_FP64) {
    if (int rc = launch_b_init<double>(s, b, s->minv64, s->x64, s->r64, o->rtol, o->atol, maxiter,
                                       nullptr, st))
      return rc;
    if (int rc = run_b_pcg<double>(s, chunk, s->x64, s->r64, s->p64, s->q64, s->t64, s->minv64, st))
      return rc;
    hipLaunchKernelGGL(k_copy_d, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, s->x64, c, n);
    LSAMD_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < ns; ++i) {
      inf[i].iterations = s->bctl_host[i].iter;
      inf[i].converged = s->bctl_host[i].converged;
      inf[i].residual = sqrt(s->bctl_host[i].rho);
      inf[i].threshold = s->bctl_host[i].thr;
    }
  } else if (o->precision == LSAMD_FP32) {
    const double inner = o->inner_rtol > 0 ? o->inner_rtol : 1e-4;
    const int64_t max_outer = o->max_outer > 0 ? o->max_outer : 50;
    hipLaunchKernelGGL(k_zero_d, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, c, n);
    // FP64 thresholds and initial residual r64 = rhs (c = 0)
    if (int rc = launch_b_init<double>(s, b, s->minv64, s->x64, s->r64, o->rtol, o->atol, maxiter,
                                       nullptr, st))
      return rc;
    LSAMD_HIP(hipMemcpyAsync(s->bctl64, s->bctl, ns * sizeof(CgCtl), hipMemcpyDeviceToDevice, st));
    std::vector<int> od(ns);
    LSAMD_HIP(hipMemcpyAsync(s->bctl_host, s->bctl, ns * sizeof(CgCtl), hipMemcpyDeviceToHost, st));
    LSAMD_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < ns; ++i) {
      od[i] = s->bctl_host[i].converged;
      inf[i].threshold = s->bctl_host[i].thr;
      inf[i].residual = sqrt(s->bctl_host[i].rho);
    }
    LSAMD_HIP(hipMemcpyAsync(s->odone, od.data(), ns * sizeof(int), hipMemcpyHostToDevice, st));
    int64_t outer = 0;
    bool all = true;
    for (int i = 0; i < ns; ++i) all = all && od[i];
    std::vector<double> res(ns);
    while (!all && outer < max_outer) {
      if (int rc = launch_b_init<float>(s, s->r64, s->minv32, s->x32, s->r32, inner, 0.0, maxiter,
                                        s->odone, st))
        return rc;
      if (int rc = run_b_pcg<float>(s, chunk, s->x32, s->r32, s->p32, s->q32, s->t32, s->minv32,
                                    st))
        return rc;
      for (int i = 0; i < ns; ++i)
        if (!od[i]) {
          inf[i].iterations += s->bctl_host[i].iter;
          inf[i].outer += 1;
        }
      ++outer;
      hipLaunchKernelGGL(k_b_accum, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, s->x32, c,
                         n);
      if (int rc = launch_apply<double>(s, c, s->q64, s->t64, s->ctl_apply, st)) return rc;
      hipLaunchKernelGGL(k_b_residual, dim3((unsigned)s->nblk), dim3(256), 0, st, b, s->q64,
                         s->minv64, s->r64, (const BSys*)s->bsys, (const int*)s->blk_sys, s->part);
      hipLaunchKernelGGL(k_b_outer_fin, dim3((unsigned)ns), dim3(256), 0, st,
                         (const double*)s->part, (const BSys*)s->bsys, (const CgCtl*)s->bctl64,
                         s->bres, s->odone);
      LSAMD_CHECK_LAUNCH("batch refinement");
      LSAMD_HIP(hipMemcpyAsync(od.data(), s->odone, ns * sizeof(int), hipMemcpyDeviceToHost, st));
      LSAMD_HIP(hipMemcpyAsync(res.data(), s->bres, ns * sizeof(double), hipMemcpyDeviceToHost, st));
      LSAMD_HIP(hipStreamSynchronize(st));
      all = true;
      for (int i = 0; i < ns; ++i) {
        if (inf[i].outer == outer) inf[i].residual = res[i];
        all = all && od[i];
      }
    }
    for (int i = 0; i < ns; ++i) inf[i].converged = od[i];
  } else {
    return fail(LSAMD_ERR_ARG, "unknown precision");
  }
  if (info) memcpy(info, inf.data(), ns * sizeof(lsamd_solve_info));
  return LSAMD_OK;
}

}  // extern "C"
