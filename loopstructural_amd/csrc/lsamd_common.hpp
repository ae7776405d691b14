// Shared device/host helpers for libloopstructural_amd (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/loopstructural_amd.h"

namespace lsamd {

// ---------------------------------------------------------------- errors --
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
// Returns LSAMD_ERR_INJECTED when the LSAMD_FAULT_INJECT environment
// variable names `site` (comma separated list), so host tests can force the
// error paths of every entry point without a faulting kernel.
int injected(const char* site);

#define LSAMD_HIP(expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return ::lsamd::fail(LSAMD_ERR_HIP, std::string(#expr) + ": " +                \
                                              hipGetErrorString(e_));                \
  } while (0)

#define LSAMD_CHECK_LAUNCH(name)                                                     \
  do {                                                                               \
    hipError_t e_ = hipGetLastError();                                               \
    if (e_ != hipSuccess)                                                            \
      return ::lsamd::fail(LSAMD_ERR_HIP, std::string(name) + ": " +                 \
                                              hipGetErrorString(e_));                \
  } while (0)

constexpr int kWave = 64;

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Device view of a structured grid.
struct GridDev {
  double ox, oy, oz;
  double hx, hy, hz;
  int nx, ny, nz;  // nodes per axis
  // node coordinates per axis of a rectilinear grid (all NULL: uniform)
  const double* xn;
  const double* yn;
  const double* zn;
};

inline GridDev to_dev(const lsamd_grid& g) {
  GridDev d;
  d.ox = g.origin[0];
  d.oy = g.origin[1];
  d.oz = g.origin[2];
  d.hx = g.step[0];
  d.hy = g.step[1];
  d.hz = g.step[2];
  d.nx = (int)g.nnodes[0];
  d.ny = (int)g.nnodes[1];
  d.nz = (int)g.nnodes[2];
  d.xn = g.axis_nodes[0];
  d.yn = g.axis_nodes[1];
  d.zn = g.axis_nodes[2];
  return d;
}

int check_grid(const lsamd_grid* g);
// check_grid, and the grid must be uniform (P1 meshes, batches)
int check_uniform_grid(const lsamd_grid* g, const char* fn);

// ------------------------------------------------ deterministic reductions --
// Fixed-order wave64 butterfly sum.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum in a fixed order (waves reduced by lane 0 of wave 0 in index
// order).  `sh` must hold blockDim.x/64 doubles.  Result valid in thread 0.
__device__ inline double block_sum(double v, double* sh) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  v = wave_sum(v);
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) s += sh[i];
  }
  __syncthreads();
  return s;
}

// Workgroup index renumbered so that each XCD takes one contiguous run of
// the grid: gfx950 hands workgroups to its 8 XCDs round-robin by linear id,
// so neighbouring workgroups (which gather from overlapping neighbourhoods)
// otherwise sit on different XCDs and each XCD's L2 fetches the shared lines
// again.  A bijection of [0, gridDim.x * gridDim.y); on = 0 keeps blockIdx.
__device__ inline void xcd_block(int on, unsigned& bx, unsigned& by) {
  bx = blockIdx.x;
  by = blockIdx.y;
  if (!on) return;
  const unsigned gx = gridDim.x, t = gx * gridDim.y;
  const unsigned l = blockIdx.y * gx + blockIdx.x;
  const unsigned q = t / 8, rem = t % 8, k = l % 8;
  const unsigned lp = k * q + (k < rem ? k : rem) + l / 8;
  bx = lp % gx;
  by = lp / gx;
}

// Device-coherent load of a value written by another workgroup.
__device__ inline double coherent_load(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int coherent_load(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum `n` partials (written by an EARLIER kernel of the stream, so plain
// loads see them) in a fixed order with the whole block; result in thread 0.
// Cross-workgroup reductions always go through a kernel boundary: an
// agent-scope fence inside a kernel writes back and invalidates the L2 of
// the XCD on gfx950, which costs far more than one extra launch.
__device__ inline double reduce_partials(const double* part, int n, double* sh) {
  double v = 0.0;
  const int per = (n + (int)blockDim.x - 1) / (int)blockDim.x;
  const int b = threadIdx.x * per;
  for (int i = 0; i < per; ++i) {
    int k = b + i;
    if (k < n) v += part[k];
  }
  return block_sum(v, sh);
}

}  // namespace lsamd
