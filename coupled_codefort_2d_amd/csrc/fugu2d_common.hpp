// fugu2d_common.hpp -- device-side grid geometry, padded-array indexing, launch and
// deterministic-reduction helpers shared by every FUGU-2D HIP kernel (gfx950).
//
// Layout: the reference's Fortran padded column-major arrays are kept byte for byte in
// HBM (i fastest), so a host array crosses the C-ABI with one copy and no transpose:
//   u-kind  (0-nh:nx+nh, 1-nh:ny+nh)   ld = nx+1+2nh   (x-face velocities, rho_u, ...)
//   v-kind  (1-nh:nx+nh, 0-nh:ny+nh)   ld = nx+2nh     (y-face velocities, rho_v, ...)
//   p-kind  (1-nh:nx+nh, 1-nh:ny+nh)   ld = nx+2nh     (cell centres: p, C, rho_c, kappa)
//   w-kind  (0-nh:nx+nh, 0-nh:ny+nh)   ld = nx+1+2nh   (vertices: zeta, eta)
// A Krylov "triplet" (u, v, p) is one contiguous allocation [u | v | p] so that the
// reference's whole-array updates are single streaming passes.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace fugu {

struct Grid {
  int nx, ny, nh;
  int per_x, per_y;
  int ldu, ldv, ldp, ldw;     // leading dimensions
  int ru, rv, rp, rw;         // number of rows of each kind
  double dx, dy, inv_dx, inv_dy, inv_dx2, inv_dy2, vol;
  size_t nu, nv, np, nw;      // element counts
};

inline Grid make_grid(int nx, int ny, int nh, double dx, double dy, int per_x, int per_y) {
  Grid g;
  g.nx = nx; g.ny = ny; g.nh = nh; g.per_x = per_x; g.per_y = per_y;
  g.ldu = nx + 1 + 2 * nh; g.ldv = nx + 2 * nh; g.ldp = nx + 2 * nh; g.ldw = nx + 1 + 2 * nh;
  g.ru = ny + 2 * nh; g.rv = ny + 1 + 2 * nh; g.rp = ny + 2 * nh; g.rw = ny + 1 + 2 * nh;
  g.dx = dx; g.dy = dy;
  g.inv_dx = 1.0 / dx; g.inv_dy = 1.0 / dy;                 // src/mod_operators.f90:43-44
  g.inv_dx2 = 1.0 / (dx * dx); g.inv_dy2 = 1.0 / (dy * dy); // src/mod_subsolvers.f90:58-59
  g.vol = dx * dy;                                          // src/mod_bicgstab2.f90:41
  g.nu = (size_t)g.ldu * g.ru; g.nv = (size_t)g.ldv * g.rv;
  g.np = (size_t)g.ldp * g.rp; g.nw = (size_t)g.ldw * g.rw;
  return g;
}

// Fortran (i, j) -> linear offset, one per field kind.
__host__ __device__ inline size_t IU(const Grid& g, int i, int j) { return (size_t)(i + g.nh) + (size_t)(j - 1 + g.nh) * g.ldu; }
__host__ __device__ inline size_t IV(const Grid& g, int i, int j) { return (size_t)(i - 1 + g.nh) + (size_t)(j + g.nh) * g.ldv; }
__host__ __device__ inline size_t IP(const Grid& g, int i, int j) { return (size_t)(i - 1 + g.nh) + (size_t)(j - 1 + g.nh) * g.ldp; }
__host__ __device__ inline size_t IW(const Grid& g, int i, int j) { return (size_t)(i + g.nh) + (size_t)(j + g.nh) * g.ldw; }

// Thread tile: 64 consecutive i (one wavefront, coalesced) x 4 rows.
constexpr int TX = 64;
constexpr int TY = 4;
constexpr int NT = TX * TY;

// Launch over the union of all padded kinds: i in [-nh, nx+nh], j in [-nh, ny+nh].
inline dim3 grid_union(const Grid& g) {
  return dim3((unsigned)((g.ldw + TX - 1) / TX), (unsigned)((g.rw + TY - 1) / TY));
}
// Launch over interior cells i in [1, nx], j in [1, ny] (plus an optional margin m).
inline dim3 grid_cells(const Grid& g, int m = 0) {
  return dim3((unsigned)((g.nx + 2 * m + TX - 1) / TX), (unsigned)((g.ny + 2 * m + TY - 1) / TY));
}

// XCD-aware tile order: consecutive workgroup ids are dealt round-robin to the 8 XCDs;
// remap them so each XCD walks a contiguous band of rows and reuses its own L2.
__device__ inline void tile_coords(int& bx, int& by) {
  const int gx = gridDim.x, gy = gridDim.y;
  const int nb = gx * gy;
  const int bid = blockIdx.x + blockIdx.y * gx;
  int t = bid;
  if (nb >= 64 && (nb % 8) == 0) {
    const int per = nb / 8;
    t = (bid % 8) * per + bid / 8;
  }
  bx = t % gx;
  by = t / gx;
}

// Union-grid point of this thread as Fortran (i, j), with i in [-nh..], j in [-nh..].
__device__ inline void union_ij(const Grid& g, int& i, int& j) {
  int bx, by;
  tile_coords(bx, by);
  i = bx * TX + (int)threadIdx.x - g.nh;
  j = by * TY + (int)threadIdx.y - g.nh;
}
// Interior-cell point (i, j) in [1-m, nx+m] x [1-m, ny+m].
__device__ inline void cell_ij(int m, int& i, int& j) {
  int bx, by;
  tile_coords(bx, by);
  i = bx * TX + (int)threadIdx.x + 1 - m;
  j = by * TY + (int)threadIdx.y + 1 - m;
}

// ---------------------------------------------------------------------------------
// Deterministic reductions: wave shuffle -> LDS -> one partial per workgroup; a second
// single-workgroup kernel folds the partials in a fixed order (two-pass scheme below).

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Sum NV values over the workgroup; thread 0 gets the totals.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double (*lds)[NV]) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int lane = tid & 63, wid = tid >> 6;
  const int nw = (blockDim.x * blockDim.y + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[wid][k] = v[k];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = lds[0][k];
      for (int w = 1; w < nw; ++w) s += lds[w][k];
      v[k] = s;
    }
  }
  __syncthreads();
}

template <int NV>
__device__ inline void block_max(double (&v)[NV], double (*lds)[NV]) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int lane = tid & 63, wid = tid >> 6;
  const int nw = (blockDim.x * blockDim.y + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_max(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[wid][k] = v[k];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = lds[0][k];
      for (int w = 1; w < nw; ++w) s = fmax(s, lds[w][k]);
      v[k] = s;
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------------
// Compensated (double-double) reductions: the fast "tree" summation order.  Every FP64 sum
// of the solver is accumulated as an unevaluated pair (hi, lo) (TwoSum per addition, then a
// renormalisation), so the total is accurate to about n u^2 sum|x| and almost always rounds
// to the same double whatever the summation order.  It is not a correctly rounded sum: under
// heavy cancellation, or when the exact sum lies within that error of a rounding boundary,
// two orders can give neighbouring doubles.  Needs -ffp-contract=off (no FMA inside TwoSum).
// The reference's own order (sequential, j outer, i inner) is the Sink path below.
struct DD { double hi, lo; };

__host__ __device__ inline DD dd_add(DD a, DD b) {
  const double s = a.hi + b.hi;
  const double v = s - a.hi;
  double e = (a.hi - (s - v)) + (b.hi - v);
  e = e + (a.lo + b.lo);
  DD r;
  r.hi = s + e;
  r.lo = e - (r.hi - s);
  return r;
}
__host__ __device__ inline DD dd_from(double x) { DD r; r.hi = x; r.lo = 0.0; return r; }
__host__ __device__ inline double dd_val(DD a) { return a.hi + a.lo; }

__device__ inline DD wave_dd(DD v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    DD w;
    w.hi = __shfl_xor(v.hi, o, 64);
    w.lo = __shfl_xor(v.lo, o, 64);
    // combine in lane order so both lanes of a pair hold the same value
    const int lane = (int)((threadIdx.x + threadIdx.y * blockDim.x) & 63);
    v = ((lane & o) == 0) ? dd_add(v, w) : dd_add(w, v);
  }
  return v;
}

template <int NV>
__device__ inline void block_dd(DD (&v)[NV], DD (*lds)[NV]) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int lane = tid & 63, wid = tid >> 6;
  const int nw = (blockDim.x * blockDim.y + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_dd(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[wid][k] = v[k];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      DD s = lds[0][k];
      for (int w = 1; w < nw; ++w) s = dd_add(s, lds[w][k]);
      v[k] = s;
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------------
// Two-pass reductions.  Pass 1: a capped grid (<= RED_MAXB workgroups, rows strided)
// reduces to one partial per workgroup -- no fences, no atomics.  Pass 2: a single
// workgroup (k_fold_*) folds the partials in a fixed order and runs the scalar logic.
// The kernel boundary orders the two passes across all XCDs.
constexpr int RED_MAXB = 1024;

inline dim3 cap_rows(dim3 gr) {
  const unsigned cap = (unsigned)(RED_MAXB / (gr.x > 0 ? gr.x : 1));
  if (gr.y > (cap > 0 ? cap : 1)) gr.y = cap > 0 ? cap : 1;
  return gr;
}
inline dim3 grid_union_red(const Grid& g) { return cap_rows(grid_union(g)); }
inline dim3 grid_cells_red(const Grid& g) { return cap_rows(grid_cells(g)); }
inline unsigned nblocks(dim3 gr) { return gr.x * gr.y; }

template <int NV>
__device__ inline void publish_dd(DD (&v)[NV], double* partials) {
  __shared__ DD lds[16][NV];
  block_dd<NV>(v, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const int bid = blockIdx.x + blockIdx.y * gridDim.x;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      partials[((size_t)bid * NV + k) * 2] = v[k].hi;
      partials[((size_t)bid * NV + k) * 2 + 1] = v[k].lo;
    }
  }
}
template <int NV>
__device__ inline void publish_max(double (&v)[NV], double* partials) {
  __shared__ double lds[16][NV];
  block_max<NV>(v, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const int bid = blockIdx.x + blockIdx.y * gridDim.x;
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)bid * NV + k] = v[k];
  }
}
// Single-workgroup fold; every thread returns the totals.
template <int NV>
__device__ inline void fold_dd(const double* partials, int nb, double (&tot)[NV]) {
  __shared__ DD lds[16][NV];
  __shared__ double out[NV];
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthr = blockDim.x * blockDim.y;
  DD acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = dd_from(0.0);
  for (int b = tid; b < nb; b += nthr) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      DD p;
      p.hi = partials[((size_t)b * NV + k) * 2];
      p.lo = partials[((size_t)b * NV + k) * 2 + 1];
      acc[k] = dd_add(acc[k], p);
    }
  }
  block_dd<NV>(acc, lds);
  if (tid == 0)
    for (int k = 0; k < NV; ++k) out[k] = dd_val(acc[k]);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = out[k];
}
// ---------------------------------------------------------------------------------
// Reference summation order.  The reference adds every global sum one term at a time,
// `loc = loc + term` from zero, j outer and i inner (u faces, then v faces, then cells for a
// coupled dot: src/mod_bicgstab2.f90:51-69; cells for the PCG sums: src/mod_subsolvers.f90:
// 93-97,108-112,152-156,165-186,194-201,209-214,228-233).  Floating-point addition is not
// associative, so only that exact sequence of roundings gives the reference's double.
//
// Sink: where a pass-1 kernel sends its terms.  n == 0: tree order -- per-workgroup
// double-double partials at p (publish_dd / fold_dd).  n > 0: reference order -- term k of chain
// q (k = its position in the reference's loop) is stored at p[q * n + k]; the sequential-sum
// pipeline (fugu2d_seqsum.hpp) then leaves chain q's sum, bit-identical to the reference's
// one-by-one loop, at tot[q].  The terms are computed in parallel with the reference's operation
// order.
struct Sink {
  double* p;
  long n;
  double* tot;
};

template <int NV>
__device__ inline void sink_add(const Sink& s, DD (&v)[NV], int q, long pos, double term) {
  if (s.n) s.p[(long)q * s.n + pos] = term;
  else v[q] = dd_add(v[q], dd_from(term));
}
template <int NV>
__device__ inline void sink_publish(const Sink& s, DD (&v)[NV]) {
  if (!s.n) publish_dd<NV>(v, s.p);   // s.n is uniform: every thread takes the same branch
}

// Fold of a Sink: nb per-workgroup partials (tree) or the reference-order sums.
template <int NV>
__device__ inline void fold_sink(const Sink& s, int nb, double (&tot)[NV]) {
  if (s.n) {
#pragma unroll
    for (int q = 0; q < NV; ++q) tot[q] = s.tot[q];
  } else {
    fold_dd<NV>(s.p, nb, tot);
  }
}

template <int NV>
__device__ inline void fold_max(const double* partials, int nb, double (&tot)[NV]) {
  __shared__ double lds[16][NV];
  __shared__ double out[NV];
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthr = blockDim.x * blockDim.y;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int b = tid; b < nb; b += nthr)
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = fmax(acc[k], partials[(size_t)b * NV + k]);
  block_max<NV>(acc, lds);
  if (tid == 0)
    for (int k = 0; k < NV; ++k) out[k] = acc[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) tot[k] = out[k];
}

}  // namespace fugu
