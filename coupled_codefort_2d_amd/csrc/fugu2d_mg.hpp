// fugu2d_mg.hpp -- geometric multigrid V-cycle for the pressure block of the PCD
// preconditioner (the variable-density operator A_p phi = sum_f c_f (phi_i - phi_nb) with
// c_f = h^-2 / rho_f, zero-flux walls, periodic wrap), used by the chosen fast solve
// forms in place of the reference's 150-iteration Jacobi-PCG (src/mod_subsolvers.f90:32-246).
//
// Cell-centred hierarchy on compact arrays (no ghosts).  Coarse cell I aggregates fine cells
// 2I and 2I+1 (the last aggregate of an odd extent has one column or row); coarsening goes on
// while both extents are >= 8 and every periodic extent is even.  A level's operator is
//     (A phi)_i = m_i phi_i + sum_f c_f (phi_i - phi_nb(f))
// with a mass term m (zero for the pressure block) and face conductances c; at a
// non-periodic edge the boundary face's conductance acts as a Dirichlet condition with a zero
// ghost (it enters the diagonal only; zero for the Neumann pressure block).  The same
// machinery serves the velocity blocks of the PCD (the u and v Helmholtz operators), where
// m = rho/dt.  Coarse operator: the Galerkin product P^T A P of the piecewise-constant
// prolongation with its conductance part halved (the rediscretisation factor of a doubled
// spacing): m_H = sum of the children's m, c_H = (sum of the fine faces on the coarse face) / 2.
// Restriction sums the children's residuals (fused with the residual).  Smoother: red-black
// Gauss-Seidel, checkerboard-compact launches (each thread owns one cell of the colour).
// The coarsest level is solved by one workgroup in LDS.  With a zero
// initial guess and fixed sweep counts the V-cycle is a fixed linear operator, which is
// what keeps the outer Krylov recurrences consistent.
#pragma once
#include "fugu2d_common.hpp"

namespace fugu {

struct MgDims {
  int nx, ny, per_x, per_y;
};

constexpr int MG_COARSE_MAX = 1024;   // coarsest level solved in one workgroup
constexpr int MG_DENSE_MAX = 64;      // coarsest level solved by its dense inverse

inline MgDims mg_coarse_dims(const MgDims& f) { return MgDims{(f.nx + 1) / 2, (f.ny + 1) / 2, f.per_x, f.per_y}; }
inline bool mg_odd_periodic(const MgDims& d) { return (d.per_x && (d.nx & 1)) || (d.per_y && (d.ny & 1)); }
// Red-black colouring needs even periodic extents (an odd one puts two neighbours across the
// seam in one colour), so a level with an odd periodic extent is only ever the coarsest one,
// and then only when the dense solve takes it.
inline bool mg_can_coarsen(const MgDims& d) {
  if (d.nx < 8 || d.ny < 8 || mg_odd_periodic(d)) return false;
  const MgDims c = mg_coarse_dims(d);
  return !mg_odd_periodic(c) || c.nx * c.ny <= MG_DENSE_MAX;
}

__device__ inline int mg_wrap(int a, int n, int per) {
  if (a < 0) return per ? a + n : 0;
  if (a >= n) return per ? a - n : n - 1;
  return a;
}

// Level-0 conductances from the face densities: cx[(nx+1) x ny], cy[nx x (ny+1)].
template <typename T>
__global__ void k_mg_coef0(Grid g, const T* __restrict__ rho_u, const T* __restrict__ rho_v,
                           T* __restrict__ cx, T* __restrict__ cy) {
  const int fi = blockIdx.x * blockDim.x + threadIdx.x;
  const int fj = blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = g.nx, ny = g.ny;
  // a periodic direction has one value per wrapped face: face n is face 0
  if (fi <= nx && fj < ny) {
    double c = 0.0;
    const int ff = (g.per_x && fi == nx) ? 0 : fi;
    if (g.per_x || (fi > 0 && fi < nx)) c = g.inv_dx2 / (double)rho_u[IU(g, ff, fj + 1)];
    cx[fi + (size_t)fj * (nx + 1)] = (T)c;
  }
  if (fi < nx && fj <= ny) {
    double c = 0.0;
    const int ff = (g.per_y && fj == ny) ? 0 : fj;
    if (g.per_y || (fj > 0 && fj < ny)) c = g.inv_dy2 / (double)rho_v[IV(g, fi + 1, ff)];
    cy[fi + (size_t)fj * nx] = (T)c;
  }
}

// Coarse mass term: the sum of the children.
template <typename T>
__global__ void k_mg_coarsen_m(MgDims f, const T* __restrict__ mf, T* __restrict__ mc) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const int J = blockIdx.y * blockDim.y + threadIdx.y;
  const int nxc = (f.nx + 1) / 2, nyc = (f.ny + 1) / 2;
  if (I >= nxc || J >= nyc) return;
  T s = T(0);
  for (int q = 0; q < 4; ++q) {
    const int i = 2 * I + (q & 1), j = 2 * J + (q >> 1);
    if (i < f.nx && j < f.ny) s += mf[i + (size_t)j * f.nx];
  }
  mc[I + (size_t)J * nxc] = s;
}

// Coarse conductances: the fine faces on each coarse face, summed and halved.
template <typename T>
__global__ void k_mg_coarsen(MgDims f, const T* __restrict__ cxf, const T* __restrict__ cyf,
                             T* __restrict__ cxc, T* __restrict__ cyc) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const int J = blockIdx.y * blockDim.y + threadIdx.y;
  const int nxc = (f.nx + 1) / 2, nyc = (f.ny + 1) / 2;
  if (I <= nxc && J < nyc) {
    const int fi = (2 * I < f.nx) ? 2 * I : f.nx;     // the last coarse face is the fine edge face
    T s = cxf[fi + (size_t)(2 * J) * (f.nx + 1)];
    if (2 * J + 1 < f.ny) s += cxf[fi + (size_t)(2 * J + 1) * (f.nx + 1)];
    cxc[I + (size_t)J * (nxc + 1)] = s * T(0.5);
  }
  if (I < nxc && J <= nyc) {
    const int fj = (2 * J < f.ny) ? 2 * J : f.ny;
    T s = cyf[2 * I + (size_t)fj * f.nx];
    if (2 * I + 1 < f.nx) s += cyf[2 * I + 1 + (size_t)fj * f.nx];
    cyc[I + (size_t)J * nxc] = s * T(0.5);
  }
}

// Inverse diagonal of a level (m may be null: no mass term).
template <typename T>
__global__ void k_mg_dinv(MgDims d, const T* __restrict__ cx, const T* __restrict__ cy, const T* __restrict__ m,
                          T* __restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= d.nx || j >= d.ny) return;
  T s = cx[i + (size_t)j * (d.nx + 1)] + cx[i + 1 + (size_t)j * (d.nx + 1)] +
        cy[i + (size_t)j * d.nx] + cy[i + (size_t)(j + 1) * d.nx];
  if (m) s += m[i + (size_t)j * d.nx];
  dinv[i + (size_t)j * d.nx] = (s > T(1.0e-14)) ? T(1) / s : T(0);
}

// Level arrays stay below 2^31 elements, so indices are 32-bit.  Every face of a level exists
// (boundary faces too) and wrapped neighbour indices are clamped at a wall, so all eight loads
// are unconditional -- issued together, no branch per load; a boundary face of a non-periodic
// edge couples to a zero ghost (its conductance enters the diagonal only) by a select.
template <typename T>
__device__ inline T mg_nbr_sum(const MgDims& d, const T* cx, const T* cy, const T* phi, int i, int j) {
  const int rx = d.nx + 1, o = j * d.nx;
  const T vW = cx[i + j * rx], vE = cx[i + 1 + j * rx], vS = cy[i + o], vN = cy[i + o + d.nx];
  const int iw = mg_wrap(i - 1, d.nx, d.per_x), ie = mg_wrap(i + 1, d.nx, d.per_x);
  const int js = mg_wrap(j - 1, d.ny, d.per_y), jn = mg_wrap(j + 1, d.ny, d.per_y);
  const T pW = phi[iw + o], pE = phi[ie + o], pS = phi[i + js * d.nx], pN = phi[i + jn * d.nx];
  const T cW = (i > 0 || d.per_x) ? vW : T(0);
  const T cE = (i < d.nx - 1 || d.per_x) ? vE : T(0);
  const T cS = (j > 0 || d.per_y) ? vS : T(0);
  const T cN = (j < d.ny - 1 || d.per_y) ? vN : T(0);
  return cW * pW + cE * pE + cS * pS + cN * pN;
}

// One colour of a red-black Gauss-Seidel sweep; thread (t, j) owns cell i = 2t + ((j + colour) & 1).
// zero_other: the other colour is known to be zero (first half-sweep from a zero guess).
template <typename T>
__global__ void __launch_bounds__(256) k_mg_rbgs(MgDims d, int colour, const T* __restrict__ cx,
                                                 const T* __restrict__ cy, const T* __restrict__ dinv,
                                                 const T* __restrict__ b, T* __restrict__ phi, int zero_other) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (j >= d.ny) return;
  const int i = 2 * t + ((j + colour) & 1);
  if (i >= d.nx) return;
  const size_t o = i + (size_t)j * d.nx;
  const T s = zero_other ? T(0) : mg_nbr_sum(d, cx, cy, phi, i, j);
  phi[o] = (b[o] + s) * dinv[o];
}

// Residual of the fine level summed onto the coarse cells (existing children only).
template <typename T>
__global__ void __launch_bounds__(256) k_mg_restrict(MgDims f, const T* __restrict__ cx, const T* __restrict__ cy,
                                                     const T* __restrict__ m, const T* __restrict__ b,
                                                     const T* __restrict__ phi, T* __restrict__ bc,
                                                     T* __restrict__ phic) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const int J = blockIdx.y * blockDim.y + threadIdx.y;
  const int nxc = (f.nx + 1) / 2, nyc = (f.ny + 1) / 2;
  if (I >= nxc || J >= nyc) return;
  T acc = T(0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = 2 * I + (q & 1), j = 2 * J + (q >> 1);
    if (i >= f.nx || j >= f.ny) continue;
    const size_t o = i + (size_t)j * f.nx;
    const size_t rx = (size_t)(f.nx + 1);
    T diag = cx[i + j * rx] + cx[i + 1 + j * rx] + cy[o] + cy[o + f.nx];
    if (m) diag += m[o];
    acc += b[o] - (diag * phi[o] - mg_nbr_sum(f, cx, cy, phi, i, j));
  }
  bc[I + (size_t)J * nxc] = acc;
  phic[I + (size_t)J * nxc] = T(0);
}

// phi_f += phi_c(parent)
template <typename T>
__global__ void __launch_bounds__(256) k_mg_prolong(MgDims f, const T* __restrict__ phic, T* __restrict__ phi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= f.nx || j >= f.ny) return;
  const size_t o = i + (size_t)j * f.nx;
  phi[o] = phi[o] + phic[(i >> 1) + (size_t)(j >> 1) * ((f.nx + 1) / 2)];
}

// Coarsest level: one workgroup.  The rhs is centred first when the operator is singular
// (project: the Neumann/periodic pressure block).  With at most MG_DENSE_MAX cells the level is
// solved exactly by its dense inverse (k_mg_coarse_inv, once per set-up; for the singular
// block the inverse of A + s 11^T / n, which maps a centred rhs to the zero-mean solution),
// otherwise by `sweeps` LDS-resident red-black sweeps from zero.
// Any workgroup size of at least one wavefront (cells strided over the threads); the rhs mean,
// when centred, is summed by the first wavefront in a fixed order (lane l: cells l, l + 64, ...;
// then across lanes), the same whatever the workgroup size.
template <typename T>
__device__ inline void mg_coarse_solve(const MgDims& d, const T* cx, const T* cy, const T* dinv, const T* cinv,
                                       const T* b, T* phi, int sweeps, int project, T* sphi, T* sb, double* red) {
  const int n = d.nx * d.ny;
  const int nt = (int)blockDim.x;
  for (int t = threadIdx.x; t < n; t += nt) { sb[t] = b[t]; sphi[t] = T(0); }
  __syncthreads();
  if (project) {
    if (threadIdx.x < 64) {
      double my = 0.0;
      for (int t = threadIdx.x; t < n; t += 64) my += (double)sb[t];
      my = wave_sum(my);
      if (threadIdx.x == 0) red[0] = my / (double)n;
    }
    __syncthreads();
    const double m = red[0];
    for (int t = threadIdx.x; t < n; t += nt) sb[t] = sb[t] - (T)m;
    __syncthreads();
  }
  if (cinv) {
    for (int t = threadIdx.x; t < n; t += nt) {
      double s = 0.0;
      for (int q = 0; q < n; ++q) s += (double)cinv[(size_t)t * n + q] * (double)sb[q];
      phi[t] = (T)s;
    }
    __syncthreads();
    return;
  }
  for (int s = 0; s < sweeps; ++s) {
    for (int colour = 0; colour < 2; ++colour) {
      for (int t = threadIdx.x; t < n; t += nt) {
        const int i = t % d.nx, j = t / d.nx;
        if (((i + j) & 1) == colour) sphi[t] = (sb[t] + mg_nbr_sum(d, cx, cy, (const T*)sphi, i, j)) * dinv[t];
      }
      __syncthreads();
    }
  }
  for (int t = threadIdx.x; t < n; t += nt) phi[t] = sphi[t];
  __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(1024) k_mg_coarse(MgDims d, const T* __restrict__ cx, const T* __restrict__ cy,
                                                    const T* __restrict__ dinv, const T* __restrict__ cinv,
                                                    const T* __restrict__ b, T* __restrict__ phi, int sweeps,
                                                    int project) {
  __shared__ T sphi[MG_COARSE_MAX];
  __shared__ T sb[MG_COARSE_MAX];
  __shared__ double red[16];
  mg_coarse_solve(d, cx, cy, dinv, cinv, b, phi, sweeps, project, sphi, sb, red);
}

// Dense inverse of the coarsest operator (n <= MG_DENSE_MAX cells), by Gauss-Jordan
// elimination in FP64 without pivoting (the operator, with the rank-one term when singular, is
// symmetric positive definite).  One workgroup; inv is n x n row-major.
template <typename T>
__global__ void __launch_bounds__(1024) k_mg_coarse_inv(MgDims d, const T* __restrict__ cx, const T* __restrict__ cy,
                                                        const T* __restrict__ m, int project, T* __restrict__ inv) {
  __shared__ double A[MG_DENSE_MAX][2 * MG_DENSE_MAX];
  __shared__ double fcol[MG_DENSE_MAX];
  __shared__ double sdiag;
  const int n = d.nx * d.ny, w = 2 * n, t = threadIdx.x;
  for (int q = t; q < n * w; q += blockDim.x) A[q / w][q % w] = ((q % w) == n + q / w) ? 1.0 : 0.0;
  __syncthreads();
  if (t < n) {
    const int i = t % d.nx, j = t / d.nx;
    const size_t rx = (size_t)(d.nx + 1);
    const double cW = (double)cx[i + j * rx], cE = (double)cx[i + 1 + j * rx];
    const double cS = (double)cy[i + (size_t)j * d.nx], cN = (double)cy[i + (size_t)(j + 1) * d.nx];
    A[t][t] = cW + cE + cS + cN + (m ? (double)m[t] : 0.0);
    // neighbours through interior or periodic faces (boundary faces enter the diagonal only)
    if (i > 0 || d.per_x) A[t][mg_wrap(i - 1, d.nx, d.per_x) + j * d.nx] -= cW;
    if (i < d.nx - 1 || d.per_x) A[t][mg_wrap(i + 1, d.nx, d.per_x) + j * d.nx] -= cE;
    if (j > 0 || d.per_y) A[t][i + mg_wrap(j - 1, d.ny, d.per_y) * d.nx] -= cS;
    if (j < d.ny - 1 || d.per_y) A[t][i + mg_wrap(j + 1, d.ny, d.per_y) * d.nx] -= cN;
  }
  __syncthreads();
  if (project) {
    if (t == 0) {
      double s = 0.0;
      for (int q = 0; q < n; ++q) s += A[q][q];
      sdiag = s / ((double)n * (double)n);
    }
    __syncthreads();
    for (int q = t; q < n * n; q += blockDim.x) A[q / n][q % n] += sdiag;
    __syncthreads();
  }
  for (int k = 0; k < n; ++k) {
    const double piv = A[k][k];
    if (t < n) fcol[t] = A[t][k];   // multipliers, read before column k is eliminated
    __syncthreads();
    for (int q = t; q < w; q += blockDim.x) A[k][q] /= piv;
    __syncthreads();
    for (int q = t; q < n * w; q += blockDim.x) {
      const int r = q / w, col = q % w;
      if (r != k) A[r][col] -= fcol[r] * A[k][col];
    }
    __syncthreads();
  }
  for (int q = t; q < n * n; q += blockDim.x) inv[q] = (T)A[q / n][n + q % n];
}

// ---------------------------------------------------------------------------------
// Fused level kernels (temporal blocking): one launch does a level's NU pre-smoothing sweeps
// and the residual restriction, another the prolongation and NU post-smoothing sweeps, instead
// of 2 NU + 1 launches each.  A workgroup owns a tile (shape below) plus a halo of
// H cells; each thread owns fixed cells of that extended tile (slot k: cell tid + NT k, row
// major) and loads their b and four face conductances into registers once (D^-1 formed from
// them), in one round of independent loads; phi of the extended tile lives in LDS.  Half-sweep h updates its colour
// on the tile grown by H - 1 - h rings, so every value the owner needs is final when read, and
// the owner writes its tile only.  Tiles are numbered so that each XCD gets a contiguous band
// of them (its L2 then holds the shared halos).  No launch reads what it writes: the
// pre-smoothing starts from zero or reads its initial iterate from the other buffer, and the
// post-smoothing writes the other buffer.  A cell's colour is that of its wrapped global index and
// each update performs the one-launch-per-half-sweep smoother's arithmetic with the same
// operands, so the result is bit-identical to it.
// Tile: 64 x TH cells, 16 TH threads (one per coarse cell of the restriction).  TH = 32 (512
// threads) cuts the halo overhead from 1.9x to 1.5x against TH = 16 and is faster at every level
// in both precisions (FP32: 7 instead of 8.5 us on small levels, 31 instead of 37 us at 2048^2;
// with the cell slots grouped by colour also the FP64 post-smoothing at 2048^2, -3 % on the
// solve); a 64 x 64 tile of 1024 threads was slower in FP32 at 2048^2 (one workgroup per CU).
constexpr int MGT_W = 64;
constexpr int MGT_TH = 32;

// A tile's global index g of a level extent n: wrapped (periodic), else clamped into the level
// (a cell outside a wall loads an in-level value it never uses).  A small periodic level can
// see several wraps (a tile wider than the level holds copies of its cells).
// Branch-free (a divergent branch between a slot's loads and the next's makes the next wait for
// them): the wrap is g - n q with q = floor(g / n) from the hardware reciprocal, off by at most
// one for |g|, n < 2^16 (relative error ~2^-22), which the two corrections absorb.
__device__ inline int mgt_index(int g, int n, int per) {
  int r = g - n * (int)floorf((float)g * __builtin_amdgcn_rcpf((float)n));
  r = r < 0 ? r + n : r;
  r = r >= n ? r - n : r;
  return per ? r : min(max(g, 0), n - 1);
}

// Element e of a level array from a 32-bit byte offset (a uniform base and one offset register
// per load; level arrays stay below 4 GiB, which the level set-up checks).
template <typename T>
__device__ inline T mgt_ld(const T* base, unsigned e) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + e * (unsigned)sizeof(T));
}

// Tile of this workgroup: dispatch order places workgroup w on XCD w mod 8; renumber so that
// XCD x owns a contiguous range of tiles.
template <int TH>
__device__ inline void mgt_tile(int& i0, int& j0) {
  const int W = gridDim.x * gridDim.y;
  const int w = blockIdx.x + blockIdx.y * gridDim.x;
  const int per = W / 8, rem = W % 8, x = w % 8, k = w / 8;
  const int t = (x < rem) ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
  i0 = (t % gridDim.x) * MGT_W;
  j0 = (t / gridDim.x) * TH;
}

// A thread's cells of the tile grown by H rings, by colour: slots 0..NSC-1 hold colour-0 cells,
// slots NSC..NS-1 colour-1 cells, so that a half-sweep visits only its own colour's slots with
// every lane busy.  The grown tile is EW (even) wide, so each of its rows holds EW/2 cells of
// either colour, and the colour of a cell, (i + j) & 1 of its wrapped global index, is (li + lj)
// & 1: i0, j0 and 2H are even, and a periodic wrap moves by an even extent.
// The conductances are held unmasked: at a wall the neighbour across the boundary face is a cell
// outside the level, whose LDS value is +0 and never updated, so c * (+0) = +0 is exactly the
// masked face's 0 * (+0) (conductances are finite and >= 0).
template <typename T, int H, int TH>
struct MgtCells {
  static constexpr int TILE_H = TH, NT = 16 * TH;
  static constexpr int EW = MGT_W + 2 * H, EH = TH + 2 * H;
  static constexpr int NSC = ((EW / 2) * EH + NT - 1) / NT, NS = 2 * NSC;
  static_assert(EW % 2 == 0, "the grown tile must be even wide");
  T b[NS], dinv[NS], cW[NS], cE[NS], cS[NS], cN[NS];
  T dg[NS];                 // diagonal (+ m): the residual of the restriction
  int off[NS], ring[NS];    // off: lj * EW + li in the LDS tile; ring: distance outside the tile
                            // (0 inside), -1: no cell or beyond the updated rings
};

// Load the thread's cells (registers) and phi (+ the parent's coarse correction) into LDS.
// All slots' loads are in flight together: nothing in the load loop uses a loaded value (a use
// there waits for every load issued before it).  Only cells up to ring rc are ever updated and
// load b and their conductances; cells beyond it are read as neighbours (their phi) or not at
// all.  D^-1 and the diagonal are formed exactly as k_mg_dinv forms them (no D^-1 array is
// read).  Offsets are 32-bit: one offset register per load against a uniform base.
template <typename T, int H, int TH>
__device__ inline void mgt_init(const MgDims& f, int i0, int j0, const T* cx, const T* cy, const T* mass,
                                const T* b, const T* phi, const T* phic, int rc, MgtCells<T, H, TH>& m, T* s) {
  using M = MgtCells<T, H, TH>;
  const int tid = threadIdx.x;
  const int nx = f.nx, rx = f.nx + 1, nxc = (f.nx + 1) / 2;
  T pv[M::NS], pc[M::NS];
  bool cell[M::NS], slot[M::NS];
#pragma unroll
  for (int k = 0; k < M::NS; ++k) {
    const int colour = k / M::NSC;
    const int q = tid + M::NT * (k - colour * M::NSC);   // the colour's q-th cell of the grown tile
    const int lj = q / (M::EW / 2), li = 2 * (q - lj * (M::EW / 2)) + ((lj + colour) & 1);
    const int gi = i0 - H + li, gj = j0 - H + lj;
    const int i = mgt_index(gi, nx, f.per_x), j = mgt_index(gj, f.ny, f.per_y);
    const int dist = max(max(max(H - li, li - (H + MGT_W - 1)), max(H - lj, lj - (H + M::TILE_H - 1))), 0);
    slot[k] = lj < M::EH;   // the last slots of the last threads can lie past the grown tile
    cell[k] = slot[k] && (f.per_x || (unsigned)gi < (unsigned)nx) && (f.per_y || (unsigned)gj < (unsigned)f.ny);
    m.off[k] = slot[k] ? lj * M::EW + li : 0;
    m.ring[k] = (cell[k] && dist <= rc) ? dist : -1;
    // 24-bit products (level extents stay below 2^24): a full 32-bit multiply-add is a 64-bit
    // instruction whose unused high addend can be a register a load in flight writes, and then
    // waits for every load issued so far
    const unsigned o = (unsigned)i + __umul24((unsigned)j, (unsigned)nx);
    const unsigned ox = (unsigned)i + __umul24((unsigned)j, (unsigned)rx);
    // b and the conductances only where the cell is updated; phi (+ phi_c) wherever it exists
    T bv = T(0), rw = T(0), re = T(0), rs = T(0), rn = T(0), mm = T(0), p = T(0), pp = T(0);
    if (m.ring[k] >= 0) {
      bv = mgt_ld(b, o);
      rw = mgt_ld(cx, ox);
      re = mgt_ld(cx, ox + 1u);
      rs = mgt_ld(cy, o);
      rn = mgt_ld(cy, o + (unsigned)nx);
      if (mass) mm = mgt_ld(mass, o);
    }
    if (slot[k]) {
      if (phi) p = mgt_ld(phi, o);
      if (phic) pp = mgt_ld(phic, (unsigned)(i >> 1) + __umul24((unsigned)(j >> 1), (unsigned)nxc));
    }
    m.b[k] = bv; m.cW[k] = rw; m.cE[k] = re; m.cS[k] = rs; m.cN[k] = rn; m.dg[k] = mm;
    pv[k] = p;
    pc[k] = pp;
  }
  // the prolonged iterate phi + phi_c; +0 outside the level
#pragma unroll
  for (int k = 0; k < M::NS; ++k) {
    const T p = phic ? pv[k] + pc[k] : pv[k];
    pv[k] = cell[k] ? p : T(0);
  }
#pragma unroll
  for (int k = 0; k < M::NS; ++k) {
    T dg = m.cW[k] + m.cE[k] + m.cS[k] + m.cN[k];
    if (mass) dg += m.dg[k];
    m.dg[k] = dg;
    m.dinv[k] = (dg > T(1.0e-14)) ? T(1) / dg : T(0);
  }
#pragma unroll
  for (int k = 0; k < M::NS; ++k)
    if (slot[k]) s[m.off[k]] = pv[k];
  __syncthreads();
}

// One half-sweep of colour COL over the tile grown by r rings (the colour's slots only).
template <typename T, int H, int TH, int COL>
__device__ inline void mgt_half(int r, int zero_other, const MgtCells<T, H, TH>& m, T* s) {
  using M = MgtCells<T, H, TH>;
#pragma unroll
  for (int k = COL * M::NSC; k < (COL + 1) * M::NSC; ++k) {
    if ((unsigned)m.ring[k] > (unsigned)r) continue;   // ring -1: never
    const int o = m.off[k];
    T sum = T(0);
    if (!zero_other) sum = m.cW[k] * s[o - 1] + m.cE[k] * s[o + 1] + m.cS[k] * s[o - M::EW] + m.cN[k] * s[o + M::EW];
    s[o] = (m.b[k] + sum) * m.dinv[k];
  }
  __syncthreads();
}

template <typename T, int H, int TH>
__device__ inline void mgt_store(const MgDims& f, int i0, int j0, T* out, const T* s) {
  using M = MgtCells<T, H, TH>;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int gi = i0 + tx;
  if (gi >= f.nx) return;
#pragma unroll
  for (int qj = ty; qj < TH; qj += TH / 4) {
    const int gj = j0 + qj;
    if (gj < f.ny) out[gi + (size_t)gj * f.nx] = s[(H + qj) * M::EW + H + tx];
  }
}

// NU pre-smoothing sweeps from zero (red first), then (bc set) the residual summed onto the
// coarse cells of the tile (coarse phi zeroed).  The residual of each tile cell is formed by the thread
// holding its coefficients and summed from LDS (children in the restriction's order).
// phi_in (optional): a non-zero initial iterate (another array than phi: tiles read their halos
// from it while others write their own cells of phi); nullptr: from zero.
template <typename T, int NU, int TH>
__global__ void __launch_bounds__(16 * TH) k_mg_pre(MgDims f, const T* __restrict__ cx, const T* __restrict__ cy,
                                                const T* __restrict__ m, const T* __restrict__ b,
                                                T* __restrict__ phi, T* __restrict__ bc, T* __restrict__ phic,
                                                const T* __restrict__ phi_in) {
  constexpr int H = 2 * NU + 1;
  using M = MgtCells<T, H, TH>;
  __shared__ T s[M::EH * M::EW];
  __shared__ T sres[TH][MGT_W];
  int i0, j0;
  mgt_tile<TH>(i0, j0);
  M cells;
  // rings 0..H-1 are updated; from zero, ring H-1 only by the first, neighbour-free half-sweep and
  // ring H is never read; from phi_in, ring H is read (its phi) by the first half-sweep
  mgt_init<T, H, TH>(f, i0, j0, cx, cy, m, b, phi_in, (const T*)nullptr, H - 1, cells, s);
#pragma unroll
  for (int h = 0; h < 2 * NU; ++h) {   // red (colour 0) first
    if (h & 1) mgt_half<T, H, TH, 1>(H - 1 - h, 0, cells, s);
    else mgt_half<T, H, TH, 0>(H - 1 - h, (h == 0 && !phi_in) ? 1 : 0, cells, s);
  }
  mgt_store<T, H, TH>(f, i0, j0, phi, s);
  if (!bc) return;   // smoothing only (a truncated hierarchy's coarsest level)
#pragma unroll
  for (int k = 0; k < M::NS; ++k) {
    if (cells.ring[k] != 0) continue;
    const int o = cells.off[k];
    const T nb = cells.cW[k] * s[o - 1] + cells.cE[k] * s[o + 1] + cells.cS[k] * s[o - M::EW] + cells.cN[k] * s[o + M::EW];
    const int lj = o / M::EW, li = o - lj * M::EW;
    sres[lj - H][li - H] = cells.b[k] - (cells.dg[k] * s[o] - nb);
  }
  __syncthreads();
  const int nxc = (f.nx + 1) / 2, nyc = (f.ny + 1) / 2;
  const int q = threadIdx.x;   // one coarse cell of the 32 x (tile height / 2) coarse tile
  const int I = i0 / 2 + (q & 31), J = j0 / 2 + (q >> 5);
  if (I >= nxc || J >= nyc) return;
  T acc = T(0);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = 2 * I + (k & 1), j = 2 * J + (k >> 1);
    if (i >= f.nx || j >= f.ny) continue;
    acc += sres[j - j0][i - i0];
  }
  bc[I + (size_t)J * nxc] = acc;
  phic[I + (size_t)J * nxc] = T(0);
}

// phi_out = phi + phi_c(parent), then NU post-smoothing sweeps (black first).
// psum (optional): the tile's sum of its outputs as a double-double partial, one per workgroup
// (the mean of the level-0 result without another pass over it).
template <typename T, int NU, int TH>
__global__ void __launch_bounds__(16 * TH) k_mg_post(MgDims f, const T* __restrict__ cx, const T* __restrict__ cy,
                                                 const T* __restrict__ m, const T* __restrict__ b,
                                                 const T* __restrict__ phi, const T* __restrict__ phic,
                                                 T* __restrict__ phi_out, double* __restrict__ psum) {
  constexpr int H = 2 * NU;
  using M = MgtCells<T, H, TH>;
  __shared__ T s[M::EH * M::EW];
  int i0, j0;
  mgt_tile<TH>(i0, j0);
  M cells;
  mgt_init<T, H, TH>(f, i0, j0, cx, cy, m, b, phi, phic, H - 1, cells, s);
#pragma unroll
  for (int h = 0; h < 2 * NU; ++h) {   // black (colour 1) first
    if (h & 1) mgt_half<T, H, TH, 0>(H - 1 - h, 0, cells, s);
    else mgt_half<T, H, TH, 1>(H - 1 - h, 0, cells, s);
  }
  mgt_store<T, H, TH>(f, i0, j0, phi_out, s);
  if (psum) {
    DD v[1] = {dd_from(0.0)};
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    if (i0 + tx < f.nx) {
#pragma unroll
      for (int qj = ty; qj < TH; qj += TH / 4)
        if (j0 + qj < f.ny) v[0] = dd_add(v[0], dd_from((double)s[(H + qj) * M::EW + H + tx]));
    }
    publish_dd<1>(v, psum);
  }
}

// Double-double partial sums of a compact level array (one per workgroup, grid-stride).
template <typename T>
__global__ void __launch_bounds__(256) k_mg_sum(MgDims d, const T* __restrict__ phi, double* __restrict__ partials) {
  DD v[1] = {dd_from(0.0)};
  const size_t n = (size_t)d.nx * d.ny, stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    v[0] = dd_add(v[0], dd_from((double)phi[k]));
  publish_dd<1>(v, partials);
}

// ---------------------------------------------------------------------------------
// Coarse tail: the V-cycle over the coarsest levels (from a level of at most a few thousand
// cells down) in ONE workgroup and one launch, with a barrier between phases instead of a
// launch per half-sweep.  STAGE: the levels' coefficients, rhs and iterates are first copied
// into LDS (when they fit MG_TAIL_POOL bytes), so every phase runs from LDS; otherwise they
// stay in global memory (L2-resident).  The staged levels lie in global memory in the pool's
// own order (blk, nblk elements), so the copy is one contiguous range whose loads are all in
// flight together: the data are cold (the finer levels streamed through the caches since the
// set-up wrote them) and a copy per array would pay the memory latency once per array.  Phase for phase the kernel performs the multi-launch
// cycle's arithmetic (same smoother update, residual sum, prolongation and coarsest solve), so
// its result is bit-identical; only the top level's phi is written back.
constexpr int MG_TAIL_LEVELS = 16;
constexpr int MG_TAIL_POOL = 136 * 1024;
constexpr int MG_TAIL_NTH = 1024;   // threads of the tail workgroup
template <typename T>
struct MgTailLevel {
  MgDims d;
  const T *cx, *cy, *m, *dinv, *cinv;
  T *b, *phi;
  int ocx, ocy, om, odinv, ob, ophi;   // offsets in the LDS pool (elements; om < 0: no mass term)
};
template <typename T>
struct MgTailArgs {
  MgTailLevel<T> L[MG_TAIL_LEVELS];
  const T* blk;   // STAGE: the pool's image in global memory (the top level's phi excepted)
  int nblk;       // its length in elements
  int nlev, nu, coarse_sweeps, project, zero_guess;
  // register-resident tail: the quads' register images (TqQuad<T>, quad t of level l at
  // l * MG_TAILQ_NTH + t) and the coarsest level's LDS image, written by k_mg_tail_q_pack
  const void* tq;
  const void* tqc;
};

// log2 of the lanes a tail phase puts along a row of w cells: the smallest power of two >= w, at
// most 64 (the rest of the threads take further rows), so that a level with no more cells per
// phase than threads gives every thread at most one.
__device__ inline int tail_lx_log2(int w) {
  int s = 0;
  while ((1 << s) < w && s < 6) ++s;
  return s;
}

// One colour of a red-black sweep in the tail (NT threads).  The cells of one colour depend only
// on the other colour, so the result does not depend on how they are spread over the threads.
// phic (the first post-smoothing half-sweep): the neighbours' values are phi + phi_c(parent), the
// prolongation done on the fly -- the only reader of the prolonged values of this colour's
// neighbours, whose own cells are overwritten before anything else reads them.
template <typename T, int NT>
__device__ inline void tail_smooth(const MgDims d, const T* cx, const T* cy, const T* dinv, const T* b, T* phi,
                                   int colour, int zero_other, const T* phic = nullptr) {
  const int half = (d.nx + 1) / 2;
  const int sx = tail_lx_log2(half);
  const int tx = threadIdx.x & ((1 << sx) - 1), ty = threadIdx.x >> sx;
  const int nxc = (d.nx + 1) / 2;
  for (int j = ty; j < d.ny; j += NT >> sx) {
    for (int t = tx; t < half; t += 1 << sx) {
      const int i = 2 * t + ((j + colour) & 1);
      if (i >= d.nx) continue;
      const int o = i + j * d.nx;
      T s = T(0);
      if (!zero_other) {
        if (phic) {
          const int rx = d.nx + 1, ro = j * d.nx;
          const T vW = cx[i + j * rx], vE = cx[i + 1 + j * rx], vS = cy[i + ro], vN = cy[i + ro + d.nx];
          const int iw = mg_wrap(i - 1, d.nx, d.per_x), ie = mg_wrap(i + 1, d.nx, d.per_x);
          const int js = mg_wrap(j - 1, d.ny, d.per_y), jn = mg_wrap(j + 1, d.ny, d.per_y);
          auto pv = [&](int ii, int jj) -> T { return phi[ii + jj * d.nx] + phic[(ii >> 1) + (jj >> 1) * nxc]; };
          const T pW = pv(iw, j), pE = pv(ie, j), pS = pv(i, js), pN = pv(i, jn);
          const T cW = (i > 0 || d.per_x) ? vW : T(0);
          const T cE = (i < d.nx - 1 || d.per_x) ? vE : T(0);
          const T cS = (j > 0 || d.per_y) ? vS : T(0);
          const T cN = (j < d.ny - 1 || d.per_y) ? vN : T(0);
          s = cW * pW + cE * pE + cS * pS + cN * pN;
        } else {
          s = mg_nbr_sum(d, cx, cy, (const T*)phi, i, j);
        }
      }
      phi[o] = (b[o] + s) * dinv[o];
    }
  }
  __syncthreads();
}

// Residual restriction of a tail level: b_c = sum of the children's residuals, phi_c = 0.
template <typename T, int NT>
__device__ inline void tail_restrict(const MgDims f, const T* cx, const T* cy, const T* m, const T* b, const T* phi,
                                     T* bc, T* phic) {
  const int nxc = (f.nx + 1) / 2, nyc = (f.ny + 1) / 2;
  const int sx = tail_lx_log2(nxc);
  const int tx = threadIdx.x & ((1 << sx) - 1), ty = threadIdx.x >> sx;
  for (int J = ty; J < nyc; J += NT >> sx) {
    for (int I = tx; I < nxc; I += 1 << sx) {
      T acc = T(0);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int i = 2 * I + (k & 1), j = 2 * J + (k >> 1);
        if (i >= f.nx || j >= f.ny) continue;
        const size_t o = i + (size_t)j * f.nx;
        const size_t rx = (size_t)(f.nx + 1);
        T diag = cx[i + j * rx] + cx[i + 1 + j * rx] + cy[o] + cy[o + f.nx];
        if (m) diag += m[o];
        acc += b[o] - (diag * phi[o] - mg_nbr_sum(f, cx, cy, phi, i, j));
      }
      bc[I + (size_t)J * nxc] = acc;
      phic[I + (size_t)J * nxc] = T(0);
    }
  }
  __syncthreads();
}

template <typename T>
__device__ inline void tail_copy(T* dst, const T* src, int n) {
  for (int q = (int)threadIdx.x; q < n; q += (int)blockDim.x) dst[q] = src[q];
}

template <typename T, bool STAGE, int NTH>
__global__ void __launch_bounds__(NTH) k_mg_tail(MgTailArgs<T> a) {
  __shared__ T sphi[MG_COARSE_MAX];
  __shared__ T sb[MG_COARSE_MAX];
  __shared__ double red[16];
  __shared__ T pool[STAGE ? MG_TAIL_POOL / sizeof(T) : 1];
  const int nu = a.nu;
  auto CX = [&](int l) -> const T* { return STAGE ? pool + a.L[l].ocx : a.L[l].cx; };
  auto CY = [&](int l) -> const T* { return STAGE ? pool + a.L[l].ocy : a.L[l].cy; };
  auto MM = [&](int l) -> const T* { return STAGE ? (a.L[l].om >= 0 ? pool + a.L[l].om : nullptr) : a.L[l].m; };
  auto DI = [&](int l) -> const T* { return STAGE ? pool + a.L[l].odinv : a.L[l].dinv; };
  auto BB = [&](int l) -> T* { return STAGE ? pool + a.L[l].ob : a.L[l].b; };
  auto PH = [&](int l) -> T* { return STAGE ? pool + a.L[l].ophi : a.L[l].phi; };
  if (STAGE) {
    // every load issued before the first LDS store (NTH threads, a fixed number per thread);
    // the top level's phi comes from its own array, or is zero for a zero guess (finite values
    // are needed even then: at a wall the smoother's neighbour sum reads the cell itself times a
    // zero conductance)
    constexpr int KMAX = (MG_TAIL_POOL / (int)sizeof(T) + NTH - 1) / NTH;
    const int p0 = a.L[0].ophi, n0 = a.L[0].d.nx * a.L[0].d.ny;
    T v[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const int q = (int)threadIdx.x + NTH * k;
      const int r = q - p0;
      v[k] = T(0);
      if (q < a.nblk) {
        if (r < 0 || r >= n0)
          v[k] = a.blk[q];
        else if (!a.zero_guess)
          v[k] = a.L[0].phi[r];
      }
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const int q = (int)threadIdx.x + NTH * k;
      if (q < a.nblk) pool[q] = v[k];
    }
    __syncthreads();
  }
  // descent: pre-smoothing and residual restriction
  for (int l = 0; l + 1 < a.nlev; ++l) {
    const MgDims f = a.L[l].d;
    const T *cx = CX(l), *cy = CY(l), *m = MM(l), *dinv = DI(l);
    const T* b = BB(l);
    T* phi = PH(l);
    const int zg = (l == 0) ? a.zero_guess : 1;
    for (int s = 0; s < nu; ++s) {
      tail_smooth<T, NTH>(f, cx, cy, dinv, b, phi, 0, (s == 0 && zg) ? 1 : 0);
      tail_smooth<T, NTH>(f, cx, cy, dinv, b, phi, 1, 0);
    }
    tail_restrict<T, NTH>(f, cx, cy, m, b, (const T*)phi, BB(l + 1), PH(l + 1));
  }
  // coarsest level
  {
    const int l = a.nlev - 1;
    mg_coarse_solve(a.L[l].d, CX(l), CY(l), DI(l), a.L[l].cinv, (const T*)BB(l), PH(l), a.coarse_sweeps, a.project,
                    sphi, sb, red);
  }
  // ascent: post-smoothing, the prolongation folded into its first half-sweep
  for (int l = a.nlev - 2; l >= 0; --l) {
    const MgDims f = a.L[l].d;
    const T *cx = CX(l), *cy = CY(l), *dinv = DI(l);
    const T* b = BB(l);
    T* phi = PH(l);
    const T* phic = PH(l + 1);
    for (int s = 0; s < nu; ++s) {
      tail_smooth<T, NTH>(f, cx, cy, dinv, b, phi, 1, 0, s == 0 ? phic : (const T*)nullptr);
      tail_smooth<T, NTH>(f, cx, cy, dinv, b, phi, 0, 0);
    }
  }
  if (STAGE) {
    const MgDims d0 = a.L[0].d;
    tail_copy(a.L[0].phi, (const T*)(pool + a.L[0].ophi), d0.nx * d0.ny);
  }
}

// ---------------------------------------------------------------------------------
// Coarse tail with the coefficients in registers (levels of even extents, at most
// 4 MG_TAILQ_NTH cells, at most MG_TAILQ_LEVELS of them above the coarsest).  The general tail
// above re-derives every cell's neighbours, wraps and masks and re-reads its coefficients from
// LDS in each half-sweep: about 70 instructions and 9 LDS reads per update, which on one
// workgroup is the whole cost.  Here thread t owns quad t of every level -- the 2 x 2 cells
// (2I + (k & 1), 2J + (k >> 1)), k = 0..3, that restrict onto coarse cell t = I + J nx/2 --
// and holds their masked conductances, D^-1 and raw diagonals in registers with the LDS
// indices of the neighbours outside the quad (those inside are at fixed offsets).  An update
// is then 5 LDS reads, the stencil and a write.  Owning all four children of its coarse cell,
// a thread forms their residuals and sums them onto the coarse cell in one phase (no residual
// array, no second barrier).  Only b and phi of the levels and the coarsest level's operator
// live in LDS.  The registers' images are built once per set-up of the levels
// (k_mg_tail_q_pack) and read back with plain loads at each V-cycle.
// The arithmetic is the general tail's, operand for operand: the neighbour sum
// cW pW + cE pE + cS pS + cN pN with clamped (wall) or wrapped (periodic) neighbour indices and
// masked wall conductances, (b + s) D^-1, the residual b - (diag phi - s) summed over the
// children in the order k = 0..3, the prolonged neighbour values phi + phi_c(parent) in the
// first post-smoothing half-sweep, and the same coarsest solve: the results are bit-identical.
constexpr int MG_TAILQ_NTH = 512;     // threads: one quad of the top level each
constexpr int MG_TAILQ_LEVELS = 3;    // smoothing levels (the coarsest not counted)
constexpr int MG_TAILQ_POOL = 8192;   // LDS elements for b and phi of every tail level
// LDS elements for the coarsest level's operator: its dense inverse (transposed), or c_x, c_y
// and D^-1 for the red-black sweeps
constexpr int MG_TAILQ_COARSE = MG_DENSE_MAX * MG_DENSE_MAX + 128;
inline size_t mg_tailq_coarse_elems(const MgDims& d, bool dense) {
  const size_t n = (size_t)d.nx * d.ny;
  return dense ? n * n : (size_t)(d.nx + 1) * d.ny + (size_t)d.nx * (d.ny + 1) + n;
}

template <typename T>
struct alignas(16) TqQuad {
  T cw[4], ce[4], cs[4], cn[4], dinv[4], diag[4];
  // the two neighbours of cell k outside the quad, LDS element indices: horizontal | vertical
  // << 16 (k = 0: W, S; 1: E, S; 2: W, N; 3: E, N)
  unsigned out[4];
  // cells 1 and 2 (the colour the post-smoothing updates first): the coarse parents of those two
  // neighbours, packed the same way (the parent of a neighbour inside the quad is coarse cell t)
  unsigned pout[2];
  unsigned own0;   // the LDS index of cell 0
};

__device__ inline unsigned tq_lo(unsigned p) { return p & 0xffffu; }
__device__ inline unsigned tq_hi(unsigned p) { return p >> 16; }

// t = I + J w (a quad of a level w quads wide; an element of an n x n matrix): J = floor((t +
// 1/2) / w) in single precision instead of an integer division (tens of instructions each, in a
// set-up whose instruction count is its cost).  (t + 1/2) / w lies at least 1/(2w) from an
// integer; the hardware reciprocal and the product err by less than 2^-22 relative, i.e. by
// less than (t + 1/2) 2^-22 / w, which is below 1/(2w) for every t < 2^21.
__device__ inline void tq_quad(int t, int w, int& I, int& J) {
  J = (int)(((float)t + 0.5f) * __builtin_amdgcn_rcpf((float)w));
  I = t - J * w;
}

// Loads of quad (I, J)'s coefficients (level arrays in global memory, 32-bit offsets against
// the level's base); no loaded value is used here, so every load of every level is in flight
// together.
template <typename T>
__device__ inline void tq_load(const MgTailLevel<T>& L, int I, int J, unsigned phi0, unsigned phic0, TqQuad<T>& q,
                               T (&mm)[4]) {
  const MgDims d = L.d;
  const int hx = d.nx >> 1;
  const unsigned rx = (unsigned)d.nx + 1u, nx = (unsigned)d.nx;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = 2 * I + (k & 1), j = 2 * J + (k >> 1);
    const unsigned o = (unsigned)i + (unsigned)j * nx, ox = (unsigned)i + (unsigned)j * rx;
    q.cw[k] = mgt_ld(L.cx, ox);
    q.ce[k] = mgt_ld(L.cx, ox + 1u);
    q.cs[k] = mgt_ld(L.cy, o);
    q.cn[k] = mgt_ld(L.cy, o + nx);
    q.dinv[k] = mgt_ld(L.dinv, o);
    mm[k] = L.m ? mgt_ld(L.m, o) : T(0);
    // the outer neighbours (mg_nbr_sum's wrapped or clamped indices) and their parents
    const int ix = (k & 1) ? mg_wrap(i + 1, d.nx, d.per_x) : mg_wrap(i - 1, d.nx, d.per_x);
    const int jy = (k >> 1) ? mg_wrap(j + 1, d.ny, d.per_y) : mg_wrap(j - 1, d.ny, d.per_y);
    q.out[k] = (phi0 + (unsigned)(ix + j * d.nx)) | ((phi0 + (unsigned)(i + jy * d.nx)) << 16);
    if (k == 1 || k == 2)
      q.pout[k - 1] = (phic0 + (unsigned)((ix >> 1) + (j >> 1) * hx)) |
                      ((phic0 + (unsigned)((i >> 1) + (jy >> 1) * hx)) << 16);
  }
}

// After the loads: raw diagonal (the residual's, k_mg_dinv's sum) and masked conductances.
template <typename T>
__device__ inline void tq_finish(const MgDims& d, int I, int J, TqQuad<T>& q, const T (&mm)[4], bool mass) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = 2 * I + (k & 1), j = 2 * J + (k >> 1);
    T dg = q.cw[k] + q.ce[k] + q.cs[k] + q.cn[k];
    if (mass) dg += mm[k];
    q.diag[k] = dg;
    if (!(i > 0 || d.per_x)) q.cw[k] = T(0);
    if (!(i < d.nx - 1 || d.per_x)) q.ce[k] = T(0);
    if (!(j > 0 || d.per_y)) q.cs[k] = T(0);
    if (!(j < d.ny - 1 || d.per_y)) q.cn[k] = T(0);
  }
}

// Neighbour values of cell k (own: its LDS index; nx: the level's row length).  PRO: plus the
// coarse correction of each neighbour's parent (pc: coarse cell t's index, for neighbours in
// the quad).
template <typename T, int K, bool PRO>
__device__ inline T tq_nbr(const TqQuad<T>& q, unsigned own, int nx, unsigned pc, const T* s) {
  const unsigned ox = tq_lo(q.out[K]), oy = tq_hi(q.out[K]);
  const unsigned w = (K & 1) ? own - 1u : ox, e = (K & 1) ? ox : own + 1u;
  const unsigned so = (K >> 1) ? own - (unsigned)nx : oy, no = (K >> 1) ? oy : own + (unsigned)nx;
  T pW = s[w], pE = s[e], pS = s[so], pN = s[no];
  if constexpr (PRO) {
    const unsigned px = tq_lo(q.pout[K - 1]), py = tq_hi(q.pout[K - 1]);
    pW = pW + s[(K & 1) ? pc : px];
    pE = pE + s[(K & 1) ? px : pc];
    pS = pS + s[(K >> 1) ? pc : py];
    pN = pN + s[(K >> 1) ? py : pc];
  }
  return q.cw[K] * pW + q.ce[K] * pE + q.cs[K] * pS + q.cn[K] * pN;
}

// Half-sweep of colour COL (cells 0 and 3: colour 0; 1 and 2: colour 1) on quad q (own0: the
// LDS index of its cell 0; db: b's index minus phi's in the level).
template <typename T, int COL, bool PRO>
__device__ inline void tq_half(const TqQuad<T>& q, bool act, unsigned own0, int nx, int db, unsigned pc,
                               bool zero_other, T* s) {
  if (act) {
    constexpr int KA = COL ? 1 : 0, KB = COL ? 2 : 3;
    const unsigned oa = own0 + (KA & 1) + (KA >> 1) * (unsigned)nx;
    const unsigned ob = own0 + (KB & 1) + (KB >> 1) * (unsigned)nx;
    T sa = T(0), sb = T(0);
    if (!zero_other) {
      sa = tq_nbr<T, KA, PRO>(q, oa, nx, pc, s);
      sb = tq_nbr<T, KB, PRO>(q, ob, nx, pc, s);
    }
    const T va = (s[oa + db] + sa) * q.dinv[KA];
    const T vb = (s[ob + db] + sb) * q.dinv[KB];
    s[oa] = va;
    s[ob] = vb;
  }
  __syncthreads();
}

// Residuals of the quad's four cells summed onto its coarse cell (bc, phic: LDS indices).
template <typename T>
__device__ inline void tq_restrict(const TqQuad<T>& q, bool act, unsigned own0, int nx, int db, unsigned bc,
                                   unsigned phic, T* s) {
  if (act) {
    T r[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const unsigned o = own0 + (k & 1) + (k >> 1) * (unsigned)nx;
      T sn;
      if (k == 0) sn = tq_nbr<T, 0, false>(q, o, nx, 0u, s);
      else if (k == 1) sn = tq_nbr<T, 1, false>(q, o, nx, 0u, s);
      else if (k == 2) sn = tq_nbr<T, 2, false>(q, o, nx, 0u, s);
      else sn = tq_nbr<T, 3, false>(q, o, nx, 0u, s);
      r[k] = s[o + db] - (q.diag[k] * s[o] - sn);
    }
    T acc = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) acc += r[k];
    s[bc] = acc;
    s[phic] = T(0);
  }
  __syncthreads();
}

// The coarsest level's solve, mg_coarse_solve's arithmetic with the operator in LDS: ci holds
// the dense inverse transposed (element (t, q) at q n + t, so that the threads' reads of one q
// are consecutive), or c_x, c_y and D^-1 back to back.
template <typename T>
__device__ inline void tq_coarse(const MgDims& d, bool dense, const T* ci, const T* b, T* phi, int sweeps,
                                 int project, T* sphi, T* sb, double* red) {
  const int n = d.nx * d.ny;
  const int nt = (int)blockDim.x;
  for (int t = threadIdx.x; t < n; t += nt) { sb[t] = b[t]; sphi[t] = T(0); }
  __syncthreads();
  if (project) {
    if (threadIdx.x < 64) {
      double my = 0.0;
      for (int t = threadIdx.x; t < n; t += 64) my += (double)sb[t];
      my = wave_sum(my);
      if (threadIdx.x == 0) red[0] = my / (double)n;
    }
    __syncthreads();
    const double m = red[0];
    for (int t = threadIdx.x; t < n; t += nt) sb[t] = sb[t] - (T)m;
    __syncthreads();
  }
  if (dense) {
    for (int t = threadIdx.x; t < n; t += nt) {
      double s = 0.0;
      for (int q = 0; q < n; ++q) s += (double)ci[q * n + t] * (double)sb[q];
      phi[t] = (T)s;
    }
    __syncthreads();
    return;
  }
  const T* cx = ci;
  const T* cy = ci + (d.nx + 1) * d.ny;
  const T* dinv = cy + d.nx * (d.ny + 1);
  for (int s = 0; s < sweeps; ++s) {
    for (int colour = 0; colour < 2; ++colour) {
      for (int t = threadIdx.x; t < n; t += nt) {
        const int i = t % d.nx, j = t / d.nx;
        if (((i + j) & 1) == colour) sphi[t] = (sb[t] + mg_nbr_sum(d, cx, cy, (const T*)sphi, i, j)) * dinv[t];
      }
      __syncthreads();
    }
  }
  for (int t = threadIdx.x; t < n; t += nt) phi[t] = sphi[t];
  __syncthreads();
}

// The tail's LDS layout: level l's b at bo[l], its phi at bo[l] + n_l.
template <int NL, typename T>
__device__ inline void tq_layout(const MgTailArgs<T>& a, int (&bo)[NL + 1], int (&nq)[NL]) {
  int off = 0;
#pragma unroll
  for (int l = 0; l <= NL; ++l) {
    bo[l] = off;
    off += 2 * a.L[l].d.nx * a.L[l].d.ny;
  }
#pragma unroll
  for (int l = 0; l < NL; ++l) nq[l] = (a.L[l].d.nx >> 1) * (a.L[l].d.ny >> 1);
}

// The register images of every quad and the coarsest level's LDS image (the dense inverse
// transposed: element p = q n + r holds inverse element (r, q); or c_x, c_y and D^-1 back to
// back), from the level arrays: once per set-up of the levels, so that each V-cycle's tail
// starts from plain 16-byte loads (its set-up's index arithmetic, wraps and packing were a
// third of its time).  One workgroup of MG_TAILQ_NTH threads.
template <typename T, int NL>
__global__ void __launch_bounds__(MG_TAILQ_NTH) k_mg_tail_q_pack(MgTailArgs<T> a, TqQuad<T>* __restrict__ qout,
                                                               T* __restrict__ cimg) {
  constexpr int NTH = MG_TAILQ_NTH;
  const int t = (int)threadIdx.x;
  int bo[NL + 1], nq[NL];
  tq_layout<NL>(a, bo, nq);
  auto PHI = [&](int l) { return (unsigned)(bo[l] + a.L[l].d.nx * a.L[l].d.ny); };
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    if (t >= nq[l]) continue;
    TqQuad<T> q;
    T mm[4];
    int I, J;
    tq_quad(t, a.L[l].d.nx >> 1, I, J);
    tq_load<T>(a.L[l], I, J, PHI(l), PHI(l + 1), q, mm);
    tq_finish<T>(a.L[l].d, I, J, q, mm, a.L[l].m != nullptr);
    q.own0 = PHI(l) + (unsigned)(2 * I + 2 * J * a.L[l].d.nx);
    qout[l * NTH + t] = q;
  }
  const MgDims dc = a.L[NL].d;
  const int ncc = dc.nx * dc.ny;
  const int ncx = (dc.nx + 1) * dc.ny, ncy = dc.nx * (dc.ny + 1);
  if (a.L[NL].cinv) {
    for (int p = t; p < ncc * ncc; p += NTH) {
      int r, qq;
      tq_quad(p, ncc, r, qq);   // p = r + qq n
      cimg[p] = a.L[NL].cinv[r * ncc + qq];
    }
  } else {
    for (int e = t; e < ncx + ncy + ncc; e += NTH)
      cimg[e] = e < ncx ? a.L[NL].cx[e] : (e < ncx + ncy ? a.L[NL].cy[e - ncx] : a.L[NL].dinv[e - ncx - ncy]);
  }
}

// NL smoothing levels a.L[0..NL-1] and the coarsest a.L[NL]; LDS: level l's b at bo[l], its phi
// at bo[l] + n_l; the coarsest level's operator in sc.  The quads' registers and the coarsest
// operator come from k_mg_tail_q_pack's images (a.tq, a.tqc).
template <typename T, int NL>
__global__ void __launch_bounds__(MG_TAILQ_NTH) k_mg_tail_q(MgTailArgs<T> a) {
  __shared__ T s[MG_TAILQ_POOL];
  __shared__ T sc[MG_TAILQ_COARSE];
  __shared__ T sphi[MG_COARSE_MAX];
  __shared__ T sb[MG_COARSE_MAX];
  __shared__ double red[16];
  constexpr int NTH = MG_TAILQ_NTH;
  constexpr int KC = (MG_TAILQ_COARSE + NTH - 1) / NTH;
  const int t = (int)threadIdx.x;
  int bo[NL + 1], nq[NL];
  tq_layout<NL>(a, bo, nq);
  auto PHI = [&](int l) { return (unsigned)(bo[l] + a.L[l].d.nx * a.L[l].d.ny); };
  // every load in flight: the top level's b and phi (zero for a zero guess; finite values are
  // needed even then, the wall terms read them times a zero conductance), the coarsest
  // operator's image and the quads' registers (indices clamped into the arrays instead of
  // branches around the loads; values loaded for no cell are never used)
  const MgDims d0 = a.L[0].d;
  const int n0 = d0.nx * d0.ny;
  T vb[4], vp[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const unsigned e = (unsigned)min(t + NTH * r, n0 - 1);
    vb[r] = mgt_ld(a.L[0].b, e);
    vp[r] = a.zero_guess ? T(0) : mgt_ld((const T*)a.L[0].phi, e);
  }
  const MgDims dc = a.L[NL].d;
  const int ncc = dc.nx * dc.ny;
  const int nco = a.L[NL].cinv ? ncc * ncc : (dc.nx + 1) * dc.ny + dc.nx * (dc.ny + 1) + ncc;
  const T* cimg = (const T*)a.tqc;
  T vc[KC];
#pragma unroll
  for (int k = 0; k < KC; ++k) vc[k] = mgt_ld(cimg, (unsigned)min(t + NTH * k, nco - 1));
  // (the quads of the lower levels occupy the first wavefronts only: the others skip their loads)
  const TqQuad<T>* qin = (const TqQuad<T>*)a.tq;
  TqQuad<T> q[NL];
#pragma unroll
  for (int l = 0; l < NL; ++l)
    if (t < nq[l]) q[l] = qin[l * NTH + t];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int e = t + NTH * r;
    if (e < n0) { s[bo[0] + e] = vb[r]; s[PHI(0) + e] = vp[r]; }
  }
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    const int e = t + NTH * k;
    if (e < nco) sc[e] = vc[k];
  }
  __syncthreads();
  unsigned own0[NL];
#pragma unroll
  for (int l = 0; l < NL; ++l) own0[l] = q[l].own0;
  const int nu = a.nu;
  // descent: pre-smoothing, residuals summed onto the coarse cells
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    const int nx = a.L[l].d.nx, db = bo[l] - (int)PHI(l);
    const bool act = t < nq[l];
    const bool zg = (l == 0) ? a.zero_guess != 0 : true;
    for (int sw = 0; sw < nu; ++sw) {
      tq_half<T, 0, false>(q[l], act, own0[l], nx, db, 0u, sw == 0 && zg, s);
      tq_half<T, 1, false>(q[l], act, own0[l], nx, db, 0u, false, s);
    }
    tq_restrict<T>(q[l], act, own0[l], nx, db, (unsigned)(bo[l + 1] + t), PHI(l + 1) + (unsigned)t, s);
  }
  tq_coarse(dc, a.L[NL].cinv != nullptr, (const T*)sc, (const T*)(s + bo[NL]), s + PHI(NL), a.coarse_sweeps, a.project, sphi, sb, red);
  // ascent: post-smoothing, the prolongation folded into its first half-sweep
#pragma unroll
  for (int l = NL - 1; l >= 0; --l) {
    const int nx = a.L[l].d.nx, db = bo[l] - (int)PHI(l);
    const bool act = t < nq[l];
    const unsigned pc = PHI(l + 1) + (unsigned)t;
    tq_half<T, 1, true>(q[l], act, own0[l], nx, db, pc, false, s);
    tq_half<T, 0, false>(q[l], act, own0[l], nx, db, 0u, false, s);
    for (int sw = 1; sw < nu; ++sw) {
      tq_half<T, 1, false>(q[l], act, own0[l], nx, db, 0u, false, s);
      tq_half<T, 0, false>(q[l], act, own0[l], nx, db, 0u, false, s);
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int e = t + NTH * r;
    if (e < n0) a.L[0].phi[e] = s[PHI(0) + e];
  }
}

// Level-0 load: b0 = z_p (- mean when both directions are walls), compact layout; also clears
// fl[done_flag] (the pressure block's early-exit flag, which the PCD reads after it).
template <typename T>
__global__ void k_mg_load(Grid g, const T* __restrict__ zp, T* __restrict__ b0, int walls_both,
                          const double* __restrict__ sc, int mean_slot, int* __restrict__ fl, int done_flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i == 0 && j == 0) fl[done_flag] = 0;
  if (i >= g.nx || j >= g.ny) return;
  const T z = zp[IP(g, i + 1, j + 1)];
  b0[i + (size_t)j * g.nx] = walls_both ? z - (T)sc[mean_slot] : z;
}

// Level-0 store into the padded phi (interior); ghosts are left at zero by the caller.
template <typename T>
__global__ void k_mg_store(Grid g, const T* __restrict__ phi0, T* __restrict__ phi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= g.nx || j >= g.ny) return;
  phi[IP(g, i + 1, j + 1)] = phi0[i + (size_t)j * g.nx];
}

}  // namespace fugu

namespace fugu {

// ---------------------------------------------------------------------------------
// Velocity-block multigrid (chosen fast forms): the u and v Helmholtz operators of the PCD's
// damped-Jacobi blocks (src/mod_subsolvers.f90:252-372) on compact node arrays.
// u block: node (ci, cj) <-> u(ci+1, cj+1) over the rows i = 1..i_u_max, j = 1..ny (an odd
// column count with walls in x); the wall faces u(0), u(nx) and the wall ghost rows are the
// zero Dirichlet values behind the boundary faces.  v block: node (ci, cj) <-> v(ci+1, cj+1)
// over i = 1..nx, j = 1..j_v_max.  Face conductances are the Jacobi coefficients (u: kappa/dx^2
// across x, (zeta - eta)/dy^2 across y; v: (zeta - eta)/dx^2 across x, kappa/dy^2 across y);
// m = rho/dt (+ penalty).
inline MgDims vmg_dims(const Grid& g, int comp) {
  if (comp == 0) return MgDims{g.per_x ? g.nx : g.nx - 1, g.ny, g.per_x, g.per_y};
  return MgDims{g.nx, g.per_y ? g.ny : g.ny - 1, g.per_x, g.per_y};
}

template <typename T>
__global__ void k_vmg_coef0(Grid g, MgDims d, int comp, const T* __restrict__ rho, const T* __restrict__ kap,
                            const T* __restrict__ zeta, const T* __restrict__ eta, T dt,
                            const T* __restrict__ pen, T* __restrict__ cx, T* __restrict__ cy,
                            T* __restrict__ m) {
  const int fi = blockIdx.x * blockDim.x + threadIdx.x;
  const int fj = blockIdx.y * blockDim.y + threadIdx.y;
  const double ix2 = g.inv_dx2, iy2 = g.inv_dy2;
  // a periodic direction has one value per wrapped face: face n is face 0
  if (fi <= d.nx && fj < d.ny) {   // x face fi (west face of node fi)
    const int ff = (d.per_x && fi == d.nx) ? 0 : fi;
    const double c = comp == 0
        ? (double)kap[IP(g, ff + 1, fj + 1)] * ix2                                        // u(fi) | u(fi+1)
        : ((double)zeta[IW(g, ff, fj + 1)] - (double)eta[IW(g, ff, fj + 1)]) * ix2;      // v(fi) | v(fi+1)
    cx[fi + (size_t)fj * (d.nx + 1)] = (T)c;
  }
  if (fi < d.nx && fj <= d.ny) {   // y face fj (south face of node row fj)
    const int ff = (d.per_y && fj == d.ny) ? 0 : fj;
    const double c = comp == 0
        ? ((double)zeta[IW(g, fi + 1, ff)] - (double)eta[IW(g, fi + 1, ff)]) * iy2       // u(., fj) | u(., fj+1)
        : (double)kap[IP(g, fi + 1, ff + 1)] * iy2;                                       // v(., fj) | v(., fj+1)
    cy[fi + (size_t)fj * d.nx] = (T)c;
  }
  if (fi < d.nx && fj < d.ny) {
    double mm;
    if (comp == 0) {
      mm = (double)rho[IU(g, fi + 1, fj + 1)] / (double)dt;
      if (pen) mm += (double)pen[IU(g, fi + 1, fj + 1)];
    } else {
      mm = (double)rho[IV(g, fi + 1, fj + 1)] / (double)dt;
      if (pen) mm += (double)pen[IV(g, fi + 1, fj + 1)];
    }
    m[fi + (size_t)fj * d.nx] = (T)mm;
  }
}

// max over the u and v blocks' level-0 cells of (sum of face conductances) / mass: how far the
// diffusion dominates the mass term, which sets how deep the velocity hierarchies must go.
template <typename T>
__global__ void __launch_bounds__(256) k_vmg_ratio(MgDims du, const T* __restrict__ cxu, const T* __restrict__ cyu,
                                                   const T* __restrict__ mu, MgDims dv, const T* __restrict__ cxv,
                                                   const T* __restrict__ cyv, const T* __restrict__ mv,
                                                   double* __restrict__ partials) {
  double r[2] = {0.0, 0.0};
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t nuc = (size_t)du.nx * du.ny, nvc = (size_t)dv.nx * dv.ny;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < nuc || q < nvc; q += stride) {
    if (q < nuc) {
      const size_t i = q % du.nx, j = q / du.nx;
      const double s = (double)cxu[i + j * (du.nx + 1)] + (double)cxu[i + 1 + j * (du.nx + 1)] + (double)cyu[q] +
                       (double)cyu[q + du.nx];
      const double m = (double)mu[q];
      if (m > 0.0) r[0] = fmax(r[0], s / m);
    }
    if (q < nvc) {
      const size_t i = q % dv.nx, j = q / dv.nx;
      const double s = (double)cxv[i + j * (dv.nx + 1)] + (double)cxv[i + 1 + j * (dv.nx + 1)] + (double)cyv[q] +
                       (double)cyv[q + dv.nx];
      const double m = (double)mv[q];
      if (m > 0.0) r[1] = fmax(r[1], s / m);
    }
  }
  publish_max<2>(r, partials);
}

// rhs rows -> compact b.
template <typename T>
__global__ void k_vmg_load(Grid g, MgDims d, const T* __restrict__ rhs, int comp, T* __restrict__ b) {
  const int ci = blockIdx.x * blockDim.x + threadIdx.x;
  const int cj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ci >= d.nx || cj >= d.ny) return;
  b[ci + (size_t)cj * d.nx] = comp == 0 ? rhs[IU(g, ci + 1, cj + 1)] : rhs[IV(g, ci + 1, cj + 1)];
}

// compact phi -> the padded output over its whole array: rows from phi, wall faces = rhs (as
// the Jacobi blocks leave them), everything else zero.
template <typename T>
__global__ void __launch_bounds__(NT) k_vmg_store(Grid g, int comp, const T* __restrict__ phi, const T* __restrict__ rhs,
                                                  T* __restrict__ out) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  if (comp == 0) {
    if (i < -nh || i > nx + nh || j < 1 - nh || j > ny + nh) return;
    const int i_u_max = g.per_x ? nx : nx - 1;
    T o = T(0);
    if (j >= 1 && j <= ny && i >= 1 && i <= i_u_max) o = phi[(i - 1) + (size_t)(j - 1) * i_u_max];
    else if (!g.per_x && j >= 1 && j <= ny && (i == 0 || i == nx)) o = rhs[IU(g, i, j)];
    out[IU(g, i, j)] = o;
  } else {
    if (i < 1 - nh || i > nx + nh || j < -nh || j > ny + nh) return;
    const int j_v_max = g.per_y ? ny : ny - 1;
    T o = T(0);
    if (i >= 1 && i <= nx && j >= 1 && j <= j_v_max) o = phi[(i - 1) + (size_t)(j - 1) * nx];
    else if (!g.per_y && i >= 1 && i <= nx && (j == 0 || j == ny)) o = rhs[IV(g, i, j)];
    out[IV(g, i, j)] = o;
  }
}

// The PCD's v-block store, its periodic halo and the cross term of the u block in one launch:
// rv (padded v array) = what k_vmg_store (comp 1) followed by the v halo leaves; bu (compact u
// rows) = zuw - cross(eta, rv), each thread reading the stored v values it needs straight from
// the compact result (same values, same operations and order as k_pcd_cross).
template <typename T>
__global__ void __launch_bounds__(NT) k_vmg_store_cross(Grid g, const T* __restrict__ phi, const T* __restrict__ rhs,
                                                        T* __restrict__ rv, const T* __restrict__ eta,
                                                        const T* __restrict__ zuw, T* __restrict__ bu) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  const int j_v_max = g.per_y ? ny : ny - 1;
  // Where the stored v at (ii, jj) comes from (periodic ghosts from their images): the compact
  // result (kind 0, index q), a wall row of rhs (kind 1, padded index q) or zero (kind 2).
  auto where = [&](int ii, int jj, int& kind, size_t& q) {
    if (g.per_x) {
      if (ii >= 1 - nh && ii <= 0) ii += nx;
      else if (ii >= nx + 1 && ii <= nx + nh) ii -= nx;
    }
    if (g.per_y) {
      if (jj >= 1 - nh && jj <= 0) jj += ny;
      else if (jj >= ny + 1 && jj <= ny + nh) jj -= ny;
    }
    kind = 2;
    q = 0;
    if (ii >= 1 && ii <= nx && jj >= 1 && jj <= j_v_max) {
      kind = 0;
      q = (ii - 1) + (size_t)(jj - 1) * nx;
    } else if (!g.per_y && ii >= 1 && ii <= nx && (jj == 0 || jj == ny)) {
      kind = 1;
      q = IV(g, ii, jj);
    }
  };
  // the four stored values this point needs, (i, j), (i+1, j), (i, j-1), (i+1, j-1): every load
  // issued before any is used
  int kd[4];
  size_t qq[4];
  where(i, j, kd[0], qq[0]);
  where(i + 1, j, kd[1], qq[1]);
  where(i, j - 1, kd[2], qq[2]);
  where(i + 1, j - 1, kd[3], qq[3]);
  T v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    T x = T(0);
    if (kd[k] == 0) x = phi[qq[k]];
    else if (kd[k] == 1) x = rhs[qq[k]];
    v[k] = x;
  }
  const int i_u_max = g.per_x ? nx : nx - 1;
  const bool urow = i >= 1 && i <= i_u_max && j >= 1 && j <= ny;
  T val = T(0), e0 = T(0), e1 = T(0);
  if (urow) {
    val = zuw[IU(g, i, j)];
    e0 = eta[IW(g, i, j)];
    e1 = eta[IW(g, i, j - 1)];
  }
  if (i >= 1 - nh && i <= nx + nh && j >= -nh && j <= ny + nh) rv[IV(g, i, j)] = v[0];
  if (urow) {
    if (i <= nx - 1) {
      const T inv_dx = (T)g.inv_dx, inv_dy = (T)g.inv_dy;
      const T cross = -inv_dy * (e0 * (v[1] - v[0]) * inv_dx - e1 * (v[3] - v[2]) * inv_dx);
      val = val - cross;
    }
    bu[(i - 1) + (size_t)(j - 1) * i_u_max] = val;
  }
}

// The PCD's u-block store, its periodic halo and the homogeneous velocity BCs of the whole
// output (k_velocity_bcs with a zero lid on ru and rv) in one launch.  ru's final values are
// formed from the compact result as store + halo + BCs leave them (the BCs read only positions
// they do not write, so each is a function of the stored values); rv, stored by the previous
// launch, gets its BC positions exactly as k_velocity_bcs sets them.
template <typename T>
__global__ void __launch_bounds__(NT) k_vmg_store_bcs(Grid g, const T* __restrict__ phi, const T* __restrict__ rhs,
                                                      T* __restrict__ ru, T* __restrict__ rv) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  const int i_u_max = g.per_x ? nx : nx - 1;
  // Each output is one stored value with signs applied: its source (the compact result, an rhs
  // wall face, or zero) is resolved first and both outputs' loads are issued before either value
  // is used (a use between loads waits for all of them).  Signs and zeros are applied with the
  // same operations as the store / halo / BC sequence: -x for a mirrored value, 0 - x at the top
  // wall (the zero lid), a literal zero on wall faces.
  // ru at (i, j)
  const bool u_here = i >= -nh && i <= nx + nh && j >= 1 - nh && j <= ny + nh;
  int ui = i, uj = j;
  int outer = 0;         // 0: x, 1: -x, 2: 0 - x
  bool inner_neg = false, inner_zero = false;
  {
    const bool xt = !g.per_x && (i <= 0 || i >= nx);
    const bool yt = !g.per_y && (j <= 0 || j >= ny + 1);
    bool ax = false;     // through u_after_x
    if (yt) {
      ax = true;
      if (j <= 0) { outer = 1; uj = 1 - j; }
      else { outer = 2; uj = 2 * ny + 1 - j; }
    } else if (xt) {
      ax = true;
    }
    if (ax && !g.per_x) {
      if (ui == 0 || ui == nx) inner_zero = true;
      else if (ui < 0 && ui >= -nh) { inner_neg = true; ui = -ui; }
      else if (ui > nx && ui <= nx + nh) { inner_neg = true; ui = 2 * nx - ui; }
    }
  }
  // the stored u at (ui, uj): periodic ghosts from their images, then the compact result, a
  // wall face of rhs, or zero
  if (g.per_x) {
    if (ui >= 1 - nh && ui <= 0) ui += nx;
    else if (ui >= nx + 1 && ui <= nx + nh) ui -= nx;
  }
  if (g.per_y) {
    if (uj >= 1 - nh && uj <= 0) uj += ny;
    else if (uj >= ny + 1 && uj <= ny + nh) uj -= ny;
  }
  const bool u_phi = uj >= 1 && uj <= ny && ui >= 1 && ui <= i_u_max;
  const bool u_rhs = !u_phi && !g.per_x && uj >= 1 && uj <= ny && (ui == 0 || ui == nx);
  // rv at (i, j): BC positions only, from rv itself (positions the BCs do not write)
  const bool v_here = i >= 1 - nh && i <= nx + nh && j >= -nh && j <= ny + nh;
  const bool vxt = !g.per_x && (i <= 0 || i >= nx + 1);
  const bool vyt = !g.per_y && (j <= 0 || j >= ny);
  const bool v_store = v_here && (vyt || vxt);
  bool v_zero = false, v_outer_neg = false, v_inner_neg = false;
  int vi = i, vj = j;
  if (vyt) {
    if (j == 0 || j == ny) v_zero = true;
    else { v_outer_neg = true; vj = (j < 0) ? -j : 2 * ny - j; }
  }
  if (!g.per_x) {
    if (vi <= 0 && vi >= 1 - nh) { v_inner_neg = true; vi = 1 - vi; }
    else if (vi > nx && vi <= nx + nh) { v_inner_neg = true; vi = 2 * nx + 1 - vi; }
  }
  // one unconditional load each, from element 0 where nothing is needed (its value is unused)
  const bool u_load = u_here && !inner_zero && (u_phi || u_rhs);
  const bool v_load = v_store && !v_zero;
  const T* usrc = u_phi ? phi : rhs;
  const size_t uq = !u_load ? 0 : (u_phi ? (ui - 1) + (size_t)(uj - 1) * i_u_max : IU(g, ui, uj));
  const T xu0 = usrc[uq];
  const T xv0 = rv[v_load ? IV(g, vi, vj) : 0];
  const T xu = u_load ? xu0 : T(0), xv = v_load ? xv0 : T(0);
  if (u_here) {
    const T y = inner_zero ? T(0) : (inner_neg ? -xu : xu);
    ru[IU(g, i, j)] = outer == 1 ? -y : (outer == 2 ? (T)(2.0 * 0.0) - y : y);
  }
  if (v_store) {
    const T y = v_inner_neg ? -xv : xv;
    rv[IV(g, i, j)] = v_zero ? T(0) : (v_outer_neg ? -y : y);
  }
}

}  // namespace fugu
