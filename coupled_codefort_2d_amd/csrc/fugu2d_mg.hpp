// fugu2d_mg.hpp -- geometric multigrid V-cycle for the pressure block of the PCD
// preconditioner (the variable-density operator A_p phi = sum_f c_f (phi_i - phi_nb) with
// c_f = h^-2 / rho_f, zero-flux walls, periodic wrap), used by the chosen fast solve
// forms in place of the reference's 150-iteration Jacobi-PCG (src/mod_subsolvers.f90:32-246).
//
// Cell-centred hierarchy on compact arrays (no ghosts), factor-2 coarsening while both
// extents are even and >= 8.  Coarse face conductances average the two fine faces they
// cover, c_H = (c_1 + c_2) / 8 (parallel faces, spacing 2h).  Smoother: red-black
// Gauss-Seidel, checkerboard-compact launches (each thread owns one cell of the colour).
// Restriction averages the four fine residuals (fused with the residual), prolongation is
// piecewise constant.  The coarsest level is solved by one workgroup in LDS.  With a zero
// initial guess and fixed sweep counts the V-cycle is a fixed linear operator, which is
// what keeps the outer Krylov recurrences consistent.
#pragma once
#include "fugu2d_common.hpp"

namespace fugu {

struct MgDims {
  int nx, ny, per_x, per_y;
};

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
  if (fi <= nx && fj < ny) {
    double c = 0.0;
    if (g.per_x || (fi > 0 && fi < nx)) c = g.inv_dx2 / (double)rho_u[IU(g, fi, fj + 1)];
    cx[fi + (size_t)fj * (nx + 1)] = (T)c;
  }
  if (fi < nx && fj <= ny) {
    double c = 0.0;
    if (g.per_y || (fj > 0 && fj < ny)) c = g.inv_dy2 / (double)rho_v[IV(g, fi + 1, fj)];
    cy[fi + (size_t)fj * nx] = (T)c;
  }
}

// Coarse conductances from the fine ones.
template <typename T>
__global__ void k_mg_coarsen(MgDims f, const T* __restrict__ cxf, const T* __restrict__ cyf,
                             T* __restrict__ cxc, T* __restrict__ cyc) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const int J = blockIdx.y * blockDim.y + threadIdx.y;
  const int nxc = f.nx / 2, nyc = f.ny / 2;
  if (I <= nxc && J < nyc) {
    const int fi = 2 * I;
    const T a = cxf[fi + (size_t)(2 * J) * (f.nx + 1)];
    const T b = cxf[fi + (size_t)(2 * J + 1) * (f.nx + 1)];
    cxc[I + (size_t)J * (nxc + 1)] = (a + b) * T(0.125);
  }
  if (I < nxc && J <= nyc) {
    const int fj = 2 * J;
    const T a = cyf[2 * I + (size_t)fj * f.nx];
    const T b = cyf[2 * I + 1 + (size_t)fj * f.nx];
    cyc[I + (size_t)J * nxc] = (a + b) * T(0.125);
  }
}

// Inverse diagonal of a level.
template <typename T>
__global__ void k_mg_dinv(MgDims d, const T* __restrict__ cx, const T* __restrict__ cy, T* __restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= d.nx || j >= d.ny) return;
  const T s = cx[i + (size_t)j * (d.nx + 1)] + cx[i + 1 + (size_t)j * (d.nx + 1)] +
              cy[i + (size_t)j * d.nx] + cy[i + (size_t)(j + 1) * d.nx];
  dinv[i + (size_t)j * d.nx] = (s > T(1.0e-14)) ? T(1) / s : T(0);
}

template <typename T>
__device__ inline T mg_nbr_sum(const MgDims& d, const T* cx, const T* cy, const T* phi, int i, int j) {
  const size_t rx = (size_t)(d.nx + 1), o = (size_t)j * d.nx;
  const T cW = cx[i + j * rx], cE = cx[i + 1 + j * rx];
  const T cS = cy[i + o], cN = cy[i + o + d.nx];
  const int iw = mg_wrap(i - 1, d.nx, d.per_x), ie = mg_wrap(i + 1, d.nx, d.per_x);
  const int js = mg_wrap(j - 1, d.ny, d.per_y), jn = mg_wrap(j + 1, d.ny, d.per_y);
  return cW * phi[iw + o] + cE * phi[ie + o] + cS * phi[i + (size_t)js * d.nx] + cN * phi[i + (size_t)jn * d.nx];
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

// Residual of the fine level averaged onto the coarse cells: bc = (1/4) sum r_children.
template <typename T>
__global__ void __launch_bounds__(256) k_mg_restrict(MgDims f, const T* __restrict__ cx, const T* __restrict__ cy,
                                                     const T* __restrict__ b, const T* __restrict__ phi,
                                                     T* __restrict__ bc, T* __restrict__ phic) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const int J = blockIdx.y * blockDim.y + threadIdx.y;
  const int nxc = f.nx / 2, nyc = f.ny / 2;
  if (I >= nxc || J >= nyc) return;
  T acc = T(0);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = 2 * I + (q & 1), j = 2 * J + (q >> 1);
    const size_t o = i + (size_t)j * f.nx;
    const size_t rx = (size_t)(f.nx + 1);
    const T diag = cx[i + j * rx] + cx[i + 1 + j * rx] + cy[o] + cy[o + f.nx];
    acc += b[o] - (diag * phi[o] - mg_nbr_sum(f, cx, cy, phi, i, j));
  }
  bc[I + (size_t)J * nxc] = acc * T(0.25);
  phic[I + (size_t)J * nxc] = T(0);
}

// phi_f += phi_c(parent)
template <typename T>
__global__ void __launch_bounds__(256) k_mg_prolong(MgDims f, const T* __restrict__ phic, T* __restrict__ phi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= f.nx || j >= f.ny) return;
  const size_t o = i + (size_t)j * f.nx;
  phi[o] = phi[o] + phic[(i >> 1) + (size_t)(j >> 1) * (f.nx / 2)];
}

// Coarsest level: one workgroup, LDS-resident, optional zero-mean projection of the rhs
// (singular Neumann operator), then `sweeps` red-black sweeps from zero.
constexpr int MG_COARSE_MAX = 1024;
template <typename T>
__global__ void __launch_bounds__(1024) k_mg_coarse(MgDims d, const T* __restrict__ cx, const T* __restrict__ cy,
                                                    const T* __restrict__ dinv, const T* __restrict__ b,
                                                    T* __restrict__ phi, int sweeps, int project) {
  __shared__ T sphi[MG_COARSE_MAX];
  __shared__ T sb[MG_COARSE_MAX];
  __shared__ double red[16];
  const int n = d.nx * d.ny;
  const int t = threadIdx.x;
  double my = 0.0;
  if (t < n) { sb[t] = b[t]; sphi[t] = T(0); my = (double)sb[t]; }
  if (project) {
    my = wave_sum(my);
    if ((t & 63) == 0) red[t >> 6] = my;
    __syncthreads();
    double m = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m += red[w];
    m /= (double)n;
    if (t < n) sb[t] = sb[t] - (T)m;
  }
  __syncthreads();
  const int i = (t < n) ? t % d.nx : 0, j = (t < n) ? t / d.nx : 0;
  for (int s = 0; s < sweeps; ++s) {
    for (int colour = 0; colour < 2; ++colour) {
      if (t < n && ((i + j) & 1) == colour) sphi[t] = (sb[t] + mg_nbr_sum(d, cx, cy, (const T*)sphi, i, j)) * dinv[t];
      __syncthreads();
    }
  }
  if (t < n) phi[t] = sphi[t];
}

// Level-0 load: b0 = z_p (- mean when both directions are walls), compact layout.
template <typename T>
__global__ void k_mg_load(Grid g, const T* __restrict__ zp, T* __restrict__ b0, int walls_both,
                          const double* __restrict__ sc, int mean_slot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
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
