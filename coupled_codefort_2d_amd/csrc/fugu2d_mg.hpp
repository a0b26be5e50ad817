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

inline MgDims mg_coarse_dims(const MgDims& f) { return MgDims{(f.nx + 1) / 2, (f.ny + 1) / 2, f.per_x, f.per_y}; }
inline bool mg_can_coarsen(const MgDims& d) {
  return d.nx >= 8 && d.ny >= 8 && !(d.per_x && (d.nx & 1)) && !(d.per_y && (d.ny & 1));
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

template <typename T>
__device__ inline T mg_nbr_sum(const MgDims& d, const T* cx, const T* cy, const T* phi, int i, int j) {
  const size_t rx = (size_t)(d.nx + 1), o = (size_t)j * d.nx;
  // boundary faces of a non-periodic edge couple to a zero ghost (diagonal only)
  const T cW = (i > 0 || d.per_x) ? cx[i + j * rx] : T(0);
  const T cE = (i < d.nx - 1 || d.per_x) ? cx[i + 1 + j * rx] : T(0);
  const T cS = (j > 0 || d.per_y) ? cy[i + o] : T(0);
  const T cN = (j < d.ny - 1 || d.per_y) ? cy[i + o + d.nx] : T(0);
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
  if (fi <= d.nx && fj < d.ny) {   // x face fi (west face of node fi)
    const double c = comp == 0
        ? (double)kap[IP(g, fi + 1, fj + 1)] * ix2                                        // u(fi) | u(fi+1)
        : ((double)zeta[IW(g, fi, fj + 1)] - (double)eta[IW(g, fi, fj + 1)]) * ix2;      // v(fi) | v(fi+1)
    cx[fi + (size_t)fj * (d.nx + 1)] = (T)c;
  }
  if (fi < d.nx && fj <= d.ny) {   // y face fj (south face of node row fj)
    const double c = comp == 0
        ? ((double)zeta[IW(g, fi + 1, fj)] - (double)eta[IW(g, fi + 1, fj)]) * iy2       // u(., fj) | u(., fj+1)
        : (double)kap[IP(g, fi + 1, fj + 1)] * iy2;                                       // v(., fj) | v(., fj+1)
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

}  // namespace fugu
