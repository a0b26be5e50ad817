// fugu2d_adjoint.hpp -- adjoint kernels (no counterpart in the reference: it has no adjoint).
//
// The step operator as the solvers see it acts on the owned degrees of freedom
// (u: j=1..ny, i=1..i_u_max; v: i=1..nx, j=1..j_v_max; p: interior):
//     A_eff = R A_pad H B E0
// (E0 embed, B velocity BCs with lid 0, H halo, A_pad padded matvec, R owned rows).  The
// transpose is evaluated in gather form: k_matvec_T computes (A_pad^T R^T lam) at every padded
// position from the rows that reference it; k_halo_T_* / k_bcs_T_* then fold ghost entries
// back onto the values they were copied or reflected from (H^T, B^T), in the reverse order
// of the forward statements, and k_mask_trip applies E0^T.
//
// The transposed linearised PCD follows the forward PCD's data flow backwards with the same
// neighbour wrap rules (periodic neighbours wrap; wall ghosts are constant zeros whose adjoints
// are dropped).  The coefficient-gradient kernel accumulates lam^T dA/dcoef X with FP64 atomics
// (one pass per step of the reverse sweep, not on the Krylov path).
//
// The Krylov-path kernels (the transposed matvec, halo and BC folds, the transposed PCD pieces)
// take their arrays in a storage type T -- double, or float for the FP32 Arnoldi cycles of the
// mixed adjoint solves -- and compute in FP64 registers; for T = double the arithmetic is the
// FP64 kernels' operation for operation.
#pragma once
#include "fugu2d_common.hpp"

namespace fugu {

__device__ inline bool own_u(const Grid& g, int i, int j) {
  return j >= 1 && j <= g.ny && i >= 1 && i <= (g.per_x ? g.nx : g.nx - 1);
}
__device__ inline bool own_v(const Grid& g, int i, int j) {
  return i >= 1 && i <= g.nx && j >= 1 && j <= (g.per_y ? g.ny : g.ny - 1);
}
__device__ inline bool own_p(const Grid& g, int i, int j) { return i >= 1 && i <= g.nx && j >= 1 && j <= g.ny; }
__device__ inline bool in_u(const Grid& g, int i, int j) {
  return i >= -g.nh && i <= g.nx + g.nh && j >= 1 - g.nh && j <= g.ny + g.nh;
}
__device__ inline bool in_v(const Grid& g, int i, int j) {
  return i >= 1 - g.nh && i <= g.nx + g.nh && j >= -g.nh && j <= g.ny + g.nh;
}
__device__ inline bool in_p(const Grid& g, int i, int j) {
  return i >= 1 - g.nh && i <= g.nx + g.nh && j >= 1 - g.nh && j <= g.ny + g.nh;
}
__device__ inline bool in_w(const Grid& g, int i, int j) {
  return i >= -g.nh && i <= g.nx + g.nh && j >= -g.nh && j <= g.ny + g.nh;
}

// E0^T / R^T: zero everything outside the owned sets (any of the pointers may be null).
template <typename T>
__global__ void __launch_bounds__(NT) k_mask_trip(Grid g, T* __restrict__ u, T* __restrict__ v, T* __restrict__ p) {
  int i, j;
  union_ij(g, i, j);
  if (u && in_u(g, i, j) && !own_u(g, i, j)) u[IU(g, i, j)] = T(0);
  if (v && in_v(g, i, j) && !own_v(g, i, j)) v[IV(g, i, j)] = T(0);
  if (p && in_p(g, i, j) && !own_p(g, i, j)) p[IP(g, i, j)] = T(0);
}

template <typename T>
struct MatvecTArgs {
  const T *Lu, *Lv, *Lp;
  T *Ou, *Ov, *Op;
  const T *uadv, *vadv, *rho_c, *rho_u, *rho_v, *kap, *zeta, *eta;
  double dt;
  int adv;
};

// A_pad^T (R^T lam) at every padded position (gather form of the reverse-mode of the matvec).
template <typename T>
__global__ void __launch_bounds__(NT) k_matvec_T(Grid g, MatvecTArgs<T> a) {
  int i, j;
  union_ij(g, i, j);
  const double ix = g.inv_dx, iy = g.inv_dy, half = 0.5, dt = a.dt;
  auto LU = [&](int ii, int jj) -> double { return own_u(g, ii, jj) ? (double)a.Lu[IU(g, ii, jj)] : 0.0; };
  auto LV = [&](int ii, int jj) -> double { return own_v(g, ii, jj) ? (double)a.Lv[IV(g, ii, jj)] : 0.0; };
  auto LP = [&](int ii, int jj) -> double { return own_p(g, ii, jj) ? (double)a.Lp[IP(g, ii, jj)] : 0.0; };
  auto KAP = [&](int ii, int jj) -> double { return in_p(g, ii, jj) ? (double)a.kap[IP(g, ii, jj)] : 0.0; };
  // adjoint stresses from lam (zero wherever no owned row references them)
  auto TXXb = [&](int ci, int cj) -> double { return (LU(ci, cj) - LU(ci - 1, cj)) * ix; };
  auto TYYb = [&](int ci, int cj) -> double { return (LV(ci, cj) - LV(ci, cj - 1)) * iy; };
  auto TXYb = [&](int vi, int vj) -> double { return (LU(vi, vj + 1) - LU(vi, vj)) * iy; };
  auto TYXb = [&](int vi, int vj) -> double { return (LV(vi + 1, vj) - LV(vi, vj)) * ix; };
  auto DUB = [&](int vi, int vj) -> double {     // d/d(du_dy) at vertex, times 1/dy
    if (!in_w(g, vi, vj)) return 0.0;
    const double ze = (double)a.zeta[IW(g, vi, vj)], et = (double)a.eta[IW(g, vi, vj)];
    return ((ze - et) * TXYb(vi, vj) + et * TYXb(vi, vj)) * iy;
  };
  auto DVB = [&](int vi, int vj) -> double {     // d/d(dv_dx) at vertex, times 1/dx
    if (!in_w(g, vi, vj)) return 0.0;
    const double ze = (double)a.zeta[IW(g, vi, vj)], et = (double)a.eta[IW(g, vi, vj)];
    return (et * TXYb(vi, vj) + (ze - et) * TYXb(vi, vj)) * ix;
  };
  // advective mass fluxes of the u row (ri, rj) and of the v row (ri, rj)
  auto UA = [&](int ii, int jj) -> double { return (double)a.uadv[IU(g, ii, jj)]; };
  auto VA = [&](int ii, int jj) -> double { return (double)a.vadv[IV(g, ii, jj)]; };
  auto RC = [&](int ii, int jj) -> double { return (double)a.rho_c[IP(g, ii, jj)]; };
  auto RU = [&](int ii, int jj) -> double { return (double)a.rho_u[IU(g, ii, jj)]; };
  auto RV = [&](int ii, int jj) -> double { return (double)a.rho_v[IV(g, ii, jj)]; };
  auto mEu = [&](int ri, int rj) { return RC(ri + 1, rj) * half * (UA(ri, rj) + UA(ri + 1, rj)); };
  auto mWu = [&](int ri, int rj) { return RC(ri, rj) * half * (UA(ri - 1, rj) + UA(ri, rj)); };
  auto mNu = [&](int ri, int rj) {
    return half * (RV(ri, rj) + RV(ri + 1, rj)) * half * (VA(ri, rj) + VA(ri + 1, rj));
  };
  auto mSu = [&](int ri, int rj) {
    return half * (RV(ri, rj - 1) + RV(ri + 1, rj - 1)) * half * (VA(ri, rj - 1) + VA(ri + 1, rj - 1));
  };
  auto mEv = [&](int ri, int rj) {
    return half * (RU(ri, rj) + RU(ri, rj + 1)) * half * (UA(ri, rj) + UA(ri, rj + 1));
  };
  auto mWv = [&](int ri, int rj) {
    return half * (RU(ri - 1, rj) + RU(ri - 1, rj + 1)) * half * (UA(ri - 1, rj) + UA(ri - 1, rj + 1));
  };
  auto mNv = [&](int ri, int rj) { return RC(ri, rj + 1) * half * (VA(ri, rj) + VA(ri, rj + 1)); };
  auto mSv = [&](int ri, int rj) { return RC(ri, rj) * half * (VA(ri, rj - 1) + VA(ri, rj)); };

  // ---- u position (i, j)
  if (in_u(g, i, j)) {
    double o = 0.0;
    o += (LP(i, j) - LP(i + 1, j)) * ix;                                  // divergence rows
    if (own_u(g, i, j)) o += (RU(i, j) / dt) * LU(i, j);                   // mass
    o += KAP(i, j) * TXXb(i, j) * ix - KAP(i + 1, j) * TXXb(i + 1, j) * ix;
    o += DUB(i, j - 1) - DUB(i, j);
    if (a.adv) {
      if (own_u(g, i, j)) {
        const double l = LU(i, j);
        o += l * half * (mEu(i, j) - mWu(i, j)) * ix + l * half * (mNu(i, j) - mSu(i, j)) * iy;
      }
      if (own_u(g, i - 1, j)) o += LU(i - 1, j) * half * mEu(i - 1, j) * ix;
      if (own_u(g, i + 1, j)) o -= LU(i + 1, j) * half * mWu(i + 1, j) * ix;
      if (own_u(g, i, j - 1)) o += LU(i, j - 1) * half * mNu(i, j - 1) * iy;
      if (own_u(g, i, j + 1)) o -= LU(i, j + 1) * half * mSu(i, j + 1) * iy;
    }
    a.Ou[IU(g, i, j)] = (T)o;
  }
  // ---- v position (i, j)
  if (in_v(g, i, j)) {
    double o = 0.0;
    o += (LP(i, j) - LP(i, j + 1)) * iy;
    if (own_v(g, i, j)) o += (RV(i, j) / dt) * LV(i, j);
    o += KAP(i, j) * TYYb(i, j) * iy - KAP(i, j + 1) * TYYb(i, j + 1) * iy;
    o += DVB(i - 1, j) - DVB(i, j);
    if (a.adv) {
      if (own_v(g, i, j)) {
        const double l = LV(i, j);
        o += l * half * (mEv(i, j) - mWv(i, j)) * ix + l * half * (mNv(i, j) - mSv(i, j)) * iy;
      }
      if (own_v(g, i - 1, j)) o += LV(i - 1, j) * half * mEv(i - 1, j) * ix;
      if (own_v(g, i + 1, j)) o -= LV(i + 1, j) * half * mWv(i + 1, j) * ix;
      if (own_v(g, i, j - 1)) o += LV(i, j - 1) * half * mNv(i, j - 1) * iy;
      if (own_v(g, i, j + 1)) o -= LV(i, j + 1) * half * mSv(i, j + 1) * iy;
    }
    a.Ov[IV(g, i, j)] = (T)o;
  }
  // ---- p position (i, j): the pressure gradient of the u and v rows
  if (in_p(g, i, j)) {
    const double o = (LU(i - 1, j) - LU(i, j)) * ix + (LV(i, j - 1) - LV(i, j)) * iy;
    a.Op[IP(g, i, j)] = (T)o;
  }
}

// H^T, y pass then x pass (the forward halo runs x then y).  kind: 0 u, 1 v, 2 p.  One thread
// per (position along the edge, layer k): adds the ghost to its source and clears the ghost.
__device__ inline size_t idx_kind(const Grid& g, int kind, int i, int j) {
  return kind == 0 ? IU(g, i, j) : kind == 1 ? IV(g, i, j) : IP(g, i, j);
}
// (a ghost folded onto its source: s += d in FP64, then d = 0)
template <typename T>
__device__ inline void fold_ghost(T* a, size_t s, size_t d) {
  a[s] = (T)((double)a[s] + (double)a[d]);
  a[d] = T(0);
}
template <typename T>
__global__ void k_halo_T_y(Grid g, T* __restrict__ a, int kind) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i0 = (kind == 0) ? -g.nh : 1 - g.nh, i1 = g.nx + g.nh;
  const int ni = i1 - i0 + 1;
  if (t >= ni * g.nh) return;
  const int i = i0 + t % ni, k = t / ni;
  const int ny = g.ny, nh = g.nh;
  // dst rows: ny+1+k (from 1+k) and 1-nh+k (from ny-nh+1+k); v rows are offset by the 0 face
  const int top = (kind == 1) ? ny + 1 + k : ny + 1 + k;
  fold_ghost(a, idx_kind(g, kind, i, 1 + k), idx_kind(g, kind, i, top));
  fold_ghost(a, idx_kind(g, kind, i, ny - nh + 1 + k), idx_kind(g, kind, i, 1 - nh + k));
}
template <typename T>
__global__ void k_halo_T_x(Grid g, T* __restrict__ a, int kind) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int j0 = (kind == 1) ? -g.nh : 1 - g.nh, j1 = g.ny + g.nh;
  const int nj = j1 - j0 + 1;
  if (t >= nj * g.nh) return;
  const int j = j0 + t % nj, k = t / nj;
  const int nx = g.nx, nh = g.nh;
  fold_ghost(a, idx_kind(g, kind, 1 + k, j), idx_kind(g, kind, nx + 1 + k, j));
  fold_ghost(a, idx_kind(g, kind, nx - nh + 1 + k, j), idx_kind(g, kind, 1 - nh + k, j));
}

// B^T: the velocity-BC statements in reverse order; y walls first (one thread per column),
// then x walls (one thread per row).
// (a mirrored ghost folded onto its source with the mirror's sign: s -= d in FP64, then d = 0)
template <typename T>
__device__ inline void fold_mirror(T* a, size_t s, size_t d) {
  a[s] = (T)((double)a[s] - (double)a[d]);
  a[d] = T(0);
}
template <typename T>
__global__ void k_bcs_T_y(Grid g, T* __restrict__ u, T* __restrict__ v) {
  const int i = -g.nh + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i > g.nx + g.nh) return;
  const int ny = g.ny, nh = g.nh;
  const bool iv = (i >= 1 - nh);
  for (int j = 1; j <= nh; ++j) fold_mirror(u, IU(g, i, ny - j + 1), IU(g, i, ny + j));
  if (iv) {
    for (int j = 1; j <= nh; ++j) fold_mirror(v, IV(g, i, ny - j), IV(g, i, ny + j));
    v[IV(g, i, ny)] = T(0);
  }
  for (int j = 1; j <= nh; ++j) fold_mirror(u, IU(g, i, j), IU(g, i, 1 - j));
  if (iv) {
    for (int j = 1; j <= nh; ++j) fold_mirror(v, IV(g, i, j), IV(g, i, -j));
    v[IV(g, i, 0)] = T(0);
  }
}
template <typename T>
__global__ void k_bcs_T_x(Grid g, T* __restrict__ u, T* __restrict__ v) {
  const int j = -g.nh + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (j > g.ny + g.nh) return;
  const int nx = g.nx, nh = g.nh;
  const bool ju = (j >= 1 - nh);
  for (int i = 1; i <= nh; ++i) fold_mirror(v, IV(g, nx - i + 1, j), IV(g, nx + i, j));
  if (ju) {
    for (int i = 1; i <= nh; ++i) fold_mirror(u, IU(g, nx - i, j), IU(g, nx + i, j));
    u[IU(g, nx, j)] = T(0);
  }
  for (int i = 1; i <= nh; ++i) fold_mirror(v, IV(g, i, j), IV(g, 1 - i, j));
  if (ju) {
    for (int i = 1; i <= nh; ++i) fold_mirror(u, IU(g, i, j), IU(g, -i, j));
    u[IU(g, 0, j)] = T(0);
  }
}

// ---------------------------------------------------------------------------------
// Transposed linearised PCD pieces.

__device__ inline int wrapi(int i, int n, int per) { return per ? (i < 1 ? i + n : (i > n ? i - n : i)) : i; }

// One transposed damped-Jacobi sweep of the u block: xb (adjoint of the sweep's output) ->
// hb (adjoint of its input; rows and wall faces only) and rhs_b += the rhs adjoint.
template <typename T>
__global__ void __launch_bounds__(NT) k_jacobi_u_T(Grid g, const T* __restrict__ xb, T* __restrict__ hb,
                                                   T* __restrict__ rb, const T* __restrict__ rho_u,
                                                   const T* __restrict__ kap, const T* __restrict__ zeta,
                                                   const T* __restrict__ eta, double dt) {
  int i, j;
  union_ij(g, i, j);
  if (!in_u(g, i, j)) return;
  const int nx = g.nx, ny = g.ny;
  const int i_u_max = g.per_x ? nx : nx - 1;
  const double ix2 = g.inv_dx2, iy2 = g.inv_dy2, omega = 0.8;
  auto row = [&](int ii, int jj) { return jj >= 1 && jj <= ny && ii >= 1 && ii <= i_u_max; };
  auto s_of = [&](int ii, int jj, double& cE, double& cW, double& cN, double& cS) -> double {
    cE = (double)kap[IP(g, ii + 1, jj)] * ix2;
    cW = (double)kap[IP(g, ii, jj)] * ix2;
    cN = ((double)zeta[IW(g, ii, jj)] - (double)eta[IW(g, ii, jj)]) * iy2;
    cS = ((double)zeta[IW(g, ii, jj - 1)] - (double)eta[IW(g, ii, jj - 1)]) * iy2;
    const double diag = ((double)rho_u[IU(g, ii, jj)] / dt) + cE + cW + cN + cS;
    return omega * (double)xb[IU(g, ii, jj)] / diag;
  };
  const bool is_row = row(i, j);
  const bool wall = !g.per_x && j >= 1 && j <= ny && (i == 0 || i == nx);
  double h = 0.0;
  if (is_row) {
    double cE, cW, cN, cS;
    const double s = s_of(i, j, cE, cW, cN, cS);
    rb[IU(g, i, j)] = (T)((double)rb[IU(g, i, j)] + s);
    h = (1.0 - omega) * (double)xb[IU(g, i, j)];
  } else if (wall) {
    rb[IU(g, i, j)] = (T)((double)rb[IU(g, i, j)] + (double)xb[IU(g, i, j)]);
  }
  if (is_row || wall) {
    double cE, cW, cN, cS;
    int ri = wrapi(i - 1, nx, g.per_x), rj = j;   // row whose east neighbour is (i, j)
    if (row(ri, rj)) { const double s = s_of(ri, rj, cE, cW, cN, cS); h += cE * s; }
    ri = wrapi(i + 1, nx, g.per_x);
    if (row(ri, rj)) { const double s = s_of(ri, rj, cE, cW, cN, cS); h += cW * s; }
    ri = i; rj = wrapi(j - 1, ny, g.per_y);
    if (row(ri, rj)) { const double s = s_of(ri, rj, cE, cW, cN, cS); h += cN * s; }
    rj = wrapi(j + 1, ny, g.per_y);
    if (row(ri, rj)) { const double s = s_of(ri, rj, cE, cW, cN, cS); h += cS * s; }
  }
  hb[IU(g, i, j)] = (T)h;
}

template <typename T>
__global__ void __launch_bounds__(NT) k_jacobi_v_T(Grid g, const T* __restrict__ xb, T* __restrict__ hb,
                                                   T* __restrict__ rb, const T* __restrict__ rho_v,
                                                   const T* __restrict__ kap, const T* __restrict__ zeta,
                                                   const T* __restrict__ eta, double dt) {
  int i, j;
  union_ij(g, i, j);
  if (!in_v(g, i, j)) return;
  const int nx = g.nx, ny = g.ny;
  const int j_v_max = g.per_y ? ny : ny - 1;
  const double ix2 = g.inv_dx2, iy2 = g.inv_dy2, omega = 0.8;
  auto row = [&](int ii, int jj) { return jj >= 1 && jj <= j_v_max && ii >= 1 && ii <= nx; };
  auto s_of = [&](int ii, int jj, double& cE, double& cW, double& cN, double& cS) -> double {
    cE = ((double)zeta[IW(g, ii, jj)] - (double)eta[IW(g, ii, jj)]) * ix2;
    cW = ((double)zeta[IW(g, ii - 1, jj)] - (double)eta[IW(g, ii - 1, jj)]) * ix2;
    cN = (double)kap[IP(g, ii, jj + 1)] * iy2;
    cS = (double)kap[IP(g, ii, jj)] * iy2;
    const double diag = ((double)rho_v[IV(g, ii, jj)] / dt) + cE + cW + cN + cS;
    return omega * (double)xb[IV(g, ii, jj)] / diag;
  };
  const bool is_row = row(i, j);
  const bool wall = !g.per_y && i >= 1 && i <= nx && (j == 0 || j == ny);
  double h = 0.0;
  if (is_row) {
    double cE, cW, cN, cS;
    const double s = s_of(i, j, cE, cW, cN, cS);
    rb[IV(g, i, j)] = (T)((double)rb[IV(g, i, j)] + s);
    h = (1.0 - omega) * (double)xb[IV(g, i, j)];
  } else if (wall) {
    rb[IV(g, i, j)] = (T)((double)rb[IV(g, i, j)] + (double)xb[IV(g, i, j)]);
  }
  if (is_row || wall) {
    double cE, cW, cN, cS;
    int ri = wrapi(i - 1, nx, g.per_x), rj = j;
    if (row(ri, rj)) { const double s = s_of(ri, rj, cE, cW, cN, cS); h += cE * s; }
    ri = wrapi(i + 1, nx, g.per_x);
    if (row(ri, rj)) { const double s = s_of(ri, rj, cE, cW, cN, cS); h += cW * s; }
    ri = i; rj = wrapi(j - 1, ny, g.per_y);
    if (row(ri, rj)) { const double s = s_of(ri, rj, cE, cW, cN, cS); h += cN * s; }
    rj = wrapi(j + 1, ny, g.per_y);
    if (row(ri, rj)) { const double s = s_of(ri, rj, cE, cW, cN, cS); h += cS * s; }
  }
  hb[IV(g, i, j)] = (T)h;
}

// rv_b += (d(-cross)/d rv)^T zuw_b: the u block's eta cross term (rows i=1..nx-1, j=1..ny;
// rv(., 0) wraps to rv(., ny) when periodic in y).
template <typename T>
__global__ void __launch_bounds__(NT) k_cross_T(Grid g, const T* __restrict__ zuwb, const T* __restrict__ eta,
                                                T* __restrict__ rvb) {
  int a, b;
  union_ij(g, a, b);
  if (!in_v(g, a, b)) return;
  const int nx = g.nx, ny = g.ny;
  const double f = g.inv_dy * g.inv_dx;
  auto row = [&](int i, int j) { return i >= 1 && i <= nx - 1 && j >= 1 && j <= ny; };
  auto ZB = [&](int i, int j) -> double { return row(i, j) ? (double)zuwb[IU(g, i, j)] : 0.0; };
  auto ET = [&](int i, int j) -> double { return (double)eta[IW(g, i, j)]; };
  double o = 0.0;
  // references of (a, b) as rv(i+1, j) and rv(i, j) of row (i, j) = (a-1, b), (a, b)
  if (b >= 1 && b <= ny) {
    o += ZB(a - 1, b) * f * ET(a - 1, b);
    o -= ZB(a, b) * f * ET(a, b);
  }
  // references as rv(i+1, j-1) and rv(i, j-1) of row (i, j) with j-1 = b (or wrapped)
  int jr = b + 1;
  if (g.per_y && b == ny) jr = 1;   // row j = 1 reads rv(., 0) = rv(., ny)
  if (jr >= 1 && jr <= ny && (b >= 0)) {
    const int bb = jr - 1;           // the vertex row of the forward term
    o -= ZB(a - 1, jr) * f * ET(a - 1, bb);
    o += ZB(a, jr) * f * ET(a, bb);
  }
  if (g.per_y && b == 0) o = 0.0;    // v(., 0) is the image of v(., ny): handled there
  rvb[IV(g, a, b)] = (T)((double)rvb[IV(g, a, b)] + o);
}

// z_b (u, v parts) and rp_b from the adjoints of z_uw, z_vw (src/mod_pcd_precond.f90:68-84);
// rp_b starts from yp_b, the adjoint of the output r_p itself.
template <typename T>
__global__ void __launch_bounds__(NT) k_zwork_T(Grid g, const T* __restrict__ zuwb, const T* __restrict__ zvwb,
                                                const T* __restrict__ ypb, T* __restrict__ zub,
                                                T* __restrict__ zvb, T* __restrict__ rpb) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny;
  if (in_u(g, i, j)) zub[IU(g, i, j)] = (j >= 1 && j <= ny && i >= 0 && i <= nx) ? zuwb[IU(g, i, j)] : T(0);
  if (in_v(g, i, j)) zvb[IV(g, i, j)] = (i >= 1 && i <= nx && j >= 0 && j <= ny) ? zvwb[IV(g, i, j)] : T(0);
  if (own_p(g, i, j)) {
    double o = (double)ypb[IP(g, i, j)];
    if (i - 1 >= 1 && i - 1 <= nx - 1) o -= (double)zuwb[IU(g, i - 1, j)] * g.inv_dx;
    if (i >= 1 && i <= nx - 1) o += (double)zuwb[IU(g, i, j)] * g.inv_dx;
    if (j - 1 >= 1 && j - 1 <= ny - 1) o -= (double)zvwb[IV(g, i, j - 1)] * g.inv_dy;
    if (j >= 1 && j <= ny - 1) o += (double)zvwb[IV(g, i, j)] * g.inv_dy;
    rpb[IP(g, i, j)] = (T)o;
  }
}

// phi_b from rp_b (the transposed convection + 1/dt term, neighbours wrap when periodic) and
// z_p,b += -visc kappa rp_b.
template <typename T>
__global__ void __launch_bounds__(NT) k_rp_T(Grid g, const T* __restrict__ rpb, const T* __restrict__ uadv,
                                             const T* __restrict__ vadv, double dt, const T* __restrict__ kap,
                                             double visc, T* __restrict__ phib, T* __restrict__ zpb) {
  int a, b;
  union_ij(g, a, b);
  if (!in_p(g, a, b)) return;
  if (!own_p(g, a, b)) { phib[IP(g, a, b)] = T(0); return; }
  const int nx = g.nx, ny = g.ny;
  const double half = 0.5;
  auto RB = [&](int i, int j) -> double { return (double)rpb[IP(g, i, j)]; };
  auto uavg = [&](int i, int j) { return half * ((double)uadv[IU(g, i, j)] + (double)uadv[IU(g, i - 1, j)]); };
  auto vavg = [&](int i, int j) { return half * ((double)vadv[IV(g, i, j)] + (double)vadv[IV(g, i, j - 1)]); };
  double o = -RB(a, b) / dt;
  // rows whose east / west / north / south neighbour is (a, b)
  int r = a - 1;
  if (r < 1 && g.per_x) r += nx;
  if (r >= 1) o -= RB(r, b) * uavg(r, b) * (half * g.inv_dx);
  r = a + 1;
  if (r > nx && g.per_x) r -= nx;
  if (r <= nx) o += RB(r, b) * uavg(r, b) * (half * g.inv_dx);
  r = b - 1;
  if (r < 1 && g.per_y) r += ny;
  if (r >= 1) o -= RB(a, r) * vavg(a, r) * (half * g.inv_dy);
  r = b + 1;
  if (r > ny && g.per_y) r -= ny;
  if (r <= ny) o += RB(a, r) * vavg(a, r) * (half * g.inv_dy);
  phib[IP(g, a, b)] = (T)o;
  if (visc != 0.0) zpb[IP(g, a, b)] = (T)((double)zpb[IP(g, a, b)] - visc * (double)kap[IP(g, a, b)] * RB(a, b));
}

// ---------------------------------------------------------------------------------
// Per-cell property path: mu_cell, rho_cell over the padded cells -> operator coefficients
// (edge conventions of src/mod_viscosity.f90:54-151; vertex viscosity = 4 / sum(1/mu_k)).
__global__ void __launch_bounds__(NT) k_cell_props(Grid g, const double* __restrict__ mu, const double* __restrict__ rho,
                                                   double* __restrict__ rho_c, double* __restrict__ rho_u,
                                                   double* __restrict__ rho_v, double* __restrict__ kap,
                                                   double* __restrict__ zeta, double* __restrict__ eta) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  if (in_p(g, i, j)) {
    rho_c[IP(g, i, j)] = rho[IP(g, i, j)];
    kap[IP(g, i, j)] = 2.0 * mu[IP(g, i, j)];
  }
  if (in_u(g, i, j)) {
    const int ii = i < 1 - nh ? 1 - nh : (i > nx + nh - 1 ? nx + nh - 1 : i);
    rho_u[IU(g, i, j)] = 0.5 * (rho[IP(g, ii, j)] + rho[IP(g, ii + 1, j)]);
  }
  if (in_v(g, i, j)) {
    const int jj = j < 1 - nh ? 1 - nh : (j > ny + nh - 1 ? ny + nh - 1 : j);
    rho_v[IV(g, i, j)] = 0.5 * (rho[IP(g, i, jj)] + rho[IP(g, i, jj + 1)]);
  }
  if (in_w(g, i, j)) {
    const int ii = i < 1 - nh ? 1 - nh : (i > nx + nh - 1 ? nx + nh - 1 : i);
    const int jj = j < 1 - nh ? 1 - nh : (j > ny + nh - 1 ? ny + nh - 1 : j);
    const double sinv = 1.0 / mu[IP(g, ii, jj)] + 1.0 / mu[IP(g, ii + 1, jj)] + 1.0 / mu[IP(g, ii, jj + 1)] +
                        1.0 / mu[IP(g, ii + 1, jj + 1)];
    const double mv = 4.0 / sinv;
    zeta[IW(g, i, j)] = 2.0 * mv;
    eta[IW(g, i, j)] = mv;
  }
}

// Transpose of k_cell_props' Jacobian: gradients of the coefficients -> g_mu, g_rho per cell
// (gather: each cell collects from the faces and vertices built from it, edge copies folded).
__global__ void __launch_bounds__(NT) k_cell_props_T(Grid g, const double* __restrict__ mu,
                                                     const double* __restrict__ grc, const double* __restrict__ gru,
                                                     const double* __restrict__ grv, const double* __restrict__ gk,
                                                     const double* __restrict__ gz, const double* __restrict__ ge,
                                                     double* __restrict__ gmu, double* __restrict__ grho) {
  int i, j;
  union_ij(g, i, j);
  if (!in_p(g, i, j)) return;
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  // u faces f = i-1 and f = i are built from cells (f, f+1); the edge faces -nh and nx+nh copy
  // faces 1-nh and nx+nh-1
  auto GU = [&](int f) -> double {
    if (f < 1 - nh || f > nx + nh - 1) return 0.0;
    double s = gru[IU(g, f, j)];
    if (f == 1 - nh) s += gru[IU(g, -nh, j)];
    if (f == nx + nh - 1) s += gru[IU(g, nx + nh, j)];
    return s;
  };
  auto GV = [&](int f) -> double {
    if (f < 1 - nh || f > ny + nh - 1) return 0.0;
    double s = grv[IV(g, i, f)];
    if (f == 1 - nh) s += grv[IV(g, i, -nh)];
    if (f == ny + nh - 1) s += grv[IV(g, i, ny + nh)];
    return s;
  };
  grho[IP(g, i, j)] += grc[IP(g, i, j)] + 0.5 * (GU(i - 1) + GU(i) + GV(j - 1) + GV(j));
  // vertices (vi, vj) in {i-1, i} x {j-1, j} built from this cell; edge vertices fold in
  auto GMV = [&](int vi, int vj) -> double {      // 2 g_zeta + g_eta of the interior vertex, edges folded
    if (vi < 1 - nh || vi > nx + nh - 1 || vj < 1 - nh || vj > ny + nh - 1) return 0.0;
    double s = 0.0;
    const int xs[2] = {vi, vi == 1 - nh ? -nh : (vi == nx + nh - 1 ? nx + nh : vi)};
    const int ys[2] = {vj, vj == 1 - nh ? -nh : (vj == ny + nh - 1 ? ny + nh : vj)};
    for (int a = 0; a < 2; ++a) {
      if (a == 1 && xs[1] == vi) continue;
      for (int b = 0; b < 2; ++b) {
        if (b == 1 && ys[1] == vj) continue;
        s += 2.0 * gz[IW(g, xs[a], ys[b])] + ge[IW(g, xs[a], ys[b])];
      }
    }
    return s;
  };
  const double m = mu[IP(g, i, j)];
  double gm = 2.0 * gk[IP(g, i, j)];
  for (int vj = j - 1; vj <= j; ++vj)
    for (int vi = i - 1; vi <= i; ++vi) {
      const double w = GMV(vi, vj);
      if (w == 0.0) continue;
      const double sinv = 1.0 / mu[IP(g, vi, vj)] + 1.0 / mu[IP(g, vi + 1, vj)] + 1.0 / mu[IP(g, vi, vj + 1)] +
                          1.0 / mu[IP(g, vi + 1, vj + 1)];
      const double mv = 4.0 / sinv;
      gm += w * mv * mv / 4.0 / (m * m);
    }
  gmu[IP(g, i, j)] += gm;
}

// ---------------------------------------------------------------------------------
// Coefficient gradients of s * lam^T A(coef; X) (X padded, as the matvec saw it), accumulated:
// g_rho_c/u/v and g_uadv/vadv by atomics (scatter of each row), g_kappa/zeta/eta by local
// gathers of the adjoint stresses.
constexpr int CG_COEF = 1;   // gradients with respect to rho_c, rho_u, rho_v, kappa, zeta, eta
constexpr int CG_ADV = 2;    // gradients with respect to the advecting velocity
struct CoefGradArgs {
  const double *Lu, *Lv, *Xu, *Xv, *uadv, *vadv, *rho_c, *rho_u, *rho_v, *kap, *zeta, *eta;
  double *grc, *gru, *grv, *gk, *gz, *ge, *gua, *gva;
  double dt, s;
  int adv, what;
};
__global__ void __launch_bounds__(NT) k_coef_grad(Grid g, CoefGradArgs a) {
  int i, j;
  union_ij(g, i, j);
  const double ix = g.inv_dx, iy = g.inv_dy, half = 0.5, s = a.s;
  auto LU = [&](int ii, int jj) -> double { return own_u(g, ii, jj) ? a.Lu[IU(g, ii, jj)] : 0.0; };
  auto LV = [&](int ii, int jj) -> double { return own_v(g, ii, jj) ? a.Lv[IV(g, ii, jj)] : 0.0; };
  auto XU = [&](int ii, int jj) -> double { return a.Xu[IU(g, ii, jj)]; };
  auto XV = [&](int ii, int jj) -> double { return a.Xv[IV(g, ii, jj)]; };
  auto UA = [&](int ii, int jj) -> double { return a.uadv[IU(g, ii, jj)]; };
  auto VA = [&](int ii, int jj) -> double { return a.vadv[IV(g, ii, jj)]; };
  const bool coef = (a.what & CG_COEF) != 0, vel = (a.what & CG_ADV) != 0;
  // u row
  if (own_u(g, i, j)) {
    const double l = s * LU(i, j);
    if (coef) atomicAdd(&a.gru[IU(g, i, j)], l * XU(i, j) / a.dt);
    if (a.adv) {
      const double uE = half * (XU(i, j) + XU(i + 1, j)), uW = half * (XU(i - 1, j) + XU(i, j));
      const double uN = half * (XU(i, j) + XU(i, j + 1)), uS = half * (XU(i, j - 1) + XU(i, j));
      const double aE = half * (UA(i, j) + UA(i + 1, j)), aW = half * (UA(i - 1, j) + UA(i, j));
      const double aN = half * (VA(i, j) + VA(i + 1, j)), aS = half * (VA(i, j - 1) + VA(i + 1, j - 1));
      const double rN = half * (a.rho_v[IV(g, i, j)] + a.rho_v[IV(g, i + 1, j)]);
      const double rS = half * (a.rho_v[IV(g, i, j - 1)] + a.rho_v[IV(g, i + 1, j - 1)]);
      if (coef) {
        atomicAdd(&a.grc[IP(g, i + 1, j)], l * aE * uE * ix);
        atomicAdd(&a.grc[IP(g, i, j)], -l * aW * uW * ix);
        atomicAdd(&a.grv[IV(g, i, j)], l * half * aN * uN * iy);
        atomicAdd(&a.grv[IV(g, i + 1, j)], l * half * aN * uN * iy);
        atomicAdd(&a.grv[IV(g, i, j - 1)], -l * half * aS * uS * iy);
        atomicAdd(&a.grv[IV(g, i + 1, j - 1)], -l * half * aS * uS * iy);
      }
      if (vel) {
        const double e = l * a.rho_c[IP(g, i + 1, j)] * half * uE * ix;
        const double w = l * a.rho_c[IP(g, i, j)] * half * uW * ix;
        atomicAdd(&a.gua[IU(g, i, j)], e - w);
        atomicAdd(&a.gua[IU(g, i + 1, j)], e);
        atomicAdd(&a.gua[IU(g, i - 1, j)], -w);
        const double n = l * rN * half * uN * iy, so = l * rS * half * uS * iy;
        atomicAdd(&a.gva[IV(g, i, j)], n);
        atomicAdd(&a.gva[IV(g, i + 1, j)], n);
        atomicAdd(&a.gva[IV(g, i, j - 1)], -so);
        atomicAdd(&a.gva[IV(g, i + 1, j - 1)], -so);
      }
    }
  }
  // v row
  if (own_v(g, i, j)) {
    const double l = s * LV(i, j);
    if (coef) atomicAdd(&a.grv[IV(g, i, j)], l * XV(i, j) / a.dt);
    if (a.adv) {
      const double vE = half * (XV(i, j) + XV(i + 1, j)), vW = half * (XV(i - 1, j) + XV(i, j));
      const double vN = half * (XV(i, j) + XV(i, j + 1)), vS = half * (XV(i, j - 1) + XV(i, j));
      const double aE = half * (UA(i, j) + UA(i, j + 1)), aW = half * (UA(i - 1, j) + UA(i - 1, j + 1));
      const double aN = half * (VA(i, j) + VA(i, j + 1)), aS = half * (VA(i, j - 1) + VA(i, j));
      const double rE = half * (a.rho_u[IU(g, i, j)] + a.rho_u[IU(g, i, j + 1)]);
      const double rW = half * (a.rho_u[IU(g, i - 1, j)] + a.rho_u[IU(g, i - 1, j + 1)]);
      if (coef) {
        atomicAdd(&a.gru[IU(g, i, j)], l * half * aE * vE * ix);
        atomicAdd(&a.gru[IU(g, i, j + 1)], l * half * aE * vE * ix);
        atomicAdd(&a.gru[IU(g, i - 1, j)], -l * half * aW * vW * ix);
        atomicAdd(&a.gru[IU(g, i - 1, j + 1)], -l * half * aW * vW * ix);
        atomicAdd(&a.grc[IP(g, i, j + 1)], l * aN * vN * iy);
        atomicAdd(&a.grc[IP(g, i, j)], -l * aS * vS * iy);
      }
      if (vel) {
        const double e = l * rE * half * vE * ix, w = l * rW * half * vW * ix;
        atomicAdd(&a.gua[IU(g, i, j)], e);
        atomicAdd(&a.gua[IU(g, i, j + 1)], e);
        atomicAdd(&a.gua[IU(g, i - 1, j)], -w);
        atomicAdd(&a.gua[IU(g, i - 1, j + 1)], -w);
        const double n = l * a.rho_c[IP(g, i, j + 1)] * half * vN * iy;
        const double so = l * a.rho_c[IP(g, i, j)] * half * vS * iy;
        atomicAdd(&a.gva[IV(g, i, j)], n - so);
        atomicAdd(&a.gva[IV(g, i, j + 1)], n);
        atomicAdd(&a.gva[IV(g, i, j - 1)], -so);
      }
    }
  }
  if (!coef) return;
  // kappa at cells (forward Txx for i = 0..nx+1, j = 1..ny; Tyy for i = 1..nx, j = 0..ny+1)
  if (in_p(g, i, j)) {
    double o = 0.0;
    if (j >= 1 && j <= g.ny && i >= 0 && i <= g.nx + 1)
      o += (LU(i, j) - LU(i - 1, j)) * ix * (XU(i, j) - XU(i - 1, j)) * ix;
    if (i >= 1 && i <= g.nx && j >= 0 && j <= g.ny + 1)
      o += (LV(i, j) - LV(i, j - 1)) * iy * (XV(i, j) - XV(i, j - 1)) * iy;
    a.gk[IP(g, i, j)] += s * o;
  }
  // zeta, eta at vertices i = 0..nx, j = 0..ny
  if (i >= 0 && i <= g.nx && j >= 0 && j <= g.ny) {
    const double txy = (LU(i, j + 1) - LU(i, j)) * iy, tyx = (LV(i + 1, j) - LV(i, j)) * ix;
    const double du = (XU(i, j + 1) - XU(i, j)) * iy, dv = (XV(i + 1, j) - XV(i, j)) * ix;
    a.gz[IW(g, i, j)] += s * (txy * du + tyx * dv);
    a.ge[IW(g, i, j)] += s * (txy - tyx) * (dv - du);
  }
}

// Elementwise helpers of the reverse sweep.
// o = a * x + b * y over n entries (either of x, y may be null).
__global__ void k_axpby(double* __restrict__ o, const double* __restrict__ x, double a, const double* __restrict__ y,
                        double b, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    double v = 0.0;
    if (x) v += a * x[k];
    if (y) v += b * y[k];
    o[k] = v;
  }
}
// o += c * w * x * y (w, y may be null -> 1) over n entries.
__global__ void k_acc_prod(double* __restrict__ o, double c, const double* __restrict__ w,
                           const double* __restrict__ x, const double* __restrict__ y, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    double v = c * x[k];
    if (w) v *= w[k];
    if (y) v *= y[k];
    o[k] += v;
  }
}

// zp(i, j) += phi(i, j) - sc[slot] over the interior (the projected pressure block).
template <typename T>
__global__ void __launch_bounds__(NT) k_add_centred(Grid g, T* __restrict__ zp, const T* __restrict__ phi,
                                                    const double* __restrict__ sc, int slot) {
  int i, j;
  union_ij(g, i, j);
  if (own_p(g, i, j)) zp[IP(g, i, j)] = (T)((double)zp[IP(g, i, j)] + ((double)phi[IP(g, i, j)] - sc[slot]));
}

// Objective band masks: owned u faces with (j - 1/2) dy >= y0, owned v faces with j dy >= y0.
__global__ void __launch_bounds__(NT) k_band_mask(Grid g, double y0, double* __restrict__ mu, double* __restrict__ mv) {
  int i, j;
  union_ij(g, i, j);
  if (in_u(g, i, j)) mu[IU(g, i, j)] = (own_u(g, i, j) && (j - 0.5) * g.dy >= y0) ? 1.0 : 0.0;
  if (in_v(g, i, j)) mv[IV(g, i, j)] = (own_v(g, i, j) && j * g.dy >= y0) ? 1.0 : 0.0;
}

}  // namespace fugu
