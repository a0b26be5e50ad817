// fugu2d_kernels.hpp -- HIP kernels of the FUGU-2D per-step hot path (gfx950).
//
// Every kernel is a memory-bound stencil or stream: no MFMA (nothing here is a dense
// contraction).  FP64 kernels are compiled with -ffp-contract=off and keep the
// reference's operation order, so element-wise results are bit-identical to the
// reference arithmetic; reductions are deterministic (fixed-shape tree).  The same
// templates instantiated with T = float are the inner kernels of the mixed solve
// (accumulation in FP64).
#pragma once
#include "fugu2d_common.hpp"

namespace fugu {

// ---------------------------------------------------------------------------------
// Device-resident solver scalars (one array per context).  Slots are written only by
// the last workgroup of a reduction kernel (thread 0) or by single-thread kernels.
enum Slot : int {
  S_NORM_B = 0, S_NORM_R, S_REL_RES, S_RHO_OLD, S_ALPHA, S_BETA, S_OMEGA,
  S_XI0, S_GAMMA, S_OM1, S_OM2, S_TOL, S_SCALE,
  // PCG (inner pressure solve)
  S_P_RZ, S_P_PAP, S_P_ALPHA, S_P_BETA, S_P_RHSN, S_P_RES, S_P_MEAN, S_P_TOL,
  // generic reduction outputs
  S_DOT0, S_DOT1, S_DOT2, S_DOT3, S_DOT4, S_MAX0,
  S_GSCALE,      // scale of the breakdown guards (1 = the reference's absolute 1e-30 tests)
  S_INVH,        // 1 / h_{k+1,k} of the outer FGMRES
  S_FG_TOL,      // the outer FGMRES cycle's absolute stopping level for its residual estimate
  S_VOL_OLD, S_VOL_NEW, S_VOL_CLAMP,   // the step's volume budget: sum C vol before, after, clamp part
  S_HBUF,        // 256 slots: dot products / combination coefficients of the outer FGMRES
  S_COUNT = S_HBUF + 256
};
enum Flag : int {
  F_BRK = 0,     // exit in the middle of a Krylov iteration: skip everything after
  F_DONE,        // no further Krylov iteration
  F_ITER,        // iteration number reported by the Krylov solver
  F_P_DONE,      // PCG: 1 = stopped (tolerance / breakdown), 2 = zero-rhs early exit
  F_P_ITER,      // PCG iteration count
  F_S_DONE,      // SVOF converged
  F_S_ITER,
  F_RELG,        // 1: breakdown guards relative to ||P^-1 b||^2 (chosen fast forms only)
  F_FG_DONE,     // outer FGMRES cycle converged
  F_REORTH,      // outer FGMRES: second Gram-Schmidt pass needed
  F_S_RBLK,      // SVOF in blocks: the block whose iterations are redone up to the converged one
  F_S_RCNT,      // ... and how many of them
  F_KITER,       // the reference Krylov solvers' iteration number (bumped by k_iter_begin)
  F_COUNT
};

// Reduction phases: scalar logic run by the last workgroup of a dot kernel.  Each one
// restates a few lines of src/mod_bicgstab2.f90 (cited at the case).
enum Phase : int {
  PH_NONE = 0,
  PH_NORM_B,      // :147-148 / :345-346
  PH_B1_INIT,     // :173-175   tot = {(r0,r), (r,r)}
  PH_B1_RV,       // :196-198
  PH_B1_TT_TS,    // :217-223   tot = {(t,t), (t,s)}
  PH_B1_RES,      // :233-242   tot = {(r,r), (r0,r)}
  PH_B2_INIT,     // :380-385   tot = {(r,r)}
  PH_B2_XI_EVEN,  // :387-401   tot = {(rt,r)}
  PH_B2_GAMMA,    // :421-423 / :472-474
  PH_B2_XI_ODD,   // :449-452   tot = {(rt,s)}
  PH_B2_GCR,      // :504-530   tot = {(r,s),(s,s),(s,t),(t,t),(r,t)}
  PH_B2_RES,      // :546-547 (+ the loop-top test :385 of the next iteration)
  PH_STORE,       // plain: sc[S_DOT0..] = tot
  PH_HSTORE = 64  // PH_HSTORE + o: sc[S_HBUF + o ..] = tot
};

__device__ inline void krylov_phase(int phase, int k, int max_iter, double* sc, int* fl,
                                    const double* tot) {
  if (k < 0) k = fl[F_KITER];   // the iteration number lives on the device (replayable launches)
  // guard thresholds for dot-product-valued scalars: the reference's absolute 1e-30, or
  // 1e-30 * ||P^-1 b||^2 when relative guards are on
  const double G = sc[S_GSCALE];
  const double tiny = 1.0e-30 * G;
  switch (phase) {
    case PH_NORM_B: {
      double nb = sqrt(tot[0]);
      if (nb < 1.0e-14 && !fl[F_RELG]) nb = 1.0;
      if (nb == 0.0) nb = 1.0;
      sc[S_NORM_B] = nb;
      sc[S_GSCALE] = fl[F_RELG] ? nb * nb : 1.0;
    } break;
    case PH_B1_INIT: {
      sc[S_RHO_OLD] = tot[0];
      sc[S_NORM_R] = sqrt(tot[1]);
      sc[S_REL_RES] = sc[S_NORM_R] / sc[S_NORM_B];
      sc[S_ALPHA] = 0.0;
      sc[S_OMEGA] = 1.0;
      if (sc[S_REL_RES] < sc[S_TOL]) { fl[F_DONE] = 1; fl[F_ITER] = 1; }
    } break;
    case PH_B1_RV: {
      const double rv = tot[0];
      if (fabs(rv) < tiny) { fl[F_BRK] = 1; fl[F_DONE] = 1; fl[F_ITER] = k; break; }
      sc[S_ALPHA] = sc[S_RHO_OLD] / rv;
    } break;
    case PH_B1_TT_TS: {
      const double tt = tot[0], ts = tot[1];
      sc[S_OMEGA] = (tt > tiny) ? ts / tt : 0.0;
    } break;
    case PH_B1_RES: {
      sc[S_NORM_R] = sqrt(tot[0]);
      sc[S_REL_RES] = sc[S_NORM_R] / sc[S_NORM_B];
      if (sc[S_REL_RES] < sc[S_TOL] || fabs(sc[S_OMEGA]) < 1.0e-30) {
        fl[F_BRK] = 1; fl[F_DONE] = 1; fl[F_ITER] = k; break;
      }
      const double rho_new = tot[1];
      if (fabs(sc[S_RHO_OLD]) < tiny) { fl[F_BRK] = 1; fl[F_DONE] = 1; fl[F_ITER] = k; break; }
      sc[S_BETA] = (rho_new / sc[S_RHO_OLD]) * (sc[S_ALPHA] / sc[S_OMEGA]);
      sc[S_RHO_OLD] = rho_new;
      if (k == max_iter) { fl[F_DONE] = 1; fl[F_ITER] = max_iter; }
    } break;
    case PH_B2_INIT: {
      sc[S_XI0] = 1.0; sc[S_OM1] = 0.0; sc[S_OM2] = 1.0; sc[S_ALPHA] = 0.0;
      sc[S_NORM_R] = sqrt(tot[0]);
      sc[S_REL_RES] = sc[S_NORM_R] / sc[S_NORM_B];
      if (sc[S_REL_RES] < sc[S_TOL]) { fl[F_DONE] = 1; fl[F_ITER] = 1; }
    } break;
    case PH_B2_XI_EVEN: {
      double xi0 = sc[S_XI0];
      if (fabs(sc[S_OM2]) > 1.0e-14) xi0 = -sc[S_OM2] * xi0;
      else if (fabs(sc[S_OM1]) > 1.0e-14) xi0 = -sc[S_OM1] * xi0;
      else xi0 = -xi0;
      const double xi1 = tot[0];
      if (fabs(xi0) < tiny) { sc[S_XI0] = xi0; fl[F_BRK] = 1; fl[F_DONE] = 1; fl[F_ITER] = k; break; }
      sc[S_BETA] = sc[S_ALPHA] * (xi1 / xi0);
      sc[S_XI0] = xi1;
    } break;
    case PH_B2_GAMMA: {
      const double gam = tot[0];
      if (fabs(gam) < tiny) { fl[F_BRK] = 1; fl[F_DONE] = 1; fl[F_ITER] = k; break; }
      sc[S_ALPHA] = sc[S_XI0] / gam;
    } break;
    case PH_B2_XI_ODD: {
      const double xi1 = tot[0];
      if (fabs(sc[S_XI0]) < tiny) { fl[F_BRK] = 1; fl[F_DONE] = 1; fl[F_ITER] = k; break; }
      sc[S_BETA] = sc[S_ALPHA] * (xi1 / sc[S_XI0]);
      sc[S_XI0] = xi1;
    } break;
    case PH_B2_GCR: {
      double om1 = tot[0];
      const double lam = tot[1], nu = tot[2], tau = tot[3];
      double om2 = tot[4];
      if (lam * tau > tiny * G) {
        const double det = lam * tau - nu * nu;
        if (det > 1.0e-14 * (lam * tau)) {
          const double c1 = om1, c2 = om2;
          om1 = (c1 * tau - c2 * nu) / det;
          om2 = (c2 * lam - c1 * nu) / det;
        } else if (lam > tiny) {
          om1 = om1 / lam; om2 = 0.0;
        } else {
          om1 = 0.0; om2 = 0.0;
        }
      } else if (lam > tiny) {
        om1 = om1 / lam; om2 = 0.0;
      } else {
        om1 = 0.0; om2 = 0.0;
      }
      sc[S_OM1] = om1; sc[S_OM2] = om2;
    } break;
    case PH_B2_RES: {
      sc[S_NORM_R] = sqrt(tot[0]);
      sc[S_REL_RES] = sc[S_NORM_R] / sc[S_NORM_B];
      // the exit test sits at the top of the next iteration (:385): X still gets its BCs
      if (sc[S_REL_RES] < sc[S_TOL]) { fl[F_DONE] = 1; fl[F_ITER] = (k + 1 < max_iter) ? k + 1 : max_iter; }
      else if (k == max_iter) { fl[F_DONE] = 1; fl[F_ITER] = max_iter; }
    } break;
    case PH_STORE:
      for (int q = 0; q < 5; ++q) sc[S_DOT0 + q] = tot[q];
      break;
    default:   // PH_HSTORE + offset: sc[S_HBUF + offset + q] = tot[q]
      if (phase >= PH_HSTORE)
        for (int q = 0; q < 5; ++q)
          if (S_HBUF + (phase - PH_HSTORE) + q < (int)S_COUNT) sc[S_HBUF + (phase - PH_HSTORE) + q] = tot[q];
      break;
  }
}

// ---------------------------------------------------------------------------------
// Ghost fills.  Periodic halo exchange at one rank (src/mod_mpi.f90:113-245): x pass
// over full columns then y pass over full rows, composed so every target reads only
// non-target points (race free in one launch).  kind: 0 = u, 1 = v, 2 = p.

template <typename T>
__global__ void k_halo(Grid g, T* __restrict__ a, int kind, const int* __restrict__ skip) {
  if (skip && *skip) return;
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  // index ranges of the kind
  const int i_lo = (kind == 0) ? -nh : 1 - nh, i_hi = nx + nh;
  const int j_lo = (kind == 1) ? -nh : 1 - nh, j_hi = ny + nh;
  if (i < i_lo || i > i_hi || j < j_lo || j > j_hi) return;
  auto at = [&](int ii, int jj) -> size_t {
    return kind == 0 ? IU(g, ii, jj) : (kind == 1 ? IV(g, ii, jj) : IP(g, ii, jj));
  };
  // x pass targets: columns 1-nh..0 and nx+1..nx+nh (u: -1..0, so u(-nh) untouched)
  auto xsrc = [&](int ii) -> int {
    if (!g.per_x) return ii;
    if (ii >= 1 - nh && ii <= 0) return ii + nx;
    if (ii >= nx + 1 && ii <= nx + nh) return ii - nx;
    return ii;
  };
  auto ysrc = [&](int jj) -> int {
    if (!g.per_y) return jj;
    if (jj >= 1 - nh && jj <= 0) return jj + ny;
    if (jj >= ny + 1 && jj <= ny + nh) return jj - ny;
    return jj;
  };
  const int si = xsrc(i), sj = ysrc(j);
  if (si == i && sj == j) return;
  a[at(i, j)] = a[at(si, sj)];
}

// Homogeneous / lid velocity BCs (src/mod_operators.f90:254-333).  x walls first, then
// y walls; the y-pass value of a corner is formed from the x-pass value of its source.
template <typename T>
__device__ inline T u_after_x(const Grid& g, const T* u, int i, int j) {
  if (g.per_x) return u[IU(g, i, j)];
  const int nx = g.nx, nh = g.nh;
  if (i == 0 || i == nx) return T(0);
  if (i < 0 && i >= -nh) return -u[IU(g, -i, j)];
  if (i > nx && i <= nx + nh) return -u[IU(g, 2 * nx - i, j)];
  return u[IU(g, i, j)];
}
template <typename T>
__device__ inline T v_after_x(const Grid& g, const T* v, int i, int j) {
  if (g.per_x) return v[IV(g, i, j)];
  const int nx = g.nx, nh = g.nh;
  if (i <= 0 && i >= 1 - nh) return -v[IV(g, 1 - i, j)];
  if (i > nx && i <= nx + nh) return -v[IV(g, 2 * nx + 1 - i, j)];
  return v[IV(g, i, j)];
}

template <typename T>
__global__ void k_velocity_bcs(Grid g, T* __restrict__ u, T* __restrict__ v, double u_lid,
                               const int* __restrict__ skip) {
  if (skip && *skip) return;
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  // u target
  if (i >= -nh && i <= nx + nh && j >= 1 - nh && j <= ny + nh) {
    const bool xt = !g.per_x && (i <= 0 || i >= nx);
    const bool yt = !g.per_y && (j <= 0 || j >= ny + 1);
    if (yt) {
      T val;
      if (j <= 0) val = -u_after_x(g, u, i, 1 - j);
      else val = (T)(2.0 * u_lid) - u_after_x(g, u, i, 2 * ny + 1 - j);
      u[IU(g, i, j)] = val;
    } else if (xt) {
      u[IU(g, i, j)] = u_after_x(g, u, i, j);
    }
  }
  // v target
  if (i >= 1 - nh && i <= nx + nh && j >= -nh && j <= ny + nh) {
    const bool xt = !g.per_x && (i <= 0 || i >= nx + 1);
    const bool yt = !g.per_y && (j <= 0 || j >= ny);
    if (yt) {
      T val;
      if (j == 0 || j == ny) val = T(0);
      else if (j < 0) val = -v_after_x(g, v, i, -j);
      else val = -v_after_x(g, v, i, 2 * ny - j);
      v[IV(g, i, j)] = val;
    } else if (xt) {
      v[IV(g, i, j)] = v_after_x(g, v, i, j);
    }
  }
}

// Neumann mirror of a cell field (src/mod_subsolvers.f90:378-416, src/mod_vof.f90:199-237):
// x walls then y walls, composed.
template <typename T>
__global__ void k_mirror_p(Grid g, T* __restrict__ a) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  if (i < 1 - nh || i > nx + nh || j < 1 - nh || j > ny + nh) return;
  int si = i, sj = j;
  if (!g.per_x) {
    if (i <= 0) si = 1 - i;
    else if (i > nx) si = 2 * nx + 1 - i;
  }
  if (!g.per_y) {
    if (j <= 0) sj = 1 - j;
    else if (j > ny) sj = 2 * ny + 1 - j;
  }
  if (si == i && sj == j) return;
  a[IP(g, i, j)] = a[IP(g, si, sj)];
}

// ---------------------------------------------------------------------------------
// coupled matvec Y = A X (src/mod_operators.f90:341-443 with :99-162, :60-88,
// :168-249, :34-53).  One pass over the union grid writes Y_u, Y_v, Y_p including
// zeroed ghosts; the stress tensors are recomputed per use (same arithmetic as the
// reference's temporaries).  The periodic halo of X is done before (k_halo).
enum MvFlags : int { MV_ADV = 1, MV_PEN = 2, MV_CONS = 4 };

template <typename T>
struct MatvecArgs {
  const T *Xu, *Xv, *Xp;
  T *Yu, *Yv, *Yp;
  const T *uadv, *vadv, *rho_c, *rho_u, *rho_v, *kap, *zeta, *eta;
  const T *penu, *penv, *mdu, *mdv;
  double dt;
  int flags;
};

// Operands of one matvec point, read straight from the padded arrays.
template <typename T>
struct MvGlobal {
  const Grid& g;
  const MatvecArgs<T>& a;
  __device__ T xu(int i, int j) const { return a.Xu[IU(g, i, j)]; }
  __device__ T xv(int i, int j) const { return a.Xv[IV(g, i, j)]; }
  __device__ T xp(int i, int j) const { return a.Xp[IP(g, i, j)]; }
  __device__ T kap(int i, int j) const { return a.kap[IP(g, i, j)]; }
  __device__ T ze(int i, int j) const { return a.zeta[IW(g, i, j)]; }
  __device__ T et(int i, int j) const { return a.eta[IW(g, i, j)]; }
  __device__ T ru(int i, int j) const { return a.rho_u[IU(g, i, j)]; }
  __device__ T rv(int i, int j) const { return a.rho_v[IV(g, i, j)]; }
  __device__ T rc(int i, int j) const { return a.rho_c[IP(g, i, j)]; }
  __device__ T ua(int i, int j) const { return a.uadv[IU(g, i, j)]; }
  __device__ T va(int i, int j) const { return a.vadv[IV(g, i, j)]; }
  __device__ T pu(int i, int j) const { return a.penu[IU(g, i, j)]; }
  __device__ T pv(int i, int j) const { return a.penv[IV(g, i, j)]; }
  __device__ T mdu(int i, int j) const { return a.mdu[IU(g, i, j)]; }
  __device__ T mdv(int i, int j) const { return a.mdv[IV(g, i, j)]; }
};

// Operands of one matvec point from LDS tiles of the arrays: the workgroup's block of points
// grown by one cell on every side (every operand of a point lies within one cell of it).
constexpr int MV_TW = TX + 2;
enum MvQ : int { MQ_XU, MQ_XV, MQ_XP, MQ_KAP, MQ_ZE, MQ_ET, MQ_RU, MQ_RV, MQ_RC, MQ_UA, MQ_VA, MQ_PU, MQ_PV,
                 MQ_MU, MQ_MV, MQ_N };
template <typename T, int TYT>
struct MvTile {
  const T (*s)[TYT + 2][MV_TW];
  int i0, j0;   // (i, j) of tile element [0][0]
  __device__ T at(int q, int i, int j) const { return s[q][j - j0][i - i0]; }
  __device__ T xu(int i, int j) const { return at(MQ_XU, i, j); }
  __device__ T xv(int i, int j) const { return at(MQ_XV, i, j); }
  __device__ T xp(int i, int j) const { return at(MQ_XP, i, j); }
  __device__ T kap(int i, int j) const { return at(MQ_KAP, i, j); }
  __device__ T ze(int i, int j) const { return at(MQ_ZE, i, j); }
  __device__ T et(int i, int j) const { return at(MQ_ET, i, j); }
  __device__ T ru(int i, int j) const { return at(MQ_RU, i, j); }
  __device__ T rv(int i, int j) const { return at(MQ_RV, i, j); }
  __device__ T rc(int i, int j) const { return at(MQ_RC, i, j); }
  __device__ T ua(int i, int j) const { return at(MQ_UA, i, j); }
  __device__ T va(int i, int j) const { return at(MQ_VA, i, j); }
  __device__ T pu(int i, int j) const { return at(MQ_PU, i, j); }
  __device__ T pv(int i, int j) const { return at(MQ_PV, i, j); }
  __device__ T mdu(int i, int j) const { return at(MQ_MU, i, j); }
  __device__ T mdv(int i, int j) const { return at(MQ_MV, i, j); }
};

// Storage type T (double: the drop-in; float: the mixed form's inner operator); the
// arithmetic is always FP64 in registers, so FP32 storage does not turn the cancellation
// inside a row (rho/dt X against the stress and pressure terms) into operator error.
template <typename T, typename S>
__device__ inline void matvec_point(const Grid& g, const MatvecArgs<T>& a, const S& s, int i, int j) {
  using R = double;
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  const R inv_dx = g.inv_dx, inv_dy = g.inv_dy, half = 0.5;
  const int i_u_max = g.per_x ? nx : nx - 1;
  const int j_v_max = g.per_y ? ny : ny - 1;
  const bool with_adv = (a.flags & MV_ADV) != 0;
  const bool cons = (a.flags & MV_CONS) != 0;
  const R dt = (R)a.dt;
  auto XU = [&](int ii, int jj) -> R { return (R)s.xu(ii, jj); };
  auto XV = [&](int ii, int jj) -> R { return (R)s.xv(ii, jj); };

  auto Txx = [&](int ii, int jj) -> R { return (R)s.kap(ii, jj) * (XU(ii, jj) - XU(ii - 1, jj)) * inv_dx; };
  auto Tyy = [&](int ii, int jj) -> R { return (R)s.kap(ii, jj) * (XV(ii, jj) - XV(ii, jj - 1)) * inv_dy; };
  auto dudy = [&](int ii, int jj) -> R { return (XU(ii, jj + 1) - XU(ii, jj)) * inv_dy; };
  auto dvdx = [&](int ii, int jj) -> R { return (XV(ii + 1, jj) - XV(ii, jj)) * inv_dx; };
  auto Txy = [&](int ii, int jj) -> R {
    const R ze = s.ze(ii, jj), et = s.et(ii, jj);
    return (ze - et) * dudy(ii, jj) + et * dvdx(ii, jj);
  };
  auto Tyx = [&](int ii, int jj) -> R {
    const R ze = s.ze(ii, jj), et = s.et(ii, jj);
    return et * dudy(ii, jj) + (ze - et) * dvdx(ii, jj);
  };

  // ---- u rows
  if (i >= -nh && i <= nx + nh && j >= 1 - nh && j <= ny + nh) {
    R y = 0.0;
    if (j >= 1 && j <= ny && i >= 1 && i <= i_u_max) {
      const R visc = -((Txx(i + 1, j) - Txx(i, j)) * inv_dx + (Txy(i, j) - Txy(i, j - 1)) * inv_dy);
      const R grad = ((R)s.xp(i + 1, j) - (R)s.xp(i, j)) * inv_dx;
      R adv = 0.0;
      if (with_adv) {
        R mE, mW, mN, mS;
        if (cons) {
          mE = half * ((R)s.mdu(i, j) + (R)s.mdu(i + 1, j));
          mW = half * ((R)s.mdu(i - 1, j) + (R)s.mdu(i, j));
          mN = half * ((R)s.mdv(i, j) + (R)s.mdv(i + 1, j));
          mS = half * ((R)s.mdv(i, j - 1) + (R)s.mdv(i + 1, j - 1));
        } else {
          mE = (R)s.rc(i + 1, j) * half * ((R)s.ua(i, j) + (R)s.ua(i + 1, j));
          mW = (R)s.rc(i, j) * half * ((R)s.ua(i - 1, j) + (R)s.ua(i, j));
          mN = half * ((R)s.rv(i, j) + (R)s.rv(i + 1, j)) * half *
               ((R)s.va(i, j) + (R)s.va(i + 1, j));
          mS = half * ((R)s.rv(i, j - 1) + (R)s.rv(i + 1, j - 1)) * half *
               ((R)s.va(i, j - 1) + (R)s.va(i + 1, j - 1));
        }
        const R uc = XU(i, j);
        const R uE = half * (uc + XU(i + 1, j));
        const R uW = half * (XU(i - 1, j) + uc);
        const R uN = half * (uc + XU(i, j + 1));
        const R uS = half * (XU(i, j - 1) + uc);
        adv = (mE * uE - mW * uW) * inv_dx + (mN * uN - mS * uS) * inv_dy;
      }
      const R x = XU(i, j);
      y = ((R)s.ru(i, j) / dt) * x + adv + visc + grad;
      if (a.flags & MV_PEN) y = y + (R)s.pu(i, j) * x;
    } else if (!g.per_x && j >= 1 && j <= ny && (i == 0 || i == nx)) {
      y = XU(i, j);
    }
    a.Yu[IU(g, i, j)] = (T)y;
  }
  // ---- v rows
  if (i >= 1 - nh && i <= nx + nh && j >= -nh && j <= ny + nh) {
    R y = 0.0;
    if (j >= 1 && j <= j_v_max && i >= 1 && i <= nx) {
      const R visc = -((Tyx(i, j) - Tyx(i - 1, j)) * inv_dx + (Tyy(i, j + 1) - Tyy(i, j)) * inv_dy);
      const R grad = ((R)s.xp(i, j + 1) - (R)s.xp(i, j)) * inv_dy;
      R adv = 0.0;
      if (with_adv) {
        R mE, mW, mN, mS;
        if (cons) {
          mE = half * ((R)s.mdu(i, j) + (R)s.mdu(i, j + 1));
          mW = half * ((R)s.mdu(i - 1, j) + (R)s.mdu(i - 1, j + 1));
          mN = half * ((R)s.mdv(i, j) + (R)s.mdv(i, j + 1));
          mS = half * ((R)s.mdv(i, j - 1) + (R)s.mdv(i, j));
        } else {
          mE = half * ((R)s.ru(i, j) + (R)s.ru(i, j + 1)) * half *
               ((R)s.ua(i, j) + (R)s.ua(i, j + 1));
          mW = half * ((R)s.ru(i - 1, j) + (R)s.ru(i - 1, j + 1)) * half *
               ((R)s.ua(i - 1, j) + (R)s.ua(i - 1, j + 1));
          mN = (R)s.rc(i, j + 1) * half * ((R)s.va(i, j) + (R)s.va(i, j + 1));
          mS = (R)s.rc(i, j) * half * ((R)s.va(i, j - 1) + (R)s.va(i, j));
        }
        const R vc = XV(i, j);
        const R vE = half * (vc + XV(i + 1, j));
        const R vW = half * (XV(i - 1, j) + vc);
        const R vN = half * (vc + XV(i, j + 1));
        const R vS = half * (XV(i, j - 1) + vc);
        adv = (mE * vE - mW * vW) * inv_dx + (mN * vN - mS * vS) * inv_dy;
      }
      const R x = XV(i, j);
      y = ((R)s.rv(i, j) / dt) * x + adv + visc + grad;
      if (a.flags & MV_PEN) y = y + (R)s.pv(i, j) * x;
    } else if (!g.per_y && i >= 1 && i <= nx && (j == 0 || j == ny)) {
      y = XV(i, j);
    }
    a.Yv[IV(g, i, j)] = (T)y;
  }
  // ---- p rows: divergence
  if (i >= 1 - nh && i <= nx + nh && j >= 1 - nh && j <= ny + nh) {
    R y = 0.0;
    if (i >= 1 && i <= nx && j >= 1 && j <= ny)
      y = (XU(i, j) - XU(i - 1, j)) * inv_dx + (XV(i, j) - XV(i, j - 1)) * inv_dy;
    a.Yp[IP(g, i, j)] = (T)y;
  }
}

template <typename T>
__global__ void __launch_bounds__(NT) k_matvec(Grid g, MatvecArgs<T> a, const int* __restrict__ skip) {
  if (skip && *skip) return;
  int i, j;
  union_ij(g, i, j);
  matvec_point(g, a, MvGlobal<T>{g, a}, i, j);
}

// Array kinds of the padded layout: 0 u-kind, 1 v-kind, 2 p-kind, 3 vertex kind.
__device__ inline bool kind_has(const Grid& g, int kind, int i, int j) {
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  const int ilo = (kind == 0 || kind == 3) ? -nh : 1 - nh, jlo = (kind == 1 || kind == 3) ? -nh : 1 - nh;
  return i >= ilo && i <= nx + nh && j >= jlo && j <= ny + nh;
}
__device__ inline size_t kind_at(const Grid& g, int kind, int i, int j) {
  return kind == 0 ? IU(g, i, j) : (kind == 1 ? IV(g, i, j) : (kind == 2 ? IP(g, i, j) : IW(g, i, j)));
}

// The same matvec with every operand staged once per workgroup: TX x TYT points, their arrays'
// tiles (one halo cell) loaded with all loads in flight, then each point computed from LDS by
// the same arithmetic (bit-identical output).  Arrays a flag switches off are not loaded.
template <typename T, int TYT>
__global__ void __launch_bounds__(TX * TYT) k_matvec_tile(Grid g, MatvecArgs<T> a, const int* __restrict__ skip) {
  if (skip && *skip) return;
  __shared__ T s[MQ_N][TYT + 2][MV_TW];
  int bx, by;
  tile_coords(bx, by);
  const int ib = bx * TX - g.nh, jb = by * TYT - g.nh;   // the block's first point
  const int i0 = ib - 1, j0 = jb - 1;                   // tile origin
  const int tid = (int)threadIdx.x + TX * (int)threadIdx.y;
  constexpr int NTILE = (TYT + 2) * MV_TW;
  constexpr int NE = (NTILE + TX * TYT - 1) / (TX * TYT);
  const bool adv = (a.flags & MV_ADV) != 0, cons = (a.flags & MV_CONS) != 0, pen = (a.flags & MV_PEN) != 0;
  const T* src[MQ_N] = {a.Xu, a.Xv, a.Xp, a.kap, a.zeta, a.eta, a.rho_u, a.rho_v,
                        (adv && !cons) ? a.rho_c : nullptr, (adv && !cons) ? a.uadv : nullptr,
                        (adv && !cons) ? a.vadv : nullptr, pen ? a.penu : nullptr, pen ? a.penv : nullptr,
                        (adv && cons) ? a.mdu : nullptr, (adv && cons) ? a.mdv : nullptr};
  constexpr int kind[MQ_N] = {0, 1, 2, 2, 3, 3, 0, 1, 2, 0, 1, 0, 1, 0, 1};
  T v[MQ_N][NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int idx = tid + TX * TYT * e;
    const int tj = idx / MV_TW, ti = idx - tj * MV_TW;
    const int ii = i0 + ti, jj = j0 + tj;
#pragma unroll
    for (int q = 0; q < MQ_N; ++q) {
      T x = T(0);
      if (src[q] && idx < NTILE && kind_has(g, kind[q], ii, jj)) x = src[q][kind_at(g, kind[q], ii, jj)];
      v[q][e] = x;
    }
  }
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int idx = tid + TX * TYT * e;
    const int tj = idx / MV_TW, ti = idx - tj * MV_TW;
    if (idx < NTILE) {
#pragma unroll
      for (int q = 0; q < MQ_N; ++q) s[q][tj][ti] = v[q][e];
    }
  }
  __syncthreads();
  matvec_point(g, a, MvTile<T, TYT>{s, i0, j0}, ib + (int)threadIdx.x, jb + (int)threadIdx.y);
}

// ---------------------------------------------------------------------------------
// coupled dots (src/mod_bicgstab2.f90:32-72), up to 5 pairs in one pass over the
// owned sets (u: j=1..ny, i=1..i_u_max; v: j=1..j_v_max, i=1..nx; p: interior), each
// term (a*b)*vol in FP64, summed in the Sink's order; the fold runs the Krylov phase.
template <typename T>
struct DotArgs {
  const T* a[5];
  const T* b[5];
  int np;
  int phase, k, max_iter;
};

template <typename T, int NP>
__global__ void __launch_bounds__(NT) k_cdot(Grid g, DotArgs<T> d, Sink out, const int* __restrict__ fl) {
  if (fl[F_BRK]) return;
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  const int i_u_max = g.per_x ? nx : nx - 1;
  const int j_v_max = g.per_y ? ny : ny - 1;
  const double vol = g.vol;
  // term positions in the reference's loop: u faces, then v faces, then cells (j outer)
  const long nU = (long)ny * i_u_max, nV = (long)j_v_max * nx;
  DD v[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) v[q] = dd_from(0.0);
  const int i = blockIdx.x * TX + (int)threadIdx.x - nh;
  const int tiles = (g.rw + TY - 1) / TY;
  for (int t = blockIdx.y; t < tiles; t += gridDim.y) {
    const int j = t * TY + (int)threadIdx.y - nh;
    if (j >= 1 && j <= ny && i >= 1 && i <= i_u_max) {
      const size_t o = IU(g, i, j);
      const long pos = (long)(j - 1) * i_u_max + (i - 1);
#pragma unroll
      for (int q = 0; q < NP; ++q) sink_add<NP>(out, v, q, pos, (double)d.a[q][o] * (double)d.b[q][o] * vol);
    }
    if (j >= 1 && j <= j_v_max && i >= 1 && i <= nx) {
      const size_t o = g.nu + IV(g, i, j);
      const long pos = nU + (long)(j - 1) * nx + (i - 1);
#pragma unroll
      for (int q = 0; q < NP; ++q) sink_add<NP>(out, v, q, pos, (double)d.a[q][o] * (double)d.b[q][o] * vol);
    }
    if (j >= 1 && j <= ny && i >= 1 && i <= nx) {
      const size_t o = g.nu + g.nv + IP(g, i, j);
      const long pos = nU + nV + (long)(j - 1) * nx + (i - 1);
#pragma unroll
      for (int q = 0; q < NP; ++q) sink_add<NP>(out, v, q, pos, (double)d.a[q][o] * (double)d.b[q][o] * vol);
    }
  }
  sink_publish<NP>(out, v);
}

template <int NP>
__global__ void __launch_bounds__(NT) k_fold_krylov(int nb, int phase, int k, int max_iter, Sink in,
                                                    double* __restrict__ sc, int* __restrict__ fl) {
  if (fl[F_BRK]) return;
  double tot[NP];
  fold_sink<NP>(in, nb, tot);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    double t5[5] = {0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < NP; ++q) t5[q] = tot[q];
    krylov_phase(phase, k, max_iter, sc, fl, t5);
  }
}

// One coupled dot over separate padded arrays (a: u1, v1, p1; b: u2, v2, p2) -- the same owned
// sets, term positions and arithmetic as k_cdot for arrays that are not one triplet allocation.
struct DotSplit {
  const double* a[3];
  const double* b[3];
};
__global__ void __launch_bounds__(NT) k_cdot_split(Grid g, DotSplit d, Sink out) {
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  const int i_u_max = g.per_x ? nx : nx - 1;
  const int j_v_max = g.per_y ? ny : ny - 1;
  const double vol = g.vol;
  const long nU = (long)ny * i_u_max, nV = (long)j_v_max * nx;
  DD v[1] = {dd_from(0.0)};
  const int i = blockIdx.x * TX + (int)threadIdx.x - nh;
  const int tiles = (g.rw + TY - 1) / TY;
  for (int t = blockIdx.y; t < tiles; t += gridDim.y) {
    const int j = t * TY + (int)threadIdx.y - nh;
    if (j >= 1 && j <= ny && i >= 1 && i <= i_u_max) {
      const size_t o = IU(g, i, j);
      sink_add<1>(out, v, 0, (long)(j - 1) * i_u_max + (i - 1), d.a[0][o] * d.b[0][o] * vol);
    }
    if (j >= 1 && j <= j_v_max && i >= 1 && i <= nx) {
      const size_t o = IV(g, i, j);
      sink_add<1>(out, v, 0, nU + (long)(j - 1) * nx + (i - 1), d.a[1][o] * d.b[1][o] * vol);
    }
    if (j >= 1 && j <= ny && i >= 1 && i <= nx) {
      const size_t o = IP(g, i, j);
      sink_add<1>(out, v, 0, nU + nV + (long)(j - 1) * nx + (i - 1), d.a[2][o] * d.b[2][o] * vol);
    }
  }
  sink_publish<1>(out, v);
}
// Fold of k_cdot_split into *out (device memory).
__global__ void __launch_bounds__(NT) k_fold_dot_out(int nb, Sink in, double* __restrict__ out) {
  double tot[1];
  fold_sink<1>(in, nb, tot);
  if (threadIdx.x == 0 && threadIdx.y == 0) out[0] = tot[0];
}

// Start of a reference Krylov iteration: the device-side iteration number (F_KITER) the phases read.
__global__ void k_iter_begin(int* __restrict__ fl) {
  if (threadIdx.x == 0) fl[F_KITER] += 1;
}

// Whole-array triplet updates (Fortran whole-array expressions, ghosts included), each
// with the reference's association order.  Coefficients are read from device slots;
// `sgn` flips a coefficient (for x - c*y written as x + (-c)*y we keep x - c*y).
enum LinOp : int {
  L_SUB1 = 0,   // o = x0 - c0*x1
  L_ADD1,       // o = x0 + c0*x1
  L_ADD3,       // o = ((x0 + c0*x1) + c1*x2) + c2*x3
  L_SUB2,       // o = (x0 - c0*x1) - c1*x2
  L_ADD2,       // o = (x0 + c0*x1) + c1*x2
  L_PUPD,       // o = x0 + c0*(x1 - c1*x2)
  L_DIFF,       // o = x0 - x1
  L_COPY,       // o = x0
  L_SCALE,      // o = c0 * x0          (c0 = sc[i0] * s0)
  L_ADDS        // o = x0 + s0 * x1     (s0 literal)
};

template <typename T>
struct LinArgs {
  T* out;
  const T *x0, *x1, *x2, *x3;
  int c0, c1, c2;        // slot indices (or -1)
  double s0;             // literal factor
  int op;
  size_t n;
  int skip_on;           // 1: skip when F_BRK set
};

template <typename To, typename T>
__global__ void __launch_bounds__(256) k_lin(LinArgs<T> a, const double* __restrict__ sc,
                                             const int* __restrict__ fl) {
  if (a.skip_on && fl[F_BRK]) return;
  const T c0 = a.c0 >= 0 ? (T)sc[a.c0] : T(0);
  const T c1 = a.c1 >= 0 ? (T)sc[a.c1] : T(0);
  const T c2 = a.c2 >= 0 ? (T)sc[a.c2] : T(0);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.n; k += stride) {
    T o;
    switch (a.op) {
      case L_SUB1: o = a.x0[k] - c0 * a.x1[k]; break;
      case L_ADD1: o = a.x0[k] + c0 * a.x1[k]; break;
      case L_ADD3: o = ((a.x0[k] + c0 * a.x1[k]) + c1 * a.x2[k]) + c2 * a.x3[k]; break;
      case L_SUB2: o = (a.x0[k] - c0 * a.x1[k]) - c1 * a.x2[k]; break;
      case L_ADD2: o = (a.x0[k] + c0 * a.x1[k]) + c1 * a.x2[k]; break;
      case L_PUPD: o = a.x0[k] + c0 * (a.x1[k] - c1 * a.x2[k]); break;
      case L_DIFF: o = a.x0[k] - a.x1[k]; break;
      case L_COPY: o = a.x0[k]; break;
      case L_SCALE: o = (T)(sc[a.c0] * a.s0) * a.x0[k]; break;
      default: o = a.x0[k] + (T)a.s0 * a.x1[k]; break;
    }
    a.out[k] = o;
  }
}

// Multi-vector update o = x0 + sgn * sum_q c_q x_q (coefficients sc[S_HBUF + c_off + q]);
// o, x0 stored as O, the x_q as V (FP64 or FP32 triplets), arithmetic in FP64; o may alias
// x0.  One pass reads every operand once.
constexpr int MAXPY_MAX = 48;
// ---------------------------------------------------------------------------------
// Outer FGMRES scalar work on the device (one thread): the Hessenberg column from the
// two classical Gram-Schmidt passes (slots S_HBUF+i and S_HBUF+64+i) and ||w||^2
// (S_DOT0), Givens rotations, residual estimate and the 1/h scaling of the next basis
// vector -- no host round trip per step.  fg layout (ld = M+1): H[0 .. (M+1)M), then
// cs[M], sn[M], g[M+1], kk.
constexpr int FG_M = 48;
constexpr int FG_CS = (FG_M + 1) * FG_M, FG_SN = FG_CS + FG_M, FG_G = FG_SN + FG_M, FG_KK = FG_G + FG_M + 1;
constexpr int FG_SIZE = FG_KK + 8;

// Called by every thread of a workgroup (uniformly): the new column's entries and the previous
// rotations are staged in LDS by all threads, then one thread applies the rotations with the
// running entry carried in a register (unrolled, so the LDS reads of several steps are in flight
// together; the dependent chain is arithmetic only) -- the same operations in the same order as
// rotating the column in place.
// hflag (optional): a host-mapped int that receives the cycle's convergence flag after this
// step (the host reads it one step behind, after an event; written with a system-scope fence).
__device__ inline void fgmres_publish(int* hflag, int done) {
  if (!hflag) return;
  *(volatile int*)hflag = done;
  __threadfence_system();
}
__device__ inline void fgmres_step_blk(int k, double* __restrict__ sc, double* __restrict__ fg, int* __restrict__ fl,
                                       int* hflag) {
  __shared__ double hcol[FG_M + 1], rc[FG_M], rs[FG_M];
  if (fl[F_FG_DONE]) {
    if (threadIdx.x == 0) fgmres_publish(hflag, 1);
    return;
  }
  for (int i = threadIdx.x; i <= k; i += blockDim.x) {
    hcol[i] = sc[S_HBUF + i] + sc[S_HBUF + 64 + i];
    if (i < k) {
      rc[i] = fg[FG_CS + i];
      rs[i] = fg[FG_SN + i];
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double tol_abs = sc[S_FG_TOL];
  double* H = fg + (size_t)k * (FG_M + 1);
  const double hn = sqrt(sc[S_DOT0]);
  double cur = hcol[0];   // H[i] as rotated so far
#pragma unroll 8
  for (int i = 0; i < k; ++i) {
    const double nxt = hcol[i + 1];
    const double c = rc[i], s = rs[i];
    const double t = c * cur + s * nxt;
    cur = -s * cur + c * nxt;
    H[i] = t;
  }
  const double den = hypot(cur, hn);
  const double c = den > 0.0 ? cur / den : 1.0, s = den > 0.0 ? hn / den : 0.0;
  fg[FG_CS + k] = c;
  fg[FG_SN + k] = s;
  H[k] = den;
  H[k + 1] = 0.0;
  fg[FG_G + k + 1] = -s * fg[FG_G + k];
  fg[FG_G + k] = c * fg[FG_G + k];
  fg[FG_KK] = (double)(k + 1);
  sc[S_INVH] = hn > 0.0 ? 1.0 / hn : 0.0;
  const int done = (fabs(fg[FG_G + k + 1]) <= tol_abs || hn == 0.0) ? 1 : 0;
  if (done) fl[F_FG_DONE] = 1;
  fgmres_publish(hflag, done);
}

template <typename O, typename V>
struct MaxpyArgs {
  O* out;
  const O* x0;
  const V* x[MAXPY_MAX];
  int nq, c_off;
  double sgn;
  size_t n;
};
// With `partials` set, also publishes per-workgroup sums of out^2 (of the stored values; grid
// capped at RED_MAXB).
template <typename O, typename V>
__global__ void __launch_bounds__(256) k_maxpy(MaxpyArgs<O, V> a, const double* __restrict__ sc,
                                               double* __restrict__ partials, const int* __restrict__ fl, int flag) {
  if (flag >= 0 && !fl[flag]) return;   // predicated pass (second Gram-Schmidt pass)
  __shared__ double c[MAXPY_MAX];
  if ((int)threadIdx.x < a.nq) c[threadIdx.x] = a.sgn * sc[S_HBUF + a.c_off + threadIdx.x];
  __syncthreads();
  double nrm[1] = {0.0};
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.n; k += stride) {
    double o = (double)a.x0[k];
#pragma unroll 4
    for (int q = 0; q < a.nq; ++q) o += c[q] * (double)a.x[q][k];
    const O os = (O)o;
    a.out[k] = os;
    nrm[0] += (double)os * (double)os;
  }
  if (partials) {
    __shared__ double lds[16][1];
    block_sum<1>(nrm, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = nrm[0];
  }
}

// First Gram-Schmidt pass fused with the second pass's products: o = x0 - sum_q c_q x_q (as
// k_maxpy, same operation order) and, in the same sweep, the per-workgroup sums of o x_q
// (q < nq) and o^2 -- the inner products the re-orthogonalisation pass needs, read while the
// basis is streamed once instead of twice.  NQ >= nq (compile-time accumulator count); the
// caller points the slots nq..NQ-1 at x0 (valid, reread from cache) and their coefficients are
// zero, so the loop has no branches and every load of an element is issued at once.
template <typename V, int NQ>
__global__ void __launch_bounds__(256) k_maxpy_dot(MaxpyArgs<V, V> a, const double* __restrict__ sc,
                                                   double* __restrict__ partials) {
  __shared__ double c[MAXPY_MAX];
  if ((int)threadIdx.x < NQ) c[threadIdx.x] = (int)threadIdx.x < a.nq ? a.sgn * sc[S_HBUF + a.c_off + threadIdx.x] : 0.0;
  __syncthreads();
  double acc[NQ + 1];
#pragma unroll
  for (int q = 0; q <= NQ; ++q) acc[q] = 0.0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.n; k += stride) {
    double x[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) x[q] = (double)a.x[q][k];
    double o = (double)a.x0[k];
#pragma unroll
    for (int q = 0; q < NQ; ++q) o += c[q] * x[q];
    const V os = (V)o;
    a.out[k] = os;
    const double od = (double)os;
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] += od * x[q];
    acc[NQ] += od * od;
  }
  __shared__ double lds[16][NQ + 1];
  block_sum<NQ + 1>(acc, lds);
  if (threadIdx.x == 0) {   // column-major (value q of every workgroup contiguous): coalesced folds
#pragma unroll
    for (int q = 0; q <= NQ; ++q) partials[(size_t)q * gridDim.x + blockIdx.x] = acc[q];
  }
}
// Fold of k_maxpy_dot: sc[S_DOT0] = vol ||o||^2, fl[F_REORTH] = (||o||^2 < frac ||w||^2, the
// Kahan-Parlett test against sc[S_DOT1] = vol ||w||^2, frac = 1/2 its usual threshold), and the
// second pass's coefficients
// sc[S_HBUF + 64 + q] = vol <o, x_q> when it runs, zero otherwise.  fg_k >= 0: when the second
// pass does not run, the FGMRES step k (fgmres_step_blk) follows here, in the same launch.
__global__ void __launch_bounds__(1024) k_fold_maxpy_dot(int nb, int nq, int stride, double vol, double frac,
                                                         const double* __restrict__ partials,
                                                         double* __restrict__ sc, int* __restrict__ fl, int fg_k,
                                                         double* __restrict__ fg, int* hflag) {
  __shared__ double red[MAXPY_MAX + 1];
  __shared__ int reorth;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q = wid; q <= nq; q += (int)(blockDim.x >> 6)) {
    const int col = q < nq ? q : stride - 1;
    // nb <= RED_MAXB: every load of the lane issued before the first add (same order of adds)
    double v[RED_MAXB / 64];
#pragma unroll
    for (int r = 0; r < RED_MAXB / 64; ++r) v[r] = (lane + 64 * r < nb) ? partials[(size_t)col * nb + lane + 64 * r] : 0.0;
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < RED_MAXB / 64; ++r)
      if (lane + 64 * r < nb) s += v[r];
    s = wave_sum(s);
    if (lane == 0) red[q] = s * vol;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    sc[S_DOT0] = red[nq];
    reorth = (red[nq] < frac * sc[S_DOT1]) ? 1 : 0;
    fl[F_REORTH] = reorth;
  }
  __syncthreads();
  if ((int)threadIdx.x < nq) sc[S_HBUF + 64 + threadIdx.x] = reorth ? red[threadIdx.x] : 0.0;
  if (fg_k >= 0 && !reorth) {
    __syncthreads();
    fgmres_step_blk(fg_k, sc, fg, fl, hflag);
  }
}

// Multi-dot of the FGMRES Gram-Schmidt: <w, v_q> for q < nv in one launch (FP64 or FP32
// storage, FP64 products and sums).  blockIdx.y selects a chunk of MD_CH basis vectors,
// blockIdx.x one of MD_BX slices of the flat arrays; plain FP64 sums in a fixed order.  The triplets the outer iteration builds
// (residuals B - A X and A z) are exactly zero outside the coupled_dot owned sets (the
// matvec zeroes ghosts and the wall-normal faces carry the zero BC value), so the flat
// product times vol is the coupled_dot inner product.
constexpr int MD_CH = 8;
constexpr int MD_BX = 256;
template <typename V>
struct MdotArgs {
  const V* w;
  const V* v[MAXPY_MAX];
  int nv;
  size_t n;
};
// Chunk 0 also sums <w, w> (partial slot MD_CH).  flag >= 0: run only when fl[flag] is set.
// 16-byte vectors of V (two doubles or four floats) for the streaming Gram-Schmidt kernels.
template <typename V>
struct Vec16;
template <>
struct Vec16<double> {
  using T = double2;
  static constexpr int W = 2;
};
template <>
struct Vec16<float> {
  using T = float4;
  static constexpr int W = 4;
};
template <typename V>
__device__ inline V vcomp(const typename Vec16<V>::T& v, int c) {
  return reinterpret_cast<const V*>(&v)[c];
}

// vec: every array is 16-byte aligned; the flat range is then read as 16-byte vectors (a lane
// sums its W elements in order) and the last n mod W elements one by one.
template <typename V>
__global__ void __launch_bounds__(256) k_mdot(MdotArgs<V> a, double* __restrict__ partials, const int* __restrict__ fl,
                                              int flag, int vec) {
  if (flag >= 0 && !fl[flag]) return;
  const int c0 = (int)blockIdx.y * MD_CH;
  const V* vp[MD_CH];
#pragma unroll
  for (int q = 0; q < MD_CH; ++q) vp[q] = (c0 + q < a.nv) ? a.v[c0 + q] : a.w;   // padding lanes reread w
  double acc[MD_CH + 1];
#pragma unroll
  for (int q = 0; q <= MD_CH; ++q) acc[q] = 0.0;
  const size_t stride = (size_t)MD_BX * blockDim.x;
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t k0 = gid;
  if (vec) {
    using VT = typename Vec16<V>::T;
    constexpr int W = Vec16<V>::W;
    const size_t nw = a.n / W;
    for (size_t k = gid; k < nw; k += stride) {
      const VT wv = reinterpret_cast<const VT*>(a.w)[k];
      VT xv[MD_CH];
#pragma unroll
      for (int q = 0; q < MD_CH; ++q) xv[q] = reinterpret_cast<const VT*>(vp[q])[k];
#pragma unroll
      for (int e = 0; e < W; ++e) {
        const double wk = (double)vcomp<V>(wv, e);
#pragma unroll
        for (int q = 0; q < MD_CH; ++q) acc[q] += wk * (double)vcomp<V>(xv[q], e);
        acc[MD_CH] += wk * wk;
      }
    }
    k0 = nw * W + gid;
  }
  for (size_t k = k0; k < a.n; k += stride) {
    const double wk = (double)a.w[k];
    double x[MD_CH];
#pragma unroll
    for (int q = 0; q < MD_CH; ++q) x[q] = (double)vp[q][k];
#pragma unroll
    for (int q = 0; q < MD_CH; ++q) acc[q] += wk * x[q];
    acc[MD_CH] += wk * wk;
  }
  __shared__ double lds[16][MD_CH + 1];
  block_sum<MD_CH + 1>(acc, lds);
  if (threadIdx.x == 0) {   // column-major per chunk: coalesced folds
#pragma unroll
    for (int q = 0; q <= MD_CH; ++q)
      partials[((size_t)blockIdx.y * (MD_CH + 1) + q) * MD_BX + blockIdx.x] = acc[q];
  }
}
// Fold of k_mdot: one wavefront per basis vector; sc[S_HBUF + off + q] = vol * sum, and
// sc[S_DOT1] = vol * <w, w>.  flag >= 0: when fl[flag] is clear the pass was skipped and the
// coefficients are written as zero.
__global__ void __launch_bounds__(1024) k_mdot_fold(int nv, int off, double vol, const double* __restrict__ partials,
                                                    double* __restrict__ sc, const int* __restrict__ fl, int flag) {
  const bool skip = flag >= 0 && !fl[flag];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int q = wid; q <= nv; q += (int)(blockDim.x >> 6)) {
    const int ch = q < nv ? q / MD_CH : 0, qq = q < nv ? q % MD_CH : MD_CH;
    double s = 0.0;
    if (!skip) {
      double v[MD_BX / 64];   // every load of the lane issued before the first add
#pragma unroll
      for (int r = 0; r < MD_BX / 64; ++r) v[r] = partials[((size_t)ch * (MD_CH + 1) + qq) * MD_BX + lane + 64 * r];
#pragma unroll
      for (int r = 0; r < MD_BX / 64; ++r) s += v[r];
    }
    s = wave_sum(s);
    if (lane == 0) {
      if (q < nv) sc[S_HBUF + off + q] = s * vol;
      else if (!skip) sc[S_DOT1] = s * vol;
    }
  }
}
// Norm fold after a Gram-Schmidt pass: sc[S_DOT0] = vol * sum.  mode 1 (after the first
// pass): fl[F_REORTH] = (||w'||^2 < ||w||^2 / 2), the Kahan-Parlett "twice is enough" test
// with its usual 1/sqrt(2) threshold.  mode 2 (after the second pass): only when it ran, and
// then (fg_k >= 0) the FGMRES step k follows in the same launch.
__global__ void __launch_bounds__(NT) k_fold_gs_norm(int nb, int mode, double vol, const double* __restrict__ partials,
                                                     double* __restrict__ sc, int* __restrict__ fl, int fg_k,
                                                     double* __restrict__ fg, int* hflag) {
  if (mode == 2 && !fl[F_REORTH]) return;
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v[0] += partials[b];
  __shared__ double lds[16][1];
  block_sum<1>(v, lds);
  if (threadIdx.x == 0) {
    sc[S_DOT0] = vol * v[0];
    if (mode == 1) fl[F_REORTH] = (vol * v[0] < 0.5 * sc[S_DOT1]) ? 1 : 0;
  }
  if (mode == 2 && fg_k >= 0) {
    __syncthreads();
    fgmres_step_blk(fg_k, sc, fg, fl, hflag);
  }
}

// Fold of per-workgroup plain sums: sc[slot] = scale * sum.
__global__ void __launch_bounds__(NT) k_fold_plain(int nb, int slot, double scale, const double* __restrict__ partials,
                                                   double* __restrict__ sc) {
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v[0] += partials[b];
  __shared__ double lds[16][1];
  block_sum<1>(v, lds);
  if (threadIdx.x == 0) sc[slot] = scale * v[0];
}

// ---------------------------------------------------------------------------------
// Outer FGMRES cycle start and end (one thread each); the steps run inside the Gram-Schmidt
// folds above (fgmres_step_blk).
__global__ void k_fgmres_init(double* __restrict__ sc, double* __restrict__ fg, int* __restrict__ fl) {
  const double beta = sqrt(sc[S_DOT0]);
  for (int i = 0; i <= FG_M; ++i) fg[FG_G + i] = 0.0;
  fg[FG_G] = beta;
  fg[FG_KK] = 0.0;
  sc[S_INVH] = beta > 0.0 ? 1.0 / beta : 0.0;
  fl[F_FG_DONE] = beta > 0.0 ? 0 : 1;
}


// y = H^-1 g (back substitution) into slots S_HBUF + 128 + i; returns kk in S_HBUF + 127.
__global__ void k_fgmres_solve(double* __restrict__ sc, const double* __restrict__ fg) {
  const int kk = (int)fg[FG_KK];
  double y[FG_M];
  for (int i = kk - 1; i >= 0; --i) {
    double s = fg[FG_G + i];
    for (int j = i + 1; j < kk; ++j) s -= fg[(size_t)j * (FG_M + 1) + i] * y[j];
    const double d = fg[(size_t)i * (FG_M + 1) + i];
    y[i] = d != 0.0 ? s / d : 0.0;
  }
  for (int i = 0; i < kk; ++i) sc[S_HBUF + 128 + i] = y[i];
  sc[S_HBUF + 127] = (double)kk;
}

// Precision conversion / scaled copy: out[k] = (To)(s * in[k]), s = sc[slot] * lit.
template <typename To, typename Ti>
__global__ void __launch_bounds__(256) k_convert(To* __restrict__ out, const Ti* __restrict__ in,
                                                 size_t n, const double* __restrict__ sc, int slot,
                                                 double lit, int invert) {
  double s = lit;
  if (slot >= 0) s *= invert ? 1.0 / sc[slot] : sc[slot];
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    out[k] = (To)(s == 1.0 ? (double)in[k] : s * (double)in[k]);
}

// X += s * D (FP64 refinement update), s = sc[slot].
template <typename Td>
__global__ void __launch_bounds__(256) k_refine_add(double* __restrict__ X, const Td* __restrict__ D,
                                                    size_t n, const double* __restrict__ sc, int slot) {
  const double s = sc[slot];
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    X[k] = X[k] + s * (double)D[k];
}

// ---------------------------------------------------------------------------------
// Jacobi-PCG pressure solve (src/mod_subsolvers.f90:32-246), two fused passes per
// iteration with device-resident scalars and predication on F_P_DONE.

// Face coefficients c = inv_h2 / rho_face with zeroed wall faces (:65-78), and D^-1 (:61-87).
template <typename T>
__global__ void k_pcg_coef(Grid g, const T* __restrict__ rho_u, const T* __restrict__ rho_v,
                           T* __restrict__ cx, T* __restrict__ cy, T* __restrict__ dinv) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  auto cxv = [&](int ii, int jj) -> double {
    if (!g.per_x && (ii == 0 || ii == nx)) return 0.0;
    return g.inv_dx2 / (double)rho_u[IU(g, ii, jj)];
  };
  auto cyv = [&](int ii, int jj) -> double {
    if (!g.per_y && (jj == 0 || jj == ny)) return 0.0;
    return g.inv_dy2 / (double)rho_v[IV(g, ii, jj)];
  };
  if (i >= -nh && i <= nx + nh && j >= 1 - nh && j <= ny + nh)
    cx[IU(g, i, j)] = (j >= 1 && j <= ny && i >= 0 && i <= nx) ? (T)cxv(i, j) : T(0);
  if (i >= 1 - nh && i <= nx + nh && j >= -nh && j <= ny + nh)
    cy[IV(g, i, j)] = (i >= 1 && i <= nx && j >= 0 && j <= ny) ? (T)cyv(i, j) : T(0);
  if (i >= 1 - nh && i <= nx + nh && j >= 1 - nh && j <= ny + nh) {
    double dv = 1.0;
    if (i >= 1 && i <= nx && j >= 1 && j <= ny) {
      const double d = cxv(i, j) + cxv(i - 1, j) + cyv(i, j) + cyv(i, j - 1);
      dv = (d > 1.0e-14) ? 1.0 / d : 1.0;
    }
    dinv[IP(g, i, j)] = (T)dv;
  }
}

// Neighbour offsets of an interior cell with periodic wrap; at walls the neighbour is
// the cell itself (the mirrored ghost), always multiplied by a zero coefficient.
struct Nbr { size_t c, e, w, n, s; };
__device__ inline Nbr cell_nbrs(const Grid& g, int i, int j) {
  const int nx = g.nx, ny = g.ny;
  int ie = i + 1, iw = i - 1, jn = j + 1, js = j - 1;
  if (ie > nx) ie = g.per_x ? 1 : nx;
  if (iw < 1) iw = g.per_x ? nx : 1;
  if (jn > ny) jn = g.per_y ? 1 : ny;
  if (js < 1) js = g.per_y ? ny : 1;
  Nbr r;
  r.c = IP(g, i, j); r.e = IP(g, ie, j); r.w = IP(g, iw, j); r.n = IP(g, i, jn); r.s = IP(g, i, js);
  return r;
}

// Host-set scalars without host memory or a stream synchronisation: the value travels as a
// kernel argument.  sc[slot] = v.
__global__ void k_set_slot(double* __restrict__ sc, int slot, double v) {
  if (threadIdx.x == 0) sc[slot] = v;
}
// The flags cleared (F_RELG = relg) and sc[S_GSCALE] = 1, as before every Krylov solve.
__global__ void k_reset_flags(int* __restrict__ fl, int relg, double* __restrict__ sc) {
  for (int i = (int)threadIdx.x; i < (int)F_COUNT; i += (int)blockDim.x) fl[i] = (i == (int)F_RELG) ? relg : 0;
  if (threadIdx.x == 0) sc[S_GSCALE] = 1.0;
}

// Sum of rhs over the interior (for the zero-mean projection, :91-99) -> S_P_MEAN.
// Interior sum of a cell field (zero-mean projection, :91-99) -> sc[slot] = sum / (nx ny).
// skip_mode: 0 never, 1 when a Krylov exit is pending (F_BRK), 2 when the PCG did not
// iterate (zero-rhs early exit or skipped: F_P_DONE >= 2).
__device__ inline bool pcg_sum_skip(int skip_mode, const int* fl) {
  return (skip_mode == 1 && fl[F_BRK]) || (skip_mode == 2 && fl[F_P_DONE] >= 2);
}
// Cell sums: term position (j-1) nx + (i-1), the reference's j-outer / i-inner loop.
__device__ inline long cell_pos(const Grid& g, int i, int j) { return (long)(j - 1) * g.nx + (i - 1); }

template <typename T>
__global__ void __launch_bounds__(NT) k_pcg_sum(Grid g, const T* __restrict__ a, Sink out,
                                                int skip_mode, const int* __restrict__ fl) {
  if (pcg_sum_skip(skip_mode, fl)) return;
  DD v[1] = {dd_from(0.0)};
  const int i = blockIdx.x * TX + (int)threadIdx.x + 1;
  const int tiles = (g.ny + TY - 1) / TY;
  for (int t = blockIdx.y; t < tiles; t += gridDim.y) {
    const int j = t * TY + (int)threadIdx.y + 1;
    if (i <= g.nx && j <= g.ny) sink_add<1>(out, v, 0, cell_pos(g, i, j), (double)a[IP(g, i, j)]);
  }
  sink_publish<1>(out, v);
}
__global__ void __launch_bounds__(NT) k_fold_sum(Grid g, int nb, int slot, int skip_mode, Sink in,
                                                 double* __restrict__ sc, const int* __restrict__ fl) {
  if (pcg_sum_skip(skip_mode, fl)) return;
  double tot[1];
  fold_sink<1>(in, nb, tot);
  if (threadIdx.x == 0 && threadIdx.y == 0) sc[slot] = tot[0] / (double)(g.nx * g.ny);
}

// PCG start: rhs_work = rhs - mean (walls) (:90-105), ||rhs_work|| with early exit
// (:107-120), r = rhs_work - A phi, z = D^-1 r, p = z, (r,z) (:126-157).
// phi_zero: the caller guarantees phi == 0 (the preconditioner's call), so A phi = 0
// and phi's interior is written as 0 here.
template <typename T>
__global__ void __launch_bounds__(NT) k_pcg_init(Grid g, const T* __restrict__ rhs, T* __restrict__ phi,
                                                 const T* __restrict__ cx, const T* __restrict__ cy,
                                                 const T* __restrict__ dinv, T* __restrict__ r,
                                                 T* __restrict__ z, T* __restrict__ p, int walls_both,
                                                 int phi_zero, Sink out,
                                                 const double* __restrict__ sc, const int* __restrict__ fl) {
  if (fl[F_BRK]) return;
  DD v[2] = {dd_from(0.0), dd_from(0.0)};
  const T mean = walls_both ? (T)sc[S_P_MEAN] : T(0);
  const int i = blockIdx.x * TX + (int)threadIdx.x + 1;
  const int tiles = (g.ny + TY - 1) / TY;
  for (int t = blockIdx.y; t < tiles; t += gridDim.y) {
    const int j = t * TY + (int)threadIdx.y + 1;
    if (i <= g.nx && j <= g.ny) {
      const Nbr nb = cell_nbrs(g, i, j);
      const T rw = walls_both ? rhs[nb.c] - mean : rhs[nb.c];
      T ap = T(0);
      if (!phi_zero) {
        const T cE = cx[IU(g, i, j)], cW = cx[IU(g, i - 1, j)];
        const T cN = cy[IV(g, i, j)], cS = cy[IV(g, i, j - 1)];
        ap = (cE + cW + cN + cS) * phi[nb.c] -
             (cE * phi[nb.e] + cW * phi[nb.w] + cN * phi[nb.n] + cS * phi[nb.s]);
      } else {
        phi[nb.c] = T(0);
      }
      const T rr = rw - ap;
      const T zz = dinv[nb.c] * rr;
      r[nb.c] = rr; z[nb.c] = zz; p[nb.c] = zz;
      const long pos = cell_pos(g, i, j);
      sink_add<2>(out, v, 0, pos, (double)rw * (double)rw);
      sink_add<2>(out, v, 1, pos, (double)rr * (double)zz);
    }
  }
  sink_publish<2>(out, v);
}
__global__ void __launch_bounds__(NT) k_fold_pcg_init(int nb, Sink in, double* __restrict__ sc,
                                                      int* __restrict__ fl) {
  if (fl[F_BRK]) {   // pending Krylov exit: mark the solve as skipped
    if (threadIdx.x == 0 && threadIdx.y == 0) fl[F_P_DONE] = 3;
    return;
  }
  double tot[2];
  fold_sink<2>(in, nb, tot);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const double rn = sqrt(tot[0]);
    sc[S_P_RHSN] = rn;
    sc[S_P_RES] = 0.0;
    fl[F_P_ITER] = 0;
    if (rn < 1.0e-14) fl[F_P_DONE] = 2;
    else { fl[F_P_DONE] = 0; sc[S_P_RZ] = tot[1]; }
  }
}

// Iteration pass 1: p = z + beta p_old (iteration > 1, computed at the 5 stencil points),
// Ap = A p, (p, Ap) -> alpha  (:161-192, :220-224).
template <typename T>
__global__ void __launch_bounds__(NT) k_pcg_iter1(Grid g, int it, const T* __restrict__ z,
                                                  const T* __restrict__ p_old, T* __restrict__ p_new,
                                                  T* __restrict__ Ap, const T* __restrict__ cx,
                                                  const T* __restrict__ cy, Sink out,
                                                  const double* __restrict__ sc, const int* __restrict__ fl) {
  if (fl[F_P_DONE]) return;
  DD v[1] = {dd_from(0.0)};
  const T beta = (T)sc[S_P_BETA];
  const int i = blockIdx.x * TX + (int)threadIdx.x + 1;
  const int tiles = (g.ny + TY - 1) / TY;
  for (int t = blockIdx.y; t < tiles; t += gridDim.y) {
    const int j = t * TY + (int)threadIdx.y + 1;
    if (i <= g.nx && j <= g.ny) {
      const Nbr nb = cell_nbrs(g, i, j);
      T pc, pe, pw, pn, ps;
      if (it == 1) {
        pc = p_old[nb.c]; pe = p_old[nb.e]; pw = p_old[nb.w]; pn = p_old[nb.n]; ps = p_old[nb.s];
      } else {
        pc = z[nb.c] + beta * p_old[nb.c];
        pe = z[nb.e] + beta * p_old[nb.e];
        pw = z[nb.w] + beta * p_old[nb.w];
        pn = z[nb.n] + beta * p_old[nb.n];
        ps = z[nb.s] + beta * p_old[nb.s];
      }
      const T cE = cx[IU(g, i, j)], cW = cx[IU(g, i - 1, j)];
      const T cN = cy[IV(g, i, j)], cS = cy[IV(g, i, j - 1)];
      const T ap = (cE + cW + cN + cS) * pc - (cE * pe + cW * pw + cN * pn + cS * ps);
      p_new[nb.c] = pc;
      Ap[nb.c] = ap;
      sink_add<1>(out, v, 0, cell_pos(g, i, j), (double)pc * (double)ap);
    }
  }
  sink_publish<1>(out, v);
}
__global__ void __launch_bounds__(NT) k_fold_pcg_iter1(int nb, int it, Sink in, double* __restrict__ sc,
                                                       int* __restrict__ fl) {
  if (fl[F_P_DONE]) return;
  double tot[1];
  fold_sink<1>(in, nb, tot);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const double pAp = tot[0];
    sc[S_P_PAP] = pAp;
    if (fabs(pAp) < 1.0e-30) { fl[F_P_DONE] = 1; fl[F_P_ITER] = it; }
    else sc[S_P_ALPHA] = sc[S_P_RZ] / pAp;
  }
}

// Iteration pass 2: phi += alpha p; r -= alpha Ap; z = D^-1 r; ||r||, (r, z) (:194-218).
template <typename T>
__global__ void __launch_bounds__(NT) k_pcg_iter2(Grid g, T* __restrict__ phi, const T* __restrict__ p,
                                                  T* __restrict__ r, const T* __restrict__ Ap, T* __restrict__ z,
                                                  const T* __restrict__ dinv, Sink out,
                                                  const double* __restrict__ sc, const int* __restrict__ fl) {
  if (fl[F_P_DONE]) return;
  DD v[2] = {dd_from(0.0), dd_from(0.0)};
  const T alpha = (T)sc[S_P_ALPHA];
  const int i = blockIdx.x * TX + (int)threadIdx.x + 1;
  const int tiles = (g.ny + TY - 1) / TY;
  for (int t = blockIdx.y; t < tiles; t += gridDim.y) {
    const int j = t * TY + (int)threadIdx.y + 1;
    if (i <= g.nx && j <= g.ny) {
      const size_t c = IP(g, i, j);
      phi[c] = phi[c] + alpha * p[c];
      const T rr = r[c] - alpha * Ap[c];
      const T zz = dinv[c] * rr;
      r[c] = rr;
      z[c] = zz;
      const long pos = cell_pos(g, i, j);
      sink_add<2>(out, v, 0, pos, (double)rr * (double)rr);
      sink_add<2>(out, v, 1, pos, (double)rr * (double)zz);
    }
  }
  sink_publish<2>(out, v);
}
__global__ void __launch_bounds__(NT) k_fold_pcg_iter2(int nb, int it, int max_iter, Sink in,
                                                       double* __restrict__ sc, int* __restrict__ fl) {
  if (fl[F_P_DONE]) return;
  double tot[2];
  fold_sink<2>(in, nb, tot);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const double res = sqrt(tot[0]);
    sc[S_P_RES] = res;
    if (res / sc[S_P_RHSN] < sc[S_P_TOL]) { fl[F_P_DONE] = 1; fl[F_P_ITER] = it; }
    else {
      const double rz_new = tot[1];
      sc[S_P_BETA] = rz_new / sc[S_P_RZ];
      sc[S_P_RZ] = rz_new;
      if (it == max_iter) fl[F_P_ITER] = max_iter;
    }
  }
}

// Final zero-mean removal of phi (:227-241) -- fine-API path; the preconditioner fuses it.
template <typename T>
__global__ void k_pcg_submean(Grid g, T* __restrict__ phi, const double* __restrict__ sc,
                              const int* __restrict__ fl) {
  if (fl[F_P_DONE] == 2) return;
  int i, j;
  cell_ij(0, i, j);
  if (i <= g.nx && j <= g.ny) {
    const size_t c = IP(g, i, j);
    phi[c] = phi[c] - (T)sc[S_P_MEAN];
  }
}

template <typename T>
__global__ void k_zero_if(T* __restrict__ a, size_t n, const int* __restrict__ fl, int flag, int value) {
  if (fl[flag] != value) return;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) a[k] = T(0);
}

// ---------------------------------------------------------------------------------
// PCD preconditioner pieces (src/mod_pcd_precond.f90).

// r_p = -((1/dt) phi + u.grad phi) on cells, zero ghosts (:54-62).  phi is the PCG
// iterate before mean removal: the mean is subtracted here on the interior; wall
// ghosts of phi are 0; periodic ghosts wrap.  A zero-rhs early exit means phi = 0.
// visc > 0 (chosen fast forms only): the Schur approximation also carries the viscous
// term, r_p -= visc * kappa * z_p (Cahouet-Chabard type), for viscous-dominated regions.
// phic (chosen fast forms): phi read from the multigrid's compact level-0 array instead.
template <typename T>
__global__ void __launch_bounds__(NT) k_pcd_rp(Grid g, const T* __restrict__ phi, const T* __restrict__ uadv,
                                               const T* __restrict__ vadv, T dt, int walls_both,
                                               T* __restrict__ rp, const double* __restrict__ sc,
                                               const int* __restrict__ fl, const T* __restrict__ zp,
                                               const T* __restrict__ kap, T visc, const T* __restrict__ phic) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  if (i < 1 - nh || i > nx + nh || j < 1 - nh || j > ny + nh) return;
  T out = T(0);
  if (i >= 1 && i <= nx && j >= 1 && j <= ny) {
    // the five phi values: neighbours wrapped (periodic) or outside (a wall: value 0); every
    // load is issued before any value is used (a use between loads waits for all of them)
    const bool early = (fl[F_P_DONE] == 2);
    const T mean = walls_both ? (T)sc[S_P_MEAN] : T(0);
    const bool okE = i < nx || g.per_x, okW = i > 1 || g.per_x;
    const bool okN = j < ny || g.per_y, okS = j > 1 || g.per_y;
    const int iE = i < nx ? i + 1 : 1, iW = i > 1 ? i - 1 : nx;
    const int jN = j < ny ? j + 1 : 1, jS = j > 1 ? j - 1 : ny;
    const T* src = phic ? phic : phi;
    auto at = [&](int ii, int jj) -> size_t { return phic ? (size_t)(ii - 1) + (size_t)(jj - 1) * nx : IP(g, ii, jj); };
    const T pC = src[at(i, j)], pE = src[at(iE, j)], pW = src[at(iW, j)], pN = src[at(i, jN)], pS = src[at(i, jS)];
    const T ua = uadv[IU(g, i, j)], ub = uadv[IU(g, i - 1, j)], va = vadv[IV(g, i, j)], vb = vadv[IV(g, i, j - 1)];
    auto ph = [&](T x, bool ok) -> T {
      if (early || !ok) return T(0);
      return walls_both ? x - mean : x;
    };
    const T half = T(0.5);
    const T u_avg = half * (ua + ub);
    const T v_avg = half * (va + vb);
    const T conv = u_avg * (ph(pE, okE) - ph(pW, okW)) * (half * (T)g.inv_dx) +
                   v_avg * (ph(pN, okN) - ph(pS, okS)) * (half * (T)g.inv_dy);
    out = -((T(1) / dt) * ph(pC, true) + conv);
    if (visc > T(0)) out = out - visc * kap[IP(g, i, j)] * zp[IP(g, i, j)];
  }
  rp[IP(g, i, j)] = out;
}

// z_u - d r_p/dx on faces i=1..nx-1 (copies at i=0, nx), z_v - d r_p/dy on j=1..ny-1
// (copies at j=0, ny), zero elsewhere (:66-84, single-rank loop bounds).  bv (optional): the
// v block's rows (i = 1..nx, j = 1..j_v_max) also written compactly, the multigrid's rhs.
template <typename T>
__global__ void __launch_bounds__(NT) k_pcd_zwork(Grid g, const T* __restrict__ zu, const T* __restrict__ zv,
                                                  const T* __restrict__ rp, T* __restrict__ zuw,
                                                  T* __restrict__ zvw, T* __restrict__ bv = nullptr) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  // every operand loaded first, from indices clamped into the arrays (a thread uses only those
  // its point needs), so that no load waits for another
  const int iu = min(max(i, 0), nx), ju = min(max(j, 1), ny), ir = min(max(i, 1), nx - 1);
  const int iv = min(max(i, 1), nx), jv = min(max(j, 0), ny), jr = min(max(j, 1), ny - 1);
  const T zu0 = zu[IU(g, iu, ju)], ru1 = rp[IP(g, ir + 1, ju)], ru0 = rp[IP(g, ir, ju)];
  const T zv0 = zv[IV(g, iv, jv)], rv1 = rp[IP(g, iv, jr + 1)], rv0 = rp[IP(g, iv, jr)];
  // both results formed before either store
  const bool urow = j >= 1 && j <= ny, vcol = i >= 1 && i <= nx;
  const T du = zu0 - (ru1 - ru0) * (T)g.inv_dx;
  const T dv = zv0 - (rv1 - rv0) * (T)g.inv_dy;
  const T ou = (urow && i >= 1 && i <= nx - 1) ? du : ((urow && (i == 0 || i == nx)) ? zu0 : T(0));
  const T ov = (vcol && j >= 1 && j <= ny - 1) ? dv : ((vcol && (j == 0 || j == ny)) ? zv0 : T(0));
  if (i >= -nh && i <= nx + nh && j >= 1 - nh && j <= ny + nh) zuw[IU(g, i, j)] = ou;
  if (i >= 1 - nh && i <= nx + nh && j >= -nh && j <= ny + nh) {
    zvw[IV(g, i, j)] = ov;
    if (bv && vcol && j >= 1 && j <= (g.per_y ? ny : ny - 1)) bv[(i - 1) + (size_t)(j - 1) * nx] = ov;
  }
}

// One damped-Jacobi sweep (omega = 0.8) of the v block (src/mod_subsolvers.f90:342-370).
// v_old == nullptr means the zero start.  Periodic neighbours wrap (the exchange at the
// top of each sweep); wall rows carry rhs; all other points copy v_old.
template <typename T>
__global__ void __launch_bounds__(NT) k_jacobi_v(Grid g, const T* __restrict__ v_old, T* __restrict__ v_new,
                                                 const T* __restrict__ rhs, const T* __restrict__ rho_v,
                                                 const T* __restrict__ kap, const T* __restrict__ zeta,
                                                 const T* __restrict__ eta, T dt, const T* __restrict__ pen) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  if (i < 1 - nh || i > nx + nh || j < -nh || j > ny + nh) return;
  const int j_v_max = g.per_y ? ny : ny - 1;
  auto old = [&](int ii, int jj) -> T {
    if (!v_old) return T(0);
    if (g.per_x && (ii < 1 || ii > nx)) ii = (ii < 1) ? ii + nx : ii - nx;
    if (g.per_y && (jj < 1 || jj > ny)) jj = (jj < 1) ? jj + ny : jj - ny;
    return v_old[IV(g, ii, jj)];
  };
  T out;
  if (j >= 1 && j <= j_v_max && i >= 1 && i <= nx) {
    const T inv_dx2 = (T)g.inv_dx2, inv_dy2 = (T)g.inv_dy2;
    const T cE = (zeta[IW(g, i, j)] - eta[IW(g, i, j)]) * inv_dx2;
    const T cW = (zeta[IW(g, i - 1, j)] - eta[IW(g, i - 1, j)]) * inv_dx2;
    const T cN = kap[IP(g, i, j + 1)] * inv_dy2;
    const T cS = kap[IP(g, i, j)] * inv_dy2;
    T diag = (rho_v[IV(g, i, j)] / dt) + cE + cW + cN + cS;
    if (pen) diag = diag + pen[IV(g, i, j)];
    const T off = cE * old(i + 1, j) + cW * old(i - 1, j) + cN * old(i, j + 1) + cS * old(i, j - 1);
    const T vo = old(i, j);
    if (diag > (T)1.0e-15) out = (T)(1.0 - 0.8) * vo + (T)0.8 * ((rhs[IV(g, i, j)] + off) / diag);
    else out = vo;
  } else if (!g.per_y && i >= 1 && i <= nx && (j == 0 || j == ny)) {
    out = rhs[IV(g, i, j)];
  } else {
    out = v_old ? v_old[IV(g, i, j)] : T(0);
  }
  v_new[IV(g, i, j)] = out;
}

// u block sweep (src/mod_subsolvers.f90:280-307).
template <typename T>
__global__ void __launch_bounds__(NT) k_jacobi_u(Grid g, const T* __restrict__ u_old, T* __restrict__ u_new,
                                                 const T* __restrict__ rhs, const T* __restrict__ rho_u,
                                                 const T* __restrict__ kap, const T* __restrict__ zeta,
                                                 const T* __restrict__ eta, T dt, const T* __restrict__ pen) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  if (i < -nh || i > nx + nh || j < 1 - nh || j > ny + nh) return;
  const int i_u_max = g.per_x ? nx : nx - 1;
  auto old = [&](int ii, int jj) -> T {
    if (!u_old) return T(0);
    if (g.per_x && (ii < 1 || ii > nx)) ii = (ii < 1) ? ii + nx : ii - nx;
    if (g.per_y && (jj < 1 || jj > ny)) jj = (jj < 1) ? jj + ny : jj - ny;
    return u_old[IU(g, ii, jj)];
  };
  T out;
  if (j >= 1 && j <= ny && i >= 1 && i <= i_u_max) {
    const T inv_dx2 = (T)g.inv_dx2, inv_dy2 = (T)g.inv_dy2;
    const T cE = kap[IP(g, i + 1, j)] * inv_dx2;
    const T cW = kap[IP(g, i, j)] * inv_dx2;
    const T cN = (zeta[IW(g, i, j)] - eta[IW(g, i, j)]) * inv_dy2;
    const T cS = (zeta[IW(g, i, j - 1)] - eta[IW(g, i, j - 1)]) * inv_dy2;
    T diag = (rho_u[IU(g, i, j)] / dt) + cE + cW + cN + cS;
    if (pen) diag = diag + pen[IU(g, i, j)];
    const T off = cE * old(i + 1, j) + cW * old(i - 1, j) + cN * old(i, j + 1) + cS * old(i, j - 1);
    const T uo = old(i, j);
    if (diag > (T)1.0e-15) out = (T)(1.0 - 0.8) * uo + (T)0.8 * ((rhs[IU(g, i, j)] + off) / diag);
    else out = uo;
  } else if (!g.per_x && j >= 1 && j <= ny && (i == 0 || i == nx)) {
    out = rhs[IU(g, i, j)];
  } else {
    out = u_old ? u_old[IU(g, i, j)] : T(0);
  }
  u_new[IU(g, i, j)] = out;
}

// u gets the eta cross term of the solved v block, i=1..nx-1, j=1..ny (:91-98).  bu (optional):
// instead of updating z_uw, write the u block's rows (i = 1..i_u_max, j = 1..ny, the cross term
// on i <= nx-1 only) compactly, the multigrid's rhs.
template <typename T>
__global__ void k_pcd_cross(Grid g, const T* __restrict__ rv, const T* __restrict__ eta, T* __restrict__ zuw,
                            T* __restrict__ bu = nullptr) {
  int i, j;
  cell_ij(0, i, j);
  const int i_u_max = g.per_x ? g.nx : g.nx - 1;
  if (bu) {
    if (i > i_u_max || j > g.ny) return;
    const size_t o = IU(g, i, j);
    T val = zuw[o];
    if (i <= g.nx - 1) {
      const T inv_dx = (T)g.inv_dx, inv_dy = (T)g.inv_dy;
      const T cross = -inv_dy * (eta[IW(g, i, j)] * (rv[IV(g, i + 1, j)] - rv[IV(g, i, j)]) * inv_dx -
                                 eta[IW(g, i, j - 1)] * (rv[IV(g, i + 1, j - 1)] - rv[IV(g, i, j - 1)]) * inv_dx);
      val = val - cross;
    }
    bu[(i - 1) + (size_t)(j - 1) * i_u_max] = val;
    return;
  }
  if (i > g.nx - 1 || j > g.ny) return;
  const T inv_dx = (T)g.inv_dx, inv_dy = (T)g.inv_dy;
  const T cross = -inv_dy * (eta[IW(g, i, j)] * (rv[IV(g, i + 1, j)] - rv[IV(g, i, j)]) * inv_dx -
                             eta[IW(g, i, j - 1)] * (rv[IV(g, i + 1, j - 1)] - rv[IV(g, i, j - 1)]) * inv_dx);
  const size_t o = IU(g, i, j);
  zuw[o] = zuw[o] - cross;
}

// ---------------------------------------------------------------------------------
// mixture properties over the full padded arrays (src/mod_viscosity.f90:28-151).
__device__ inline double clamp01(double x) {
  const double m = (1.0 < x) ? 1.0 : x;     // Fortran min(one, x)
  return (m > 0.0) ? m : 0.0;               // Fortran max(zero, .)
}
__device__ inline double harmonic_mu(double c, double mu1, double mu2) {
  const double cc = clamp01(c);
  if (fabs(mu1 - mu2) < 1.0e-14) return mu1;
  const double den = mu2 * cc + mu1 * (1.0 - cc);
  if (den > 1.0e-15) return (mu1 * mu2) / den;
  return (mu2 < mu1) ? mu2 : mu1;
}

__global__ void k_mixture(Grid g, const double* __restrict__ c, double rho1, double rho2, double mu1,
                          double mu2, double* __restrict__ rho_c, double* __restrict__ rho_u,
                          double* __restrict__ rho_v, double* __restrict__ kap, double* __restrict__ zeta,
                          double* __restrict__ eta) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  auto rc = [&](int ii, int jj) -> double {
    const double cv = clamp01(c[IP(g, ii, jj)]);
    return rho1 * cv + rho2 * (1.0 - cv);
  };
  if (i >= 1 - nh && i <= nx + nh && j >= 1 - nh && j <= ny + nh) {
    rho_c[IP(g, i, j)] = rc(i, j);
    kap[IP(g, i, j)] = 2.0 * harmonic_mu(c[IP(g, i, j)], mu1, mu2);
  }
  if (i >= -nh && i <= nx + nh && j >= 1 - nh && j <= ny + nh) {
    int ii = i;
    if (ii < 1 - nh) ii = 1 - nh;
    if (ii > nx + nh - 1) ii = nx + nh - 1;
    rho_u[IU(g, i, j)] = 0.5 * (rc(ii, j) + rc(ii + 1, j));
  }
  if (i >= 1 - nh && i <= nx + nh && j >= -nh && j <= ny + nh) {
    int jj = j;
    if (jj < 1 - nh) jj = 1 - nh;
    if (jj > ny + nh - 1) jj = ny + nh - 1;
    rho_v[IV(g, i, j)] = 0.5 * (rc(i, jj) + rc(i, jj + 1));
  }
  if (i >= -nh && i <= nx + nh && j >= -nh && j <= ny + nh) {
    int ii = i, jj = j;
    if (ii < 1 - nh) ii = 1 - nh;
    if (ii > nx + nh - 1) ii = nx + nh - 1;
    if (jj < 1 - nh) jj = 1 - nh;
    if (jj > ny + nh - 1) jj = ny + nh - 1;
    const double cv = 0.25 * (c[IP(g, ii, jj)] + c[IP(g, ii + 1, jj)] + c[IP(g, ii, jj + 1)] + c[IP(g, ii + 1, jj + 1)]);
    const double m = harmonic_mu(cv, mu1, mu2);
    zeta[IW(g, i, j)] = 2.0 * m;
    eta[IW(g, i, j)] = m;
  }
}

// ---------------------------------------------------------------------------------
// VOF-TVD advection (src/mod_vof.f90:25-140).  Face fluxes are recomputed per use
// (each face by its two cells) with the same arithmetic as the reference's Flux arrays.
__device__ inline double van_leer(double r) { return (r <= 0.0) ? 0.0 : (2.0 * r) / (1.0 + r); }
// Superbee, psi(r) = max(0, min(2r, 1), min(r, 2)) (Roe 1986; the upper edge of Sweby's TVD region):
// an opt-in limiter, not the reference's.  Fortran min/max order, so the oracle's value bit for bit.
__device__ inline double superbee(double r) {
  const double a = (1.0 < 2.0 * r) ? 1.0 : 2.0 * r;
  const double b = (2.0 < r) ? 2.0 : r;
  const double m = (b > a) ? b : a;
  return (m > 0.0) ? m : 0.0;
}
// the face limiter: 0 van Leer (the reference), 1 Superbee
__device__ inline double vof_psi(double r, int lim) { return lim == 1 ? superbee(r) : van_leer(r); }

__device__ inline double vof_flux_x(const Grid& g, const double* c, const double* u, double dt, int i, int j,
                                    int lim) {
  const double uf = u[IU(g, i, j)];
  double num, den, rr, cfl, cf;
  if (uf >= 0.0) {
    num = c[IP(g, i, j)] - c[IP(g, i - 1, j)];
    den = c[IP(g, i + 1, j)] - c[IP(g, i, j)];
    rr = (fabs(den) > 1.0e-12) ? num / den : 0.0;
    cfl = fabs(uf) * dt * g.inv_dx;
    cf = c[IP(g, i, j)] + 0.5 * (1.0 - cfl) * vof_psi(rr, lim) * den;
  } else {
    num = c[IP(g, i + 2, j)] - c[IP(g, i + 1, j)];
    den = c[IP(g, i + 1, j)] - c[IP(g, i, j)];
    rr = (fabs(den) > 1.0e-12) ? num / den : 0.0;
    cfl = fabs(uf) * dt * g.inv_dx;
    cf = c[IP(g, i + 1, j)] - 0.5 * (1.0 - cfl) * vof_psi(rr, lim) * den;
  }
  return uf * clamp01(cf);
}
__device__ inline double vof_flux_y(const Grid& g, const double* c, const double* v, double dt, int i, int j,
                                    int lim) {
  const double vf = v[IV(g, i, j)];
  double num, den, rr, cfl, cf;
  if (vf >= 0.0) {
    num = c[IP(g, i, j)] - c[IP(g, i, j - 1)];
    den = c[IP(g, i, j + 1)] - c[IP(g, i, j)];
    rr = (fabs(den) > 1.0e-12) ? num / den : 0.0;
    cfl = fabs(vf) * dt * g.inv_dy;
    cf = c[IP(g, i, j)] + 0.5 * (1.0 - cfl) * vof_psi(rr, lim) * den;
  } else {
    num = c[IP(g, i, j + 2)] - c[IP(g, i, j + 1)];
    den = c[IP(g, i, j + 1)] - c[IP(g, i, j)];
    rr = (fabs(den) > 1.0e-12) ? num / den : 0.0;
    cfl = fabs(vf) * dt * g.inv_dy;
    cf = c[IP(g, i, j + 1)] - 0.5 * (1.0 - cfl) * vof_psi(rr, lim) * den;
  }
  return vf * clamp01(cf);
}

// dclamp (optional, p-kind): the volume fraction the final clamp to [0, 1] adds to each interior
// cell (clamped - unclamped update), for the step's volume budget.
__global__ void __launch_bounds__(NT) k_vof(Grid g, const double* __restrict__ cin, double* __restrict__ cout,
                                            const double* __restrict__ u, const double* __restrict__ v,
                                            double dt, double* __restrict__ fx, double* __restrict__ fy, int lim,
                                            double* __restrict__ dclamp) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  if (i >= 1 - nh && i <= nx + nh && j >= 1 - nh && j <= ny + nh) {
    double o = cin[IP(g, i, j)];
    if (i >= 1 && i <= nx && j >= 1 && j <= ny) {
      const double fe = vof_flux_x(g, cin, u, dt, i, j, lim), fw = vof_flux_x(g, cin, u, dt, i - 1, j, lim);
      const double fn = vof_flux_y(g, cin, v, dt, i, j, lim), fs = vof_flux_y(g, cin, v, dt, i, j - 1, lim);
      o = cin[IP(g, i, j)] - dt * ((fe - fw) * g.inv_dx + (fn - fs) * g.inv_dy);
      const double oc = clamp01(o);
      if (dclamp) dclamp[IP(g, i, j)] = oc - o;
      o = oc;
    }
    cout[IP(g, i, j)] = o;
  }
  if (fx && i >= -nh && i <= nx + nh && j >= 1 - nh && j <= ny + nh)
    fx[IU(g, i, j)] = (j >= 1 && j <= ny && i >= 0 && i <= nx) ? vof_flux_x(g, cin, u, dt, i, j, lim) : 0.0;
  if (fy && i >= 1 - nh && i <= nx + nh && j >= -nh && j <= ny + nh)
    fy[IV(g, i, j)] = (i >= 1 && i <= nx && j >= 0 && j <= ny) ? vof_flux_y(g, cin, v, dt, i, j, lim) : 0.0;
}

// Volume budget of one VOF update over the interior cells: sum C_old, sum C_new, sum of the clamp
// corrections (double-double, per-workgroup partials; k_fold_vol writes S_VOL_*).
__global__ void __launch_bounds__(NT) k_vol_sums(Grid g, const double* __restrict__ cold,
                                                 const double* __restrict__ cnew, const double* __restrict__ dclamp,
                                                 double* __restrict__ partials) {
  DD v[3] = {dd_from(0.0), dd_from(0.0), dd_from(0.0)};
  const int i = blockIdx.x * TX + (int)threadIdx.x + 1;
  const int tiles = (g.ny + TY - 1) / TY;
  for (int t = blockIdx.y; t < tiles; t += gridDim.y) {
    const int j = t * TY + (int)threadIdx.y + 1;
    if (i <= g.nx && j <= g.ny) {
      const size_t o = IP(g, i, j);
      v[0] = dd_add(v[0], dd_from(cold[o]));
      v[1] = dd_add(v[1], dd_from(cnew[o]));
      v[2] = dd_add(v[2], dd_from(dclamp[o]));
    }
  }
  publish_dd<3>(v, partials);
}
__global__ void __launch_bounds__(NT) k_fold_vol(int nb, double vol, const double* __restrict__ partials,
                                                 double* __restrict__ sc) {
  double tot[3];
  fold_dd<3>(partials, nb, tot);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    sc[S_VOL_OLD] = tot[0] * vol; sc[S_VOL_NEW] = tot[1] * vol; sc[S_VOL_CLAMP] = tot[2] * vol;
  }
}

// one SVOF Jacobi sweep (src/mod_vof.f90:168-190) with mirror/wrap neighbour reads
// of the previous iterate and a max-reduction of the update; predicated on F_S_DONE.
__global__ void __launch_bounds__(NT) k_svof_sweep(Grid g, const double* __restrict__ craw,
                                                   const double* __restrict__ cold, double* __restrict__ cnew,
                                                   double cx, double cy, double diag, double* __restrict__ partials,
                                                   const int* __restrict__ fl) {
  if (fl[F_S_DONE]) return;
  double v[1] = {0.0};
  const int i = blockIdx.x * TX + (int)threadIdx.x + 1;
  const int tiles = (g.ny + TY - 1) / TY;
  for (int t = blockIdx.y; t < tiles; t += gridDim.y) {
    const int j = t * TY + (int)threadIdx.y + 1;
    if (i <= g.nx && j <= g.ny) {
      const Nbr nb = cell_nbrs(g, i, j);
      // at walls the mirrored ghost equals the cell itself: cell_nbrs already maps to it
      const double o = (craw[nb.c] + cx * (cold[nb.e] + cold[nb.w]) + cy * (cold[nb.n] + cold[nb.s])) / diag;
      cnew[nb.c] = o;
      v[0] = fmax(v[0], fabs(o - cold[nb.c]));
    }
  }
  publish_max<1>(v, partials);
}
__global__ void __launch_bounds__(NT) k_fold_svof(int nb, int it, double tol, const double* __restrict__ partials,
                                                  double* __restrict__ sc, int* __restrict__ fl) {
  if (fl[F_S_DONE]) return;
  double tot[1];
  fold_max<1>(partials, nb, tot);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    sc[S_MAX0] = tot[0];
    fl[F_S_ITER] = it;
    if (tot[0] < tol) fl[F_S_DONE] = 1;
  }
}

// SVOF Jacobi sweeps in blocks of up to SVOF_S iterations per launch (temporal blocking):
// a 64 x 32 tile of owned cells with an SVOF_S-cell halo in LDS, every cell updated from the
// previous iterate exactly as k_svof_sweep does (same neighbours: wall neighbours are the cell
// itself, periodic ones wrap; same expression and order), the valid region shrinking by one
// cell per iteration.  Per iteration and tile, the max |new - old| over the owned cells goes to
// partials[tile * SVOF_S + t]; k_fold_svof_blk finds the first iteration below tol.  When that
// is not the block's last, the next launch (or the extra one after the loop) recomputes the
// block from its input -- which the ping-pong left intact -- up to that iteration: the result
// and the iteration count are those of one sweep per launch.
constexpr int SVOF_S = 8, SVOF_TW = 64, SVOF_TH = 32;
constexpr int SVOF_EW = SVOF_TW + 2 * SVOF_S, SVOF_EH = SVOF_TH + 2 * SVOF_S;
constexpr int SVOF_NT = 1024;                                       // threads per tile
constexpr int SVOF_CPT = (SVOF_EW * SVOF_EH + SVOF_NT - 1) / SVOF_NT;   // cells per thread
__global__ void __launch_bounds__(SVOF_NT) k_svof_block(Grid g, const double* __restrict__ craw,
                                                    const double* __restrict__ cin, double* __restrict__ cout,
                                                    double cx, double cy, double diag, int blk, int cnt,
                                                    double* __restrict__ partials, const int* __restrict__ fl) {
  bool redo = false;
  if (fl[F_S_DONE]) {
    if (fl[F_S_RBLK] + 1 != blk) return;
    redo = true;   // block blk - 1 again, from its input (this launch's output buffer)
    const double* t = cin;
    cin = cout;
    cout = const_cast<double*>(t);
    cnt = fl[F_S_RCNT];
  }
  if (cnt <= 0) return;
  // two iterates in LDS (read one, write the other): one barrier per iteration
  __shared__ double LB[2][SVOF_EW * SVOF_EH];
  const int nx = g.nx, ny = g.ny;
  const int i0 = (int)blockIdx.x * SVOF_TW, j0 = (int)blockIdx.y * SVOF_TH;   // 0-based owned corner
  double cr[SVOF_CPT], nw[SVOF_CPT];
  unsigned present = 0, owned = 0, clamp_e = 0, clamp_w = 0, clamp_n = 0, clamp_s = 0;
  // load: every load issued before the first LDS store
#pragma unroll
  for (int r = 0; r < SVOF_CPT; ++r) {
    const int q = (int)threadIdx.x + SVOF_NT * r;
    cr[r] = 0.0;
    nw[r] = 0.0;
    if (q >= SVOF_EW * SVOF_EH) continue;
    const int a = q % SVOF_EW, b = q / SVOF_EW;
    int gi = i0 - SVOF_S + a + 1, gj = j0 - SVOF_S + b + 1;   // 1-based cell indices
    if (g.per_x) gi = ((gi - 1) % nx + nx) % nx + 1;
    if (g.per_y) gj = ((gj - 1) % ny + ny) % ny + 1;
    if (gi < 1 || gi > nx || gj < 1 || gj > ny) continue;
    const unsigned bit = 1u << r;
    present |= bit;
    if (a >= SVOF_S && a < SVOF_S + SVOF_TW && b >= SVOF_S && b < SVOF_S + SVOF_TH && i0 + a - SVOF_S < nx &&
        j0 + b - SVOF_S < ny)
      owned |= bit;
    // a wall neighbour is the cell itself (as cell_nbrs); so is one past the region's edge,
    // which lies outside the valid cone anyway
    if ((!g.per_x && gi == nx) || a + 1 == SVOF_EW) clamp_e |= bit;
    if ((!g.per_x && gi == 1) || a == 0) clamp_w |= bit;
    if ((!g.per_y && gj == ny) || b + 1 == SVOF_EH) clamp_n |= bit;
    if ((!g.per_y && gj == 1) || b == 0) clamp_s |= bit;
    cr[r] = craw[IP(g, gi, gj)];
    nw[r] = cin[IP(g, gi, gj)];
  }
#pragma unroll
  for (int r = 0; r < SVOF_CPT; ++r)
    if (present & (1u << r)) LB[0][threadIdx.x + SVOF_NT * r] = nw[r];
  __syncthreads();
  __shared__ double wmax[SVOF_S][SVOF_NT / 64];   // per iteration and wavefront: max |new - old| over owned cells
  // rows of the extended tile a wavefront's 64 cells of slot r span (uniform per wavefront)
  const int wq0 = (int)(threadIdx.x & ~63u);
  for (int t = 0; t < cnt; ++t) {
    const double* L = LB[t & 1];
    double* Lo = LB[(t + 1) & 1];
    double mt = 0.0;
    // sweep t + 1 only matters on the rows within SVOF_S - t - 1 of the owned ones: rows
    // t + 1 .. EH - t - 2 (the next sweep reads one row further in, the owned rows last); a
    // wavefront whose cells all lie outside skips (its cells are never read again)
    const int rlo = t + 1, rhi = SVOF_EH - t - 2;
#pragma unroll
    for (int r = 0; r < SVOF_CPT; ++r) {
      const unsigned bit = 1u << r;
      const int b0 = (wq0 + SVOF_NT * r) / SVOF_EW, b1 = (wq0 + SVOF_NT * r + 63) / SVOF_EW;
      if (b1 < rlo || b0 > rhi) continue;
      if (!(present & bit)) continue;
      const int q = (int)threadIdx.x + SVOF_NT * r;
      const double e = L[(clamp_e & bit) ? q : q + 1], w = L[(clamp_w & bit) ? q : q - 1];
      const double n = L[(clamp_n & bit) ? q : q + SVOF_EW], s = L[(clamp_s & bit) ? q : q - SVOF_EW];
      const double o = (cr[r] + cx * (e + w) + cy * (n + s)) / diag;
      if (owned & bit) mt = fmax(mt, fabs(o - nw[r]));   // nw[r]: this cell's previous iterate
      nw[r] = o;
      Lo[q] = o;
    }
    mt = wave_max(mt);
    if ((threadIdx.x & 63) == 0) wmax[t][threadIdx.x >> 6] = mt;
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < SVOF_CPT; ++r) {
    if (!(owned & (1u << r))) continue;
    const int q = (int)threadIdx.x + SVOF_NT * r;
    const int gi = i0 + q % SVOF_EW - SVOF_S + 1, gj = j0 + q / SVOF_EW - SVOF_S + 1;
    cout[IP(g, gi, gj)] = nw[r];
  }
  if (!redo && threadIdx.x < SVOF_S) {
    const int t = (int)threadIdx.x;
    double m = 0.0;
    if (t < cnt)
      for (int w = 0; w < SVOF_NT / 64; ++w) m = fmax(m, wmax[t][w]);
    partials[((size_t)blockIdx.x + (size_t)blockIdx.y * gridDim.x) * SVOF_S + t] = m;
  }
}
// Fold of block blk (iterations it0 .. it0 + cnt - 1): the first iteration whose max change
// is below tol ends the smoothing (fl[F_S_DONE], fl[F_S_ITER], sc[S_MAX0] as k_fold_svof), and
// when it is not the block's last, fl[F_S_RBLK] / fl[F_S_RCNT] ask for the block's recomputation.
__global__ void __launch_bounds__(NT) k_fold_svof_blk(int nb, int blk, int it0, int cnt, double tol,
                                                      const double* __restrict__ partials, double* __restrict__ sc,
                                                      int* __restrict__ fl) {
  if (fl[F_S_DONE]) return;
  double tot[SVOF_S];
  fold_max<SVOF_S>(partials, nb, tot);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    for (int t = 0; t < cnt; ++t) {
      sc[S_MAX0] = tot[t];
      fl[F_S_ITER] = it0 + t;
      if (tot[t] < tol) {
        fl[F_S_DONE] = 1;
        fl[F_S_RBLK] = (t + 1 < cnt) ? blk : -10;
        fl[F_S_RCNT] = t + 1;
        break;
      }
    }
  }
}

// face normals from the smoothed C (src/mod_surface_tension.f90:41-60).
__global__ void k_normals(Grid g, const double* __restrict__ cs, double* __restrict__ nxu, double* __restrict__ nyv) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny;
  const double eps = 1.0e-12;
  if (j >= 1 && j <= ny && i >= 0 && i <= nx) {
    const double gx = (cs[IP(g, i + 1, j)] - cs[IP(g, i, j)]) * g.inv_dx;
    const double gy = 0.25 * (cs[IP(g, i, j + 1)] + cs[IP(g, i + 1, j + 1)] - cs[IP(g, i, j - 1)] - cs[IP(g, i + 1, j - 1)]) * g.inv_dy;
    nxu[IU(g, i, j)] = gx / sqrt(gx * gx + gy * gy + eps);
  }
  if (i >= 1 && i <= nx && j >= 0 && j <= ny) {
    const double gx = 0.25 * (cs[IP(g, i + 1, j)] + cs[IP(g, i + 1, j + 1)] - cs[IP(g, i - 1, j)] - cs[IP(g, i - 1, j + 1)]) * g.inv_dx;
    const double gy = (cs[IP(g, i, j + 1)] - cs[IP(g, i, j)]) * g.inv_dy;
    nyv[IV(g, i, j)] = gy / sqrt(gx * gx + gy * gy + eps);
  }
}

// kappa = -div n at centres, zero ghosts (:63-69).
__global__ void k_curv(Grid g, const double* __restrict__ nxu, const double* __restrict__ nyv, double* __restrict__ kc) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  if (i < 1 - nh || i > nx + nh || j < 1 - nh || j > ny + nh) return;
  double o = 0.0;
  if (i >= 1 && i <= nx && j >= 1 && j <= ny)
    o = -((nxu[IU(g, i, j)] - nxu[IU(g, i - 1, j)]) * g.inv_dx + (nyv[IV(g, i, j)] - nyv[IV(g, i, j - 1)]) * g.inv_dy);
  kc[IP(g, i, j)] = o;
}

// CSF face force sigma * kappa_face * grad C (src/mod_surface_tension.f90:78-109).
__global__ void k_csf(Grid g, const double* __restrict__ craw, const double* __restrict__ kc, double sigma,
                      double* __restrict__ fu, double* __restrict__ fv) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  if (i >= -nh && i <= nx + nh && j >= 1 - nh && j <= ny + nh) {
    double o = 0.0;
    if (j >= 1 && j <= ny && i >= 0 && i <= nx) {
      const double kf = 0.5 * (kc[IP(g, i, j)] + kc[IP(g, i + 1, j)]);
      o = sigma * kf * (craw[IP(g, i + 1, j)] - craw[IP(g, i, j)]) * g.inv_dx;
    }
    fu[IU(g, i, j)] = o;
  }
  if (i >= 1 - nh && i <= nx + nh && j >= -nh && j <= ny + nh) {
    double o = 0.0;
    if (i >= 1 && i <= nx && j >= 0 && j <= ny) {
      const double kf = 0.5 * (kc[IP(g, i, j)] + kc[IP(g, i, j + 1)]);
      o = sigma * kf * (craw[IP(g, i, j + 1)] - craw[IP(g, i, j)]) * g.inv_dy;
    }
    fv[IV(g, i, j)] = o;
  }
}

// ---------------------------------------------------------------------------------
// step preparation (src/mod_solver_coupled.f90:209-222).
// u_adv = 2 u^n - u^(n-1) (or u^n at step 1) and u_prev = u^n, over the full arrays.
__global__ void k_ab2(double* __restrict__ uadv, double* __restrict__ uprev, const double* __restrict__ un,
                      size_t n, int first) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const double a = un[k];
    uadv[k] = first ? a : 2.0 * a - uprev[k];
    uprev[k] = a;
  }
}
// x <- 2 x - x_prev and x_prev <- x (old value), elementwise (the extrapolated initial guess).
__global__ void k_extrap(double* x, double* x_prev, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const double a = x[k];
    x[k] = 2.0 * a - x_prev[k];
    x_prev[k] = a;
  }
}
// mdot = (rho1 - rho2) flux + rho2 u_adv (:383-384)
__global__ void k_amp(double* __restrict__ md, const double* __restrict__ flux, const double* __restrict__ uadv,
                      size_t n, double drho, double rho2) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride)
    md[k] = drho * flux[k] + rho2 * uadv[k];
}
// pen = alpha * phi_s (:410-411)
__global__ void k_scale(double* __restrict__ o, const double* __restrict__ a, size_t n, double s) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) o[k] = s * a[k];
}

struct RhsArgs {
  const double *un, *unm1, *vn, *vnm1, *rho_u, *rho_v, *fu, *fv, *penu, *penv, *zeta, *eta;
  double *Bu, *Bv, *Bp;
  double dt, gx, gy, u_solid, v_solid, u_lid;
  int bdf2, has_lid;
};
// B_u, B_v, B_p (src/mod_solver_coupled.f90:261-303) over the full padded arrays (zero outside owned rows).
__global__ void __launch_bounds__(NT) k_rhs(Grid g, RhsArgs a) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny, nh = g.nh;
  const int i_u_max = g.per_x ? nx : nx - 1;
  const int j_v_max = g.per_y ? ny : ny - 1;
  if (i >= -nh && i <= nx + nh && j >= 1 - nh && j <= ny + nh) {
    double b = 0.0;
    if (j >= 1 && j <= ny && i >= 1 && i <= i_u_max) {
      const size_t o = IU(g, i, j);
      if (a.bdf2) b = (a.rho_u[o] / a.dt) * (2.0 * a.un[o] - 0.5 * a.unm1[o]) + a.rho_u[o] * a.gx + a.fu[o];
      else b = (a.rho_u[o] / a.dt) * a.un[o] + a.rho_u[o] * a.gx + a.fu[o];
      if (a.penu) b = b + a.penu[o] * a.u_solid;
      if (a.has_lid && !g.per_y && j == ny)
        b = b + (2.0 * (a.zeta[IW(g, i, ny)] - a.eta[IW(g, i, ny)]) / (g.dy * g.dy)) * a.u_lid;
    }
    a.Bu[IU(g, i, j)] = b;
  }
  if (i >= 1 - nh && i <= nx + nh && j >= -nh && j <= ny + nh) {
    double b = 0.0;
    if (j >= 1 && j <= j_v_max && i >= 1 && i <= nx) {
      const size_t o = IV(g, i, j);
      if (a.bdf2) b = (a.rho_v[o] / a.dt) * (2.0 * a.vn[o] - 0.5 * a.vnm1[o]) + a.rho_v[o] * a.gy + a.fv[o];
      else b = (a.rho_v[o] / a.dt) * a.vn[o] + a.rho_v[o] * a.gy + a.fv[o];
      if (a.penv) b = b + a.penv[o] * a.v_solid;
    }
    a.Bv[IV(g, i, j)] = b;
  }
  if (i >= 1 - nh && i <= nx + nh && j >= 1 - nh && j <= ny + nh) a.Bp[IP(g, i, j)] = 0.0;
}

// max |u|, |v| over faces for the adaptive dt (src/mod_solver_coupled.f90:54-66).
__global__ void __launch_bounds__(NT) k_absmax(Grid g, const double* __restrict__ u, const double* __restrict__ v,
                                               double* __restrict__ partials) {
  double m[2] = {0.0, 0.0};
  const int i = blockIdx.x * TX + (int)threadIdx.x - g.nh;
  const int tiles = (g.rw + TY - 1) / TY;
  for (int t = blockIdx.y; t < tiles; t += gridDim.y) {
    const int j = t * TY + (int)threadIdx.y - g.nh;
    if (j >= 1 && j <= g.ny && i >= 0 && i <= g.nx) m[0] = fmax(m[0], fabs(u[IU(g, i, j)]));
    if (i >= 1 && i <= g.nx && j >= 0 && j <= g.ny) m[1] = fmax(m[1], fabs(v[IV(g, i, j)]));
  }
  publish_max<2>(m, partials);
}
__global__ void __launch_bounds__(NT) k_fold_absmax(int nb, const double* __restrict__ partials, double* __restrict__ sc) {
  double tot[2];
  fold_max<2>(partials, nb, tot);
  if (threadIdx.x == 0 && threadIdx.y == 0) { sc[S_DOT0] = tot[0]; sc[S_DOT1] = tot[1]; }
}

}  // namespace fugu
