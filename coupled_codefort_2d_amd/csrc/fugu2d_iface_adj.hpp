// fugu2d_iface_adj.hpp -- transposes of the interface chain (no counterpart in the reference: it
// has no adjoint): the reverse mode of the VOF-TVD advection (src/mod_vof.f90:38-140), the SVOF
// filter (:145-194) at its sweep count, the curvature (src/mod_surface_tension.f90:25-72) and the
// CSF force (:78-109), at a given forward state.  Everything is in gather form (each output
// position sums what the forward statements that read it send back), so the results do not
// depend on thread order.
//
// Kinks take one side, the same convention throughout: the clamp to [0, 1] passes the derivative
// only strictly inside (0, 1); the limiters their one-sided slope at r (van Leer 2 / (1 + r)^2 for
// r > 0, else 0; Superbee the slope of the branch the forward select takes); the upwind switch the
// branch the forward takes (d|u|/du = 1 at u = 0); a face with |den| <= 1e-12 (r = 0, psi(0) den
// = 0) only its upwind cell.  The SVOF sweep count, which the forward's stopping test picks, is
// held at the forward's value: the sweeps are then linear.
#pragma once
#include "fugu2d_kernels.hpp"
#include "fugu2d_adjoint.hpp"

namespace fugu {

// ---- CSF^T ------------------------------------------------------------------------
// f_u(i, j) = sigma kf (c(i+1, j) - c(i, j)) / dx, kf = (k(i, j) + k(i+1, j)) / 2 on the u faces
// i = 0..nx, j = 1..ny (v alike).  cbar += dF/dc^T fbar; kbar = dF/dk^T fbar (whole p array).
__global__ void __launch_bounds__(NT) k_csf_T(Grid g, const double* __restrict__ craw, const double* __restrict__ kc,
                                              double sigma, const double* __restrict__ fub,
                                              const double* __restrict__ fvb, double* __restrict__ cbar,
                                              double* __restrict__ kbar) {
  int i, j;
  union_ij(g, i, j);
  if (!in_p(g, i, j)) return;
  const int nx = g.nx, ny = g.ny;
  double cb = 0.0, kb = 0.0;
  if (j >= 1 && j <= ny) {
    if (i - 1 >= 0 && i - 1 <= nx) {   // face (i-1, j): this cell is its east cell
      const double fb = fub[IU(g, i - 1, j)];
      const double kf = 0.5 * (kc[IP(g, i - 1, j)] + kc[IP(g, i, j)]);
      const double d = craw[IP(g, i, j)] - craw[IP(g, i - 1, j)];
      cb += sigma * kf * g.inv_dx * fb;
      kb += 0.5 * sigma * d * g.inv_dx * fb;
    }
    if (i >= 0 && i <= nx) {           // face (i, j): its west cell
      const double fb = fub[IU(g, i, j)];
      const double kf = 0.5 * (kc[IP(g, i, j)] + kc[IP(g, i + 1, j)]);
      const double d = craw[IP(g, i + 1, j)] - craw[IP(g, i, j)];
      cb -= sigma * kf * g.inv_dx * fb;
      kb += 0.5 * sigma * d * g.inv_dx * fb;
    }
  }
  if (i >= 1 && i <= nx) {
    if (j - 1 >= 0 && j - 1 <= ny) {
      const double fb = fvb[IV(g, i, j - 1)];
      const double kf = 0.5 * (kc[IP(g, i, j - 1)] + kc[IP(g, i, j)]);
      const double d = craw[IP(g, i, j)] - craw[IP(g, i, j - 1)];
      cb += sigma * kf * g.inv_dy * fb;
      kb += 0.5 * sigma * d * g.inv_dy * fb;
    }
    if (j >= 0 && j <= ny) {
      const double fb = fvb[IV(g, i, j)];
      const double kf = 0.5 * (kc[IP(g, i, j)] + kc[IP(g, i, j + 1)]);
      const double d = craw[IP(g, i, j + 1)] - craw[IP(g, i, j)];
      cb -= sigma * kf * g.inv_dy * fb;
      kb += 0.5 * sigma * d * g.inv_dy * fb;
    }
  }
  cbar[IP(g, i, j)] += cb;
  kbar[IP(g, i, j)] = kb;
}

// ---- curvature^T ------------------------------------------------------------------
// Face pass: kb is the curvature adjoint with its periodic ghosts already folded (H^T); the
// u faces' normal n_x = g_x / |g| gives (g_x bar, g_y bar) per face (v faces alike), zero on
// positions that are not faces of the forward loops.
__global__ void __launch_bounds__(NT) k_curv_T_faces(Grid g, const double* __restrict__ cs,
                                                     const double* __restrict__ kb, double* __restrict__ gux,
                                                     double* __restrict__ guy, double* __restrict__ gvx,
                                                     double* __restrict__ gvy) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny;
  const double eps = 1.0e-12;
  if (in_u(g, i, j)) {
    double ax = 0.0, ay = 0.0;
    if (i >= 0 && i <= nx && j >= 1 && j <= ny) {
      const double nb = ((i + 1 <= nx) ? kb[IP(g, i + 1, j)] : 0.0) * g.inv_dx - ((i >= 1) ? kb[IP(g, i, j)] : 0.0) * g.inv_dx;
      const double gx = (cs[IP(g, i + 1, j)] - cs[IP(g, i, j)]) * g.inv_dx;
      const double gy = 0.25 * (cs[IP(g, i, j + 1)] + cs[IP(g, i + 1, j + 1)] - cs[IP(g, i, j - 1)] - cs[IP(g, i + 1, j - 1)]) * g.inv_dy;
      const double n2 = gx * gx + gy * gy + eps, n3 = n2 * sqrt(n2);
      ax = nb * (gy * gy + eps) / n3;
      ay = -nb * gx * gy / n3;
    }
    gux[IU(g, i, j)] = ax;
    guy[IU(g, i, j)] = ay;
  }
  if (in_v(g, i, j)) {
    double ax = 0.0, ay = 0.0;
    if (i >= 1 && i <= nx && j >= 0 && j <= ny) {
      const double nb = ((j + 1 <= ny) ? kb[IP(g, i, j + 1)] : 0.0) * g.inv_dy - ((j >= 1) ? kb[IP(g, i, j)] : 0.0) * g.inv_dy;
      const double gx = 0.25 * (cs[IP(g, i + 1, j)] + cs[IP(g, i + 1, j + 1)] - cs[IP(g, i - 1, j)] - cs[IP(g, i - 1, j + 1)]) * g.inv_dx;
      const double gy = (cs[IP(g, i, j + 1)] - cs[IP(g, i, j)]) * g.inv_dy;
      const double n2 = gx * gx + gy * gy + eps, n3 = n2 * sqrt(n2);
      ay = nb * (gx * gx + eps) / n3;
      ax = -nb * gx * gy / n3;
    }
    gvx[IV(g, i, j)] = ax;
    gvy[IV(g, i, j)] = ay;
  }
}
// Cell pass: csbar += what the face gradients read from each cell.
__global__ void __launch_bounds__(NT) k_curv_T_cells(Grid g, const double* __restrict__ gux,
                                                     const double* __restrict__ guy, const double* __restrict__ gvx,
                                                     const double* __restrict__ gvy, double* __restrict__ csbar) {
  int k, l;
  union_ij(g, k, l);
  if (!in_p(g, k, l)) return;
  auto U_ = [&](const double* a, int i, int j) { return in_u(g, i, j) ? a[IU(g, i, j)] : 0.0; };
  auto V_ = [&](const double* a, int i, int j) { return in_v(g, i, j) ? a[IV(g, i, j)] : 0.0; };
  double s = 0.0;
  // u faces: g_x = (cs(i+1, j) - cs(i, j)) / dx; g_y = (cs(i, j+1) + cs(i+1, j+1) - cs(i, j-1) - cs(i+1, j-1)) / 4dy
  s += (U_(gux, k - 1, l) - U_(gux, k, l)) * g.inv_dx;
  s += 0.25 * g.inv_dy * (U_(guy, k, l - 1) + U_(guy, k - 1, l - 1) - U_(guy, k, l + 1) - U_(guy, k - 1, l + 1));
  // v faces: g_y = (cs(i, j+1) - cs(i, j)) / dy; g_x = (cs(i+1, j) + cs(i+1, j+1) - cs(i-1, j) - cs(i-1, j+1)) / 4dx
  s += (V_(gvy, k, l - 1) - V_(gvy, k, l)) * g.inv_dy;
  s += 0.25 * g.inv_dx * (V_(gvx, k - 1, l) + V_(gvx, k - 1, l - 1) - V_(gvx, k + 1, l) - V_(gvx, k + 1, l - 1));
  csbar[IP(g, k, l)] += s;
}

// ---- SVOF^T -----------------------------------------------------------------------
// The filter's output is interior values copied to every ghost by the halo exchange (periodic)
// and the Neumann mirror (walls), separately in x and y: ghost a takes the column src(a) = a + nx
// or a - nx (periodic), 1 - a or 2 nx + 1 - a (walls).  The transpose sums, for each interior
// cell, the adjoints of every padded position whose value it became (at most 3 x 3 of them).
__device__ inline int svof_pre(int i, int n, int nh, int per, int* out) {
  int m = 0;
  out[m++] = i;
  if (per) {
    if (i + n <= n + nh) out[m++] = i + n;
    if (i - n >= 1 - nh) out[m++] = i - n;
  } else {
    if (1 - i >= 1 - nh) out[m++] = 1 - i;
    if (2 * n + 1 - i <= n + nh) out[m++] = 2 * n + 1 - i;
  }
  return m;
}
__global__ void __launch_bounds__(NT) k_fill_T_p(Grid g, const double* __restrict__ abar, double* __restrict__ sb) {
  int i, j;
  union_ij(g, i, j);
  if (!in_p(g, i, j)) return;
  double s = 0.0;
  if (own_p(g, i, j)) {
    int xi[3], yj[3];
    const int mx = svof_pre(i, g.nx, g.nh, g.per_x, xi), my = svof_pre(j, g.ny, g.nh, g.per_y, yj);
    for (int b = 0; b < my; ++b)
      for (int a = 0; a < mx; ++a) s += abar[IP(g, xi[a], yj[b])];
  }
  sb[IP(g, i, j)] = s;
}
// One transposed sweep: y = sb_in / diag is the sweep's output adjoint; acc += y (the c_raw
// term), sb_out = N y with the forward's neighbour map (walls: the cell itself; periodic: wrap),
// which is symmetric.
__global__ void __launch_bounds__(NT) k_svof_T_sweep(Grid g, const double* __restrict__ sb_in,
                                                     double* __restrict__ sb_out, double* __restrict__ acc, double cx,
                                                     double cy, double diag) {
  int i, j;
  union_ij(g, i, j);
  if (!own_p(g, i, j)) return;
  const Nbr nb = cell_nbrs(g, i, j);
  acc[nb.c] += sb_in[nb.c] / diag;
  sb_out[nb.c] = cx * (sb_in[nb.e] / diag + sb_in[nb.w] / diag) + cy * (sb_in[nb.n] / diag + sb_in[nb.s] / diag);
}
// crawbar (interior) += acc + sb (the initial copy c_smooth = c_raw).
__global__ void __launch_bounds__(NT) k_svof_T_final(Grid g, const double* __restrict__ acc,
                                                     const double* __restrict__ sb, double* __restrict__ crawbar) {
  int i, j;
  union_ij(g, i, j);
  if (!own_p(g, i, j)) return;
  const size_t o = IP(g, i, j);
  crawbar[o] += acc[o] + sb[o];
}

// ---- VOF^T ------------------------------------------------------------------------
__device__ inline double vof_dpsi(double r, int lim) {
  if (lim == 1) {
    const double a = (1.0 < 2.0 * r) ? 1.0 : 2.0 * r;
    const double b = (2.0 < r) ? 2.0 : r;
    const double m = (b > a) ? b : a;
    if (!(m > 0.0)) return 0.0;
    if (b > a) return (2.0 < r) ? 0.0 : 1.0;
    return (1.0 < 2.0 * r) ? 0.0 : 2.0;
  }
  return (r <= 0.0) ? 0.0 : 2.0 / ((1.0 + r) * (1.0 + r));
}
// One face of the advection: stencil c0, c1, c2 (uf >= 0: c_{i-1}, c_i, c_{i+1}; uf < 0: c_i,
// c_{i+1}, c_{i+2}), dt and 1 / h (the Courant number formed as the forward forms it), fbar the
// flux adjoint.  Returns the velocity adjoint and the stencil's adjoints in d[0..2].
__device__ inline double vof_face_T(double uf, double c0, double c1, double c2, double dt, double inv_h, int lim,
                                    double fbar, double d[3]) {
  double num, den;
  const double sgn = (uf >= 0.0) ? 1.0 : -1.0;
  if (uf >= 0.0) { num = c1 - c0; den = c2 - c1; }
  else { num = c2 - c1; den = c1 - c0; }
  const bool big = fabs(den) > 1.0e-12;
  const double rr = big ? num / den : 0.0;
  const double cfl = fabs(uf) * dt * inv_h;
  const double A = 0.5 * (1.0 - cfl);
  const double k = dt * inv_h;
  const double psi = vof_psi(rr, lim);
  const double cf = c1 + sgn * A * psi * den;
  const double dcl = (cf > 0.0 && cf < 1.0) ? 1.0 : 0.0;
  const double ub = fbar * (clamp01(cf) + uf * dcl * (-0.5 * k * psi * den));
  const double cfb = fbar * uf * dcl;
  const double dp = big ? vof_dpsi(rr, lim) : 0.0;
  const double pn = big ? A * dp : 0.0;
  const double pd = big ? A * (psi - rr * dp) : 0.0;
  d[1] = cfb * (1.0 + pn - pd);
  if (uf >= 0.0) { d[0] = -cfb * pn; d[2] = cfb * pd; }
  else { d[0] = cfb * pd; d[2] = -cfb * pn; }
  return ub;
}
// Cell pass: cb is the output adjoint with its periodic ghosts folded (H^T of the final halo);
// tb = cb clamp'(unclamped update) on the interior (the unclamped update recomputed with the
// forward's own flux functions), 0 elsewhere; cbar_in = tb on the interior, cb elsewhere (the
// forward copies c_in there).
__global__ void __launch_bounds__(NT) k_vof_T_cells(Grid g, const double* __restrict__ cin,
                                                    const double* __restrict__ u, const double* __restrict__ v,
                                                    double dt, int lim, const double* __restrict__ cb,
                                                    double* __restrict__ tb, double* __restrict__ cbar_in) {
  int i, j;
  union_ij(g, i, j);
  if (!in_p(g, i, j)) return;
  const size_t o = IP(g, i, j);
  if (own_p(g, i, j)) {
    const double fe = vof_flux_x(g, cin, u, dt, i, j, lim), fw = vof_flux_x(g, cin, u, dt, i - 1, j, lim);
    const double fn = vof_flux_y(g, cin, v, dt, i, j, lim), fs = vof_flux_y(g, cin, v, dt, i, j - 1, lim);
    const double pre = cin[o] - dt * ((fe - fw) * g.inv_dx + (fn - fs) * g.inv_dy);
    const double t = (pre > 0.0 && pre < 1.0) ? cb[o] : 0.0;
    tb[o] = t;
    cbar_in[o] = t;
  } else {
    tb[o] = 0.0;
    cbar_in[o] = cb[o];
  }
}
// Face pass: flux adjoints from tb, the velocity adjoints (ubar, vbar: zero off the faces) and
// each face's three stencil adjoints (dx3[q], dy3[q], q = 0..2: the stencil's first, second and
// third cell).
struct VofTFaces {
  double* ubar;
  double* vbar;
  double* dx3[3];
  double* dy3[3];
};
__global__ void __launch_bounds__(NT) k_vof_T_faces(Grid g, const double* __restrict__ cin,
                                                    const double* __restrict__ u, const double* __restrict__ v,
                                                    double dt, int lim, const double* __restrict__ tb, VofTFaces f) {
  int i, j;
  union_ij(g, i, j);
  const int nx = g.nx, ny = g.ny;
  if (in_u(g, i, j)) {
    const size_t o = IU(g, i, j);
    double ub = 0.0, d[3] = {0.0, 0.0, 0.0};
    if (i >= 0 && i <= nx && j >= 1 && j <= ny) {
      const double fb = dt * g.inv_dx * (tb[IP(g, i + 1, j)] - tb[IP(g, i, j)]);
      const double uf = u[o];
      const int b = (uf >= 0.0) ? i - 1 : i;
      ub = vof_face_T(uf, cin[IP(g, b, j)], cin[IP(g, b + 1, j)], cin[IP(g, b + 2, j)], dt, g.inv_dx, lim, fb, d);
    }
    f.ubar[o] = ub;
    f.dx3[0][o] = d[0]; f.dx3[1][o] = d[1]; f.dx3[2][o] = d[2];
  }
  if (in_v(g, i, j)) {
    const size_t o = IV(g, i, j);
    double vb = 0.0, d[3] = {0.0, 0.0, 0.0};
    if (i >= 1 && i <= nx && j >= 0 && j <= ny) {
      const double fb = dt * g.inv_dy * (tb[IP(g, i, j + 1)] - tb[IP(g, i, j)]);
      const double vf = v[o];
      const int b = (vf >= 0.0) ? j - 1 : j;
      vb = vof_face_T(vf, cin[IP(g, i, b)], cin[IP(g, i, b + 1)], cin[IP(g, i, b + 2)], dt, g.inv_dy, lim, fb, d);
    }
    f.vbar[o] = vb;
    f.dy3[0][o] = d[0]; f.dy3[1][o] = d[1]; f.dy3[2][o] = d[2];
  }
}
// Cell gather: the x faces i = k-2..k+1 and y faces j = l-2..l+1 whose stencils hold the cell.
__global__ void __launch_bounds__(NT) k_vof_T_gather(Grid g, const double* __restrict__ u,
                                                     const double* __restrict__ v, VofTFaces f,
                                                     double* __restrict__ cbar_in) {
  int k, l;
  union_ij(g, k, l);
  if (!in_p(g, k, l)) return;
  const int nx = g.nx, ny = g.ny;
  double s = 0.0;
  if (l >= 1 && l <= ny)
    for (int fi = k - 2; fi <= k + 1; ++fi) {
      if (fi < 0 || fi > nx) continue;
      const size_t o = IU(g, fi, l);
      const int q = k - ((u[o] >= 0.0) ? fi - 1 : fi);
      if (q >= 0 && q <= 2) s += f.dx3[q][o];
    }
  if (k >= 1 && k <= nx)
    for (int fj = l - 2; fj <= l + 1; ++fj) {
      if (fj < 0 || fj > ny) continue;
      const size_t o = IV(g, k, fj);
      const int q = l - ((v[o] >= 0.0) ? fj - 1 : fj);
      if (q >= 0 && q <= 2) s += f.dy3[q][o];
    }
  cbar_in[IP(g, k, l)] += s;
}

}  // namespace fugu
