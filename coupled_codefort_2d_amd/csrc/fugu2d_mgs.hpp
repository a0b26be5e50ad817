// fugu2d_mgs.hpp -- register-strip form of the fused multigrid level kernels (k_mg_pre /
// k_mg_post of fugu2d_mg.hpp): the same temporally blocked red-black smoothing, residual
// restriction and prolongation, with every operand of a cell held in registers instead of an
// LDS tile.
//
// Layout.  A workgroup of NW wavefronts owns an extended tile of 128 x NW R cells: lane l of wave w
// holds columns 2l and 2l+1 of rows wR .. wR + R - 1, with phi, b, D^-1 and the west and south
// face conductances of each of its cells (plus the east face of its odd column and the north
// face of its last row).  Within a row the two columns of a lane are a red and a black cell, so
// a half-sweep updates one cell per row in every lane.  The west neighbour of an even column
// and the east neighbour of an odd column sit in the next lane: they come over DPP wave shifts.
// The south / north neighbours of a wave's first / last row sit in the next wave: after each
// half-sweep every wave publishes its first and last rows in LDS (double-buffered, one barrier).
// Loads are coalesced (consecutive lanes read consecutive cell pairs) and each operand is read
// once per tile.
//
// Temporal blocking.  The interior (the cells a tile writes) is the extended tile less HP rings
// on each side; HP is even, so a lane's columns and a wave's row pairs are the children of one
// coarse cell.  Every cell of the extended tile is updated in every half-sweep: a cell's update
// reads its four neighbours only, so after half-sweep h (from 0) the cells within HP - 1 - h
// rings of the interior hold exactly the values of the one-launch-per-half-sweep smoother, and
// whatever the outer rings hold stays finite (outside the tile the shifts and the edge waves
// read +0).  Pre-smoothing: HP = 2 NU + 2 (the rings within 1 of the interior are exact after
// 2 NU half-sweeps, which the residual needs); post-smoothing: HP = 2 NU.
//
// Cells outside the level (beyond a wall) hold +0 and are never updated; wall faces keep their
// conductance, which then multiplies that +0, exactly as in the LDS-tile kernels.  A periodic
// direction wraps (its extent is even, so the colour of a cell is that of its unwrapped index).
// Each update performs the same operations on the same operands, in the same order, as the
// one-launch-per-half-sweep smoother (mg_nbr_sum, k_mg_rbgs), the residual restriction
// (k_mg_restrict) and the prolongation (k_mg_prolong), so the results are bit-identical.
#pragma once
#include "fugu2d_mg.hpp"

namespace fugu {

constexpr int MGS_EW = 128;   // extended tile width: two columns per lane

// DPP wave shifts by one lane (GFX9 DPP controls wave_shr:1 / wave_shl:1); a lane whose source
// lies outside the wavefront reads 0.
template <int CTRL>
__device__ inline float mgs_dpp(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ inline double mgs_dpp(double v) {
  const unsigned long long x = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)x, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(x >> 32), CTRL, 0xF, 0xF, true);
  return __builtin_bit_cast(double, (unsigned long long)(unsigned)lo | ((unsigned long long)(unsigned)hi << 32));
}
// lane l receives the value of lane l - 1 / lane l + 1
template <typename T> __device__ inline T mgs_from_left(T v) { return mgs_dpp<0x138>(v); }
template <typename T> __device__ inline T mgs_from_right(T v) { return mgs_dpp<0x130>(v); }

// Face index of the face west (south) of unwrapped cell index g: wrapped in a periodic
// direction (face n is stored with face 0's value), else clamped into 0..n (a face outside the
// level is read by cells that are never updated).
__device__ inline int mgs_face(int g, int n, int per) {
  return per ? mgt_index(g, n, 1) : min(max(g, 0), n);
}

template <int R, int HP, int NW>
struct MgsGeom {
  static_assert(HP % 2 == 0 && R % 2 == 0, "even halo and rows: lanes and row pairs own whole coarse cells");
  static constexpr int EH = NW * R;               // extended tile height
  static constexpr int TW = MGS_EW - 2 * HP;      // interior width
  static constexpr int TH = EH - 2 * HP;          // interior height
  static_assert(TH > 0, "tile taller than its halo");
};

// Interior origin (i0, j0) of this workgroup's tile: dispatch order places workgroup w on XCD
// w mod 8; renumbered so that XCD x owns a contiguous range of tiles (its L2 holds their shared
// halos).
template <int TW, int TH>
__device__ inline void mgs_tile(int& i0, int& j0) {
  const int W = gridDim.x * gridDim.y;
  const int w = blockIdx.x + blockIdx.y * gridDim.x;
  const int per = W / 8, rem = W % 8, x = w % 8, k = w / 8;
  const int t = (x < rem) ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
  i0 = (t % gridDim.x) * TW;
  j0 = (t / gridDim.x) * TH;
}

template <typename T, int R>
struct MgsCells {
  T ph[R][2];      // iterate
  T b[R][2];       // right-hand side
  T di[R][2];      // D^-1
  T cw[R][2];      // west face conductance (the east face of column 0 is column 1's west face)
  T cs[R + 1][2];  // south face conductance; row R: the north face of the last row
  T ce[R];         // east face conductance of column 1
  T gs[2], gn[2];  // the south neighbours of row 0 and the north neighbours of row R - 1
  int ic[2];       // level column of each column (wrapped or clamped)
  bool colin[2];   // the column lies in the level
  bool rowin[R];   // the row lies in the level (uniform over the wavefront)
};

// The wavefront's index in the workgroup, as a uniform value: row indices, offsets and row base
// pointers then live in scalar registers, and each load is a scalar base plus the lane's 32-bit
// column offset.
__device__ inline int mgs_wave() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// Rows of a wavefront: level row index j (wrapped in a periodic direction, else clamped) and
// south face index fs (the face below the row: wrapped, else clamped into 0..ny) of the
// unwrapped rows g0, g0 + 1, ...  Uniform values (scalar code): the first row is wrapped once,
// the next ones step by one.
struct MgsRows {
  int g, j, fs, ny, per;
  __device__ MgsRows(int g0, int ny_, int per_) : g(g0), ny(ny_), per(per_) {
    j = g0;
    if (per) {
      j = g0 % ny_;   // one scalar division per wavefront
      if (j < 0) j += ny_;
    } else {
      j = min(max(g0, 0), ny_ - 1);
    }
    fs = per ? j : min(max(g0, 0), ny_);
  }
  __device__ void next() {
    ++g;
    if (per) {
      j = (j + 1 == ny) ? 0 : j + 1;
      fs = j;
    } else {
      j = min(max(g, 0), ny - 1);
      fs = min(max(g, 0), ny);
    }
  }
};

// Load the cells of the thread (all loads in flight together) and form D^-1 as k_mg_dinv does.
// HAS_M: mass term; HAS_PHI: initial iterate phi0 (else zero); HAS_PC: the parent's coarse
// correction phic added to it.
template <typename T, int R, int HP, bool HAS_M, bool HAS_PHI, bool HAS_PC>
__device__ inline void mgs_init(const MgDims& f, int i0, int j0, const T* cx, const T* cy, const T* mass,
                                const T* b, const T* phi0, const T* phic, MgsCells<T, R>& s) {
  const int lane = threadIdx.x & 63, w = mgs_wave();
  const int nx = f.nx, ny = f.ny, rx = f.nx + 1, nxc = (f.nx + 1) / 2;
  int fw[2], icc[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int gi = i0 - HP + 2 * lane + c;
    s.ic[c] = mgt_index(gi, nx, f.per_x);
    fw[c] = mgs_face(gi, nx, f.per_x);
    icc[c] = s.ic[c] >> 1;
    s.colin[c] = f.per_x || (unsigned)gi < (unsigned)nx;
  }
  const int fe = mgs_face(i0 - HP + 2 * lane + 2, nx, f.per_x);
  T mm[R][2], pc[R][2];
  MgsRows rows(j0 - HP + w * R, ny, f.per_y);
#pragma unroll
  for (int r = 0; r <= R; ++r, rows.next()) {
    const int gj = rows.g;
    const T* cyr = cy + (size_t)rows.fs * nx;
#pragma unroll
    for (int c = 0; c < 2; ++c) s.cs[r][c] = mgt_ld(cyr, (unsigned)s.ic[c]);
    if (r == R) break;
    const int j = rows.j;
    s.rowin[r] = f.per_y || (unsigned)gj < (unsigned)ny;
    const T* cxr = cx + (size_t)j * rx;
    const T* br = b + (size_t)j * nx;
    s.ce[r] = mgt_ld(cxr, (unsigned)fe);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      s.b[r][c] = mgt_ld(br, (unsigned)s.ic[c]);
      s.cw[r][c] = mgt_ld(cxr, (unsigned)fw[c]);
      mm[r][c] = HAS_M ? mgt_ld(mass + (size_t)j * nx, (unsigned)s.ic[c]) : T(0);
      s.ph[r][c] = HAS_PHI ? mgt_ld(phi0 + (size_t)j * nx, (unsigned)s.ic[c]) : T(0);
      pc[r][c] = HAS_PC ? mgt_ld(phic + (size_t)(j >> 1) * nxc, (unsigned)icc[c]) : T(0);
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const T p = HAS_PC ? s.ph[r][c] + pc[r][c] : s.ph[r][c];
      s.ph[r][c] = (s.colin[c] && s.rowin[r]) ? p : T(0);
      T dg = s.cw[r][c] + (c ? s.ce[r] : s.cw[r][1]) + s.cs[r][c] + s.cs[r + 1][c];
      if (HAS_M) dg += mm[r][c];
      s.di[r][c] = (dg > T(1.0e-14)) ? T(1) / dg : T(0);
    }
  }
  s.gs[0] = s.gs[1] = s.gn[0] = s.gn[1] = T(0);
}

// Neighbour rows between wavefronts.  Row 0 of column c has colour c and row R - 1 colour 1 - c
// (R even), so after a half-sweep of colour K a wave's changed boundary values are row 0 of
// column K and row R - 1 of column 1 - K: the north neighbours gn[K] of the wave below and the
// south neighbours gs[1 - K] of the wave above.  xs: [2 buffers][waves][first/last row][column]
// [lane]; the buffers alternate, so one barrier per exchange suffices.
template <typename T, int R, int NW, int K>
__device__ inline void mgs_exchange(MgsCells<T, R>& s, T (*xs)[NW][2][2][64], int buf) {
  const int lane = threadIdx.x & 63, w = mgs_wave();
  xs[buf][w][0][K][lane] = s.ph[0][K];
  xs[buf][w][1][1 - K][lane] = s.ph[R - 1][1 - K];
  __syncthreads();
  if (w > 0) s.gs[1 - K] = xs[buf][w - 1][1][1 - K][lane];
  if (w < NW - 1) s.gn[K] = xs[buf][w + 1][0][K][lane];
}
// Both colours (the loaded iterate).
template <typename T, int R, int NW>
__device__ inline void mgs_exchange_all(MgsCells<T, R>& s, T (*xs)[NW][2][2][64], int buf) {
  const int lane = threadIdx.x & 63, w = mgs_wave();
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    xs[buf][w][0][c][lane] = s.ph[0][c];
    xs[buf][w][1][c][lane] = s.ph[R - 1][c];
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if (w > 0) s.gs[c] = xs[buf][w - 1][1][c][lane];
    if (w < NW - 1) s.gn[c] = xs[buf][w + 1][0][c][lane];
  }
}

// The four neighbour values and conductances of cell (r, c) in the current state.
template <typename T, int R>
__device__ inline T mgs_nbr_sum(const MgsCells<T, R>& s, int r, int c) {
  const T pW = c ? s.ph[r][0] : mgs_from_left(s.ph[r][1]);
  const T pE = c ? mgs_from_right(s.ph[r][0]) : s.ph[r][1];
  const T pS = r > 0 ? s.ph[r - 1][c] : s.gs[c];
  const T pN = r < R - 1 ? s.ph[r + 1][c] : s.gn[c];
  const T ce = c ? s.ce[r] : s.cw[r][1];
  return s.cw[r][c] * pW + ce * pE + s.cs[r][c] * pS + s.cs[r + 1][c] * pN;
}

// One half-sweep of colour K over the whole extended tile.  zero_other: the other colour is
// known to be zero (the first half-sweep from a zero iterate).
template <typename T, int R, int K>
__device__ inline void mgs_half(MgsCells<T, R>& s, bool zero_other) {
  T nv[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = (K + r) & 1;   // a cell's colour is (column + row) & 1 in the extended tile
    const T sum = zero_other ? T(0) : mgs_nbr_sum(s, r, c);
    nv[r] = (s.b[r][c] + sum) * s.di[r][c];
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int c = (K + r) & 1;
    if (s.colin[c] && s.rowin[r]) s.ph[r][c] = nv[r];
  }
}

// Store the interior cells of the tile that lie in the level.
template <typename T, int R, int HP, int NW>
__device__ inline void mgs_store(const MgDims& f, int i0, int j0, const MgsCells<T, R>& s, T* out) {
  using G = MgsGeom<R, HP, NW>;
  const int lane = threadIdx.x & 63, w = mgs_wave();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int lj = w * R + r, gj = j0 - HP + lj;
    if (lj < HP || lj >= HP + G::TH || gj >= f.ny) continue;
    T* orow = out + (size_t)gj * f.nx;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int li = 2 * lane + c, gi = i0 - HP + li;
      if (li >= HP && li < HP + G::TW && gi < f.nx) orow[gi] = s.ph[r][c];
    }
  }
}

// NU pre-smoothing sweeps (red first) from zero (HAS_PHI false) or from phi_in, then (bc set)
// the residual summed onto the coarse cells of the tile (coarse phi zeroed).
template <typename T, int NU, int R, int NW, bool HAS_M, bool HAS_PHI>
__device__ inline void mgs_pre_body(const MgDims& f, const T* cx, const T* cy, const T* m, const T* b, T* phi,
                                    T* bc, T* phic, const T* phi_in, T (*xs)[NW][2][2][64]) {
  constexpr int HP = 2 * NU + 2;
  using G = MgsGeom<R, HP, NW>;
  int i0, j0;
  mgs_tile<G::TW, G::TH>(i0, j0);
  MgsCells<T, R> s;
  mgs_init<T, R, HP, HAS_M, HAS_PHI, false>(f, i0, j0, cx, cy, m, b, phi_in, (const T*)nullptr, s);
  if (HAS_PHI) mgs_exchange_all<T, R, NW>(s, xs, 1);
#pragma unroll
  for (int h = 0; h < 2 * NU; ++h) {   // red (colour 0) first
    const bool last = h + 1 == 2 * NU && !bc;   // the residual needs the final neighbours
    if (h & 1) {
      mgs_half<T, R, 1>(s, false);
      if (!last) mgs_exchange<T, R, NW, 1>(s, xs, 1);
    } else {
      mgs_half<T, R, 0>(s, h == 0 && !HAS_PHI);
      if (!last) mgs_exchange<T, R, NW, 0>(s, xs, 0);
    }
  }
  mgs_store<T, R, HP, NW>(f, i0, j0, s, phi);
  if (!bc) return;   // smoothing only (a truncated hierarchy's coarsest level)
  const int lane = threadIdx.x & 63, w = mgs_wave();
  const int nxc = (f.nx + 1) / 2, nyc = (f.ny + 1) / 2;
  const int I = (i0 - HP) / 2 + lane;
  const bool lane_in = 2 * lane >= HP && 2 * lane < HP + G::TW && I < nxc;
  MgsRows rows(j0 - HP + w * R, f.ny, f.per_y);
#pragma unroll
  for (int q = 0; q < R / 2; ++q) {
    const int lj = w * R + 2 * q;
    const int J = (j0 - HP + lj) / 2;
    T res[2][2];
#pragma unroll
    for (int rr = 0; rr < 2; ++rr, rows.next()) {
      const int r = 2 * q + rr;
      T mm[2] = {T(0), T(0)};
      if (HAS_M) {   // the mass term again (an L2 hit) instead of a register per cell
        const T* mr = m + (size_t)rows.j * f.nx;
#pragma unroll
        for (int c = 0; c < 2; ++c) mm[c] = mgt_ld(mr, (unsigned)s.ic[c]);
      }
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const T ce = c ? s.ce[r] : s.cw[r][1];
        T dg = s.cw[r][c] + ce + s.cs[r][c] + s.cs[r + 1][c];
        if (HAS_M) dg += mm[c];
        const T nb = mgs_nbr_sum(s, r, c);
        res[rr][c] = s.b[r][c] - (dg * s.ph[r][c] - nb);
      }
    }
    if (lj < HP || lj >= HP + G::TH || J >= nyc || !lane_in) continue;
    T acc = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int i = 2 * I + (k & 1), j = 2 * J + (k >> 1);
      if (i >= f.nx || j >= f.ny) continue;
      acc += res[k >> 1][k & 1];
    }
    bc[I + (size_t)J * nxc] = acc;
    phic[I + (size_t)J * nxc] = T(0);
  }
}

// Wavefronts per SIMD the register allocation targets: four for a 512-thread FP32 workgroup (at
// most 128 VGPRs: two workgroups per CU, so one workgroup's loads overlap another's sweeps),
// else what one workgroup needs.
template <typename T, int NW>
constexpr int mgs_waves_per_eu() { return (sizeof(T) == 4 && NW <= 8) ? 4 : (NW + 3) / 4; }

// Same arguments and results as k_mg_pre (HAS_M: m is not null).
template <typename T, int NU, int R, int NW, bool HAS_M>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(mgs_waves_per_eu<T, NW>()))) k_mgs_pre(MgDims f, const T* __restrict__ cx, const T* __restrict__ cy,
                                                    const T* __restrict__ m, const T* __restrict__ b,
                                                    T* __restrict__ phi, T* __restrict__ bc, T* __restrict__ phic,
                                                    const T* __restrict__ phi_in) {
  __shared__ T xs[2][NW][2][2][64];
  if (phi_in) mgs_pre_body<T, NU, R, NW, HAS_M, true>(f, cx, cy, m, b, phi, bc, phic, phi_in, xs);
  else mgs_pre_body<T, NU, R, NW, HAS_M, false>(f, cx, cy, m, b, phi, bc, phic, phi_in, xs);
}

// phi_out = phi + phi_c(parent), then NU post-smoothing sweeps (black first).  psum (optional):
// the tile's sum of its outputs as a double-double partial, one per workgroup.  Same arguments
// and results as k_mg_post (HAS_M: m is not null).
template <typename T, int NU, int R, int NW, bool HAS_M>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(mgs_waves_per_eu<T, NW>()))) k_mgs_post(MgDims f, const T* __restrict__ cx, const T* __restrict__ cy,
                                                     const T* __restrict__ m, const T* __restrict__ b,
                                                     const T* __restrict__ phi, const T* __restrict__ phic,
                                                     T* __restrict__ phi_out, double* __restrict__ psum) {
  constexpr int HP = 2 * NU;
  using G = MgsGeom<R, HP, NW>;
  __shared__ T xs[2][NW][2][2][64];
  int i0, j0;
  mgs_tile<G::TW, G::TH>(i0, j0);
  MgsCells<T, R> s;
  mgs_init<T, R, HP, HAS_M, true, true>(f, i0, j0, cx, cy, m, b, phi, phic, s);
  mgs_exchange_all<T, R, NW>(s, xs, 1);
#pragma unroll
  for (int h = 0; h < 2 * NU; ++h) {   // black (colour 1) first
    const bool last = h + 1 == 2 * NU;
    if (h & 1) {
      mgs_half<T, R, 0>(s, false);
      if (!last) mgs_exchange<T, R, NW, 0>(s, xs, 1);
    } else {
      mgs_half<T, R, 1>(s, false);
      if (!last) mgs_exchange<T, R, NW, 1>(s, xs, 0);
    }
  }
  mgs_store<T, R, HP, NW>(f, i0, j0, s, phi_out);
  if (psum) {
    const int lane = threadIdx.x & 63, w = mgs_wave();
    DD v[1] = {dd_from(0.0)};
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int lj = w * R + r, gj = j0 - HP + lj;
      if (lj < HP || lj >= HP + G::TH || gj >= f.ny) continue;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int li = 2 * lane + c, gi = i0 - HP + li;
        if (li >= HP && li < HP + G::TW && gi < f.nx) v[0] = dd_add(v[0], dd_from((double)s.ph[r][c]));
      }
    }
    publish_dd<1>(v, psum);
  }
}

// Launch grid of a strip level kernel with NU sweeps, R rows per wavefront and NW wavefronts.
template <int NU, int R, int NW>
inline dim3 mgs_grid_pre(const MgDims& d) {
  using G = MgsGeom<R, 2 * NU + 2, NW>;
  return dim3((unsigned)((d.nx + G::TW - 1) / G::TW), (unsigned)((d.ny + G::TH - 1) / G::TH));
}
template <int NU, int R, int NW>
inline dim3 mgs_grid_post(const MgDims& d) {
  using G = MgsGeom<R, 2 * NU, NW>;
  return dim3((unsigned)((d.nx + G::TW - 1) / G::TW), (unsigned)((d.ny + G::TH - 1) / G::TH));
}

}  // namespace fugu
