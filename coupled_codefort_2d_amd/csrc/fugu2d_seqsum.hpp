// fugu2d_seqsum.hpp -- the reference's sequential FP64 sums, reproduced bit for bit in parallel.
//
// The reference adds every global sum one term at a time: s_0 = 0, s_k = fl(s_{k-1} + t_k)
// (round to nearest even), j outer and i inner.  Addition is not associative, so a tree gives
// other roundings.  This file computes exactly the sequential s_n with almost all of the work in
// parallel, from one observation:
//
//   while the exact partial value x_k = s_{k-1} + t_k stays inside one binade
//   [2^(e), 2^(e+1)) of one sign, every representable double near it is a multiple of
//   Q = 2^(e-52), s_{k-1} is itself a multiple of Q, and so
//       fl(s_{k-1} + t_k) = s_{k-1} + Q * rint(t_k / Q)
//   unless t_k / Q is exactly halfway between two integers (then the tie is broken by the parity
//   of s_{k-1}).  The integers rint(t_k / Q) add exactly and in any order.
//
// Which binade each x_k lies in is certified without knowing s: the exact prefix P_k = t_1 + ... +
// t_k (double-double, any order) and the classical bound for recursive summation,
// |s_k - P_k| <= gamma_k * (|t_1| + ... + |t_k|), gamma_k = k u / (1 - k u), u = 2^-53
// (N. J. Higham, "Accuracy and Stability of Numerical Algorithms", 2nd ed., section 4.2), give an
// interval around x_k; when the whole interval lies in one binade of one sign the binade is certain.
//
// Pipeline per reduction (after a pass-1 kernel has written the terms, chain q's term k at
// t[q n + k]), in chunks of CH terms:
//   k_seq_chunks   chunk totals: double-double sum of t and the sum of |t|;
//   k_seq_compact  each chunk certifies its terms and splits them into runs (consecutive terms
//                  of one certified binade, sign and no tie).  A run becomes two values, its first
//                  term t_a (added by a real FP add, which puts s in the binade) and
//                  G = Q * sum_{a<k<=b} rint(t_k / Q) (exact: |G| < 2^(e+1) is a multiple of Q);
//                  an uncertain term or a tie becomes one value, itself.  The values overwrite
//                  the chunk's terms in place, in order, with their count;
//   k_seq_fold     one wavefront lane per chain adds every chunk's values one by one from zero
//                  -- the same sequence of FP additions as the reference, with each run's
//                  middle collapsed into one exact addition.  (Folding in the last chunk
//                  workgroup to finish instead, behind a ticket and device-scope fences, measured
//                  slower at every size: 1.05x at 2^16 terms, 2.2x at 2^22.)
// Adding G is exact, and adding +0.0 changes nothing (s is never -0.0: from +0.0, an FP sum is
// -0.0 only when both operands are).  Terms that are NaN or infinite fail certification and are
// added one by one, so they propagate as in the reference.
#pragma once
#include "fugu2d_common.hpp"

namespace fugu {

// Terms per chunk: a template parameter CH of the pipeline's kernels.  Short sums use short
// chunks (more workgroups, each thread's certification chain shorter), long sums long ones (fewer
// run values for the serial fold).  Measured on the drop-in's sums (one launch triple each, the
// same doubles whatever the chunk): the 256^2 droplet's reference-algorithm solve runs 0.95 /
// 1.34 / 1.40 solves/s with 4096 / 1024 / 512-term chunks; the 2048^2 150-iteration PCG takes
// 51 / 65 / 94 ms.
constexpr int SQ_CH_TINY = 512;
constexpr int SQ_CH_SHORT = 1024;
constexpr int SQ_CH_LONG = 4096;
constexpr int SQ_CH_MIN = SQ_CH_TINY;     // the smallest chunk (sizes the per-chunk buffers)
inline int sq_chunk(long n) { return n <= (1L << 16) ? SQ_CH_TINY : (n <= (1L << 18) ? SQ_CH_SHORT : SQ_CH_LONG); }
constexpr int SQ_NT = 256;              // threads per chunk workgroup
constexpr int SQ_NW = SQ_NT / 64;       // its wavefronts
template <int CH> struct SqGeom {
  static constexpr int PT = CH / SQ_NT;   // consecutive terms per thread
  static constexpr int LDP = PT + 1;      // padded LDS row of one thread's terms (odd: no bank conflicts)
};
constexpr int SQ_WIN = 4096;            // chunks per window of the serial fold
constexpr int SQ_BUF = 2048;            // values per staged piece of the serial fold

// A reference-order reduction: nv chains of n terms (chain q at t + q n), chunk totals ctot
// (3 doubles per chunk: hi, lo, sum |t|), value counts cnt and the chains' sums tot.
struct SeqArgs {
  double* t;
  long n;
  int nch;
  double* ctot;
  int* cnt;
  double* tot;
};
// Skip a launch whose pass-1 kernel skipped (fl[fi] >= fv); fi < 0: never.
struct Skip { int fi, fv; };
__device__ inline bool seq_skip(const Skip& s, const int* fl) { return s.fi >= 0 && fl[s.fi] >= s.fv; }

__device__ inline int seq_expo(double x) {   // unbiased binary exponent of a finite |x| > 0
  return (int)((__double_as_longlong(x) >> 52) & 0x7ff) - 1023;
}
__device__ inline double seq_pow2(int e) { return __longlong_as_double((long long)(e + 1023) << 52); }

// A prefix: the double-double sum of the terms and the sum of their magnitudes.
struct PA {
  DD p;
  double a;
};
__device__ inline PA pa_zero() { PA r; r.p = dd_from(0.0); r.a = 0.0; return r; }
__device__ inline PA pa_add(PA x, PA y) { PA r; r.p = dd_add(x.p, y.p); r.a = x.a + y.a; return r; }
__device__ inline PA pa_shfl_up(PA v, int o) {
  PA r;
  r.p.hi = __shfl_up(v.p.hi, o, 64); r.p.lo = __shfl_up(v.p.lo, o, 64); r.a = __shfl_up(v.a, o, 64);
  return r;
}
__device__ inline PA pa_shfl_xor(PA v, int o) {
  PA r;
  r.p.hi = __shfl_xor(v.p.hi, o, 64); r.p.lo = __shfl_xor(v.p.lo, o, 64); r.a = __shfl_xor(v.a, o, 64);
  return r;
}
// Sum over the wavefront (every lane gets it; the grouping does not matter: only an
// approximation of the prefix is needed, with the error bound below).
__device__ inline PA pa_wave_sum(PA v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = pa_add(v, pa_shfl_xor(v, o));
  return v;
}

// Workgroup (SQ_NT threads) exclusive scans: wavefront scans by lane shuffles, then one LDS
// exchange of the wavefront totals (`lds`: SQ_NW entries, not reused before a later barrier).
__device__ inline PA sq_scan_pa(PA v, PA* lds) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  PA inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const PA y = pa_shfl_up(inc, o);
    if (lane >= o) inc = pa_add(y, inc);
  }
  PA ex = pa_shfl_up(inc, 1);
  if (lane == 0) ex = pa_zero();
  if (lane == 63) lds[wid] = inc;
  __syncthreads();
  PA pre = pa_zero();
  for (int w = 0; w < wid; ++w) pre = pa_add(pre, lds[w]);
  return pa_add(pre, ex);
}
struct I2 { int a, b; };
__device__ inline I2 sq_scan_i2(I2 v, I2* lds, I2* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  I2 inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int ya = __shfl_up(inc.a, o, 64), yb = __shfl_up(inc.b, o, 64);
    if (lane >= o) { inc.a += ya; inc.b += yb; }
  }
  I2 ex;
  ex.a = __shfl_up(inc.a, 1, 64); ex.b = __shfl_up(inc.b, 1, 64);
  if (lane == 0) { ex.a = 0; ex.b = 0; }
  if (lane == 63) lds[wid] = inc;
  __syncthreads();
  I2 pre{0, 0}, tot{0, 0};
  for (int w = 0; w < SQ_NW; ++w) {
    if (w < wid) { pre.a += lds[w].a; pre.b += lds[w].b; }
    tot.a += lds[w].a; tot.b += lds[w].b;
  }
  *total = tot;
  ex.a += pre.a; ex.b += pre.b;
  return ex;
}

// Chunk totals (grid: nch x nv).
template <int CH>
__global__ void __launch_bounds__(SQ_NT) k_seq_chunks(SeqArgs a, Skip sk, const int* __restrict__ fl) {
  if (seq_skip(sk, fl)) return;
  const int c = blockIdx.x, q = blockIdx.y, tid = threadIdx.x;
  const double* t = a.t + (long)q * a.n;
  PA v = pa_zero();
#pragma unroll
  for (int m = 0; m < SqGeom<CH>::PT; ++m) {
    const long idx = (long)c * CH + tid + m * SQ_NT;
    if (idx < a.n) {
      const double x = t[idx];
      v.p = dd_add(v.p, dd_from(x));
      v.a = v.a + fabs(x);
    }
  }
  v = pa_wave_sum(v);
  __shared__ PA lw[SQ_NW];
  if ((tid & 63) == 0) lw[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) {
    PA s = lw[0];
    for (int w = 1; w < SQ_NW; ++w) s = pa_add(s, lw[w]);
    double* o = a.ctot + ((long)q * a.nch + c) * 3;
    o[0] = s.p.hi; o[1] = s.p.lo; o[2] = s.a;
  }
}

// The serial fold of chain q (one workgroup): lane 0 adds every chunk's values in order while
// wavefronts 1.. stage the next piece in LDS (all of them stage the first).  off: SQ_WIN + 1 ints;
// buf: 2 x SQ_BUF doubles.
template <int CH>
__device__ inline void sq_fold(const SeqArgs& a, int q, int* off, double* buf, I2* lsi) {
  const int tid = threadIdx.x;
  const double* t = a.t + (long)q * a.n;
  const int* cnt = a.cnt + (long)q * a.nch;
  constexpr int PW = SQ_WIN / SQ_NT;   // chunk counts per thread
  double s = 0.0;
  for (int w0 = 0; w0 < a.nch; w0 += SQ_WIN) {
    const int wn = min(SQ_WIN, a.nch - w0);
    // offsets of the window's chunks in the virtual concatenation of their values
    int my[PW];
    int loc = 0;
#pragma unroll
    for (int m = 0; m < PW; ++m) {
      const int cc = tid * PW + m;
      my[m] = cc < wn ? cnt[w0 + cc] : 0;
      loc += my[m];
    }
    I2 T2;
    const I2 e2 = sq_scan_i2(I2{loc, 0}, lsi, &T2);
    int ex = e2.a;
    const int T = T2.a;
#pragma unroll
    for (int m = 0; m < PW; ++m) {
      off[tid * PW + m] = ex;
      ex += my[m];
    }
    if (tid == 0) off[SQ_WIN] = T;
    __syncthreads();
    const int npc = (T + SQ_BUF - 1) / SQ_BUF;
    // the last chunk with off[cc] <= j (off is non-decreasing): a fixed-length search, so that a
    // thread's four searches and then its four loads are in flight together
    auto chunk_of = [&](int j) {
      int lo = 0;
#pragma unroll
      for (int s = SQ_WIN >> 1; s > 0; s >>= 1)
        if (lo + s < wn && off[lo + s] <= j) lo += s;
      return lo;
    };
    // stage piece pc into buffer b with threads t0.. (nt of them), four values per thread at a time
    auto stage = [&](int pc, int b, int t0, int nt) {
      const int j0 = pc * SQ_BUF, j1 = min(T, j0 + SQ_BUF);
      for (int j = j0 + (tid - t0); j < j1; j += 4 * nt) {
        int cc[4];
        double v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) cc[q] = (j + q * nt < j1) ? chunk_of(j + q * nt) : 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int jj = j + q * nt;
          v[q] = (jj < j1) ? t[(long)(w0 + cc[q]) * CH + (jj - off[cc[q]])] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (j + q * nt < j1) buf[b * SQ_BUF + (j + q * nt - j0)] = v[q];
      }
    };
    if (npc > 0) stage(0, 0, 0, SQ_NT);   // every wavefront: the fold waits for this piece anyway
    __syncthreads();
    for (int pc = 0; pc < npc; ++pc) {
      if (tid >= 64 && pc + 1 < npc) stage(pc + 1, (pc + 1) & 1, 64, SQ_NT - 64);
      if (tid == 0) {
        // the additions in order, 32 values at a time read into registers first: the LDS latency
        // is paid once per batch, not once per dependent addition
        const double* bb = buf + (pc & 1) * SQ_BUF;
        const int nn = min(SQ_BUF, T - pc * SQ_BUF);
        int m = 0;
        for (; m + 32 <= nn; m += 32) {
          double r[32];
#pragma unroll
          for (int q = 0; q < 32; ++q) r[q] = bb[m + q];
#pragma unroll
          for (int q = 0; q < 32; ++q) s = s + r[q];
        }
        for (; m < nn; ++m) s = s + bb[m];
      }
      __syncthreads();
    }
  }
  if (tid == 0) a.tot[q] = s;
}

// Certify, split into runs and compact one chunk in place (grid: nch x nv).
template <int CH>
__global__ void __launch_bounds__(SQ_NT) k_seq_compact(SeqArgs a, Skip sk, const int* __restrict__ fl) {
  if (seq_skip(sk, fl)) return;
  constexpr int SQ_CH = CH, SQ_PT = SqGeom<CH>::PT, SQ_LDP = SqGeom<CH>::LDP;
  const int c = blockIdx.x, q = blockIdx.y, tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  double* t = a.t + (long)q * a.n;
  const long base = (long)c * SQ_CH;
  const int len = (int)min((long)SQ_CH, a.n - base);
  const double u = 0x1p-53;

  // the chunk's terms, one padded row per thread; then the runs' integer sums
  __shared__ double lt[SQ_NT * SQ_LDP];
  __shared__ PA lpre[SQ_NW], lthr[SQ_NW];
  __shared__ int lkf[SQ_NW], lkl[SQ_NW];
  __shared__ I2 lcnt[SQ_NW];
  static_assert(SQ_NT * SQ_LDP >= SQ_CH, "LDS reuse");

  // coalesced load of the chunk: term L -> row L / SQ_PT, column L % SQ_PT
  for (int m = tid; m < SQ_CH; m += SQ_NT) lt[(m / SQ_PT) * SQ_LDP + (m % SQ_PT)] = m < len ? t[base + m] : 0.0;
  // prefix of the earlier chunks (any order: only an approximation of P is needed)
  {
    PA e = pa_zero();
    for (int cc = tid; cc < c; cc += SQ_NT) {
      const double* o = a.ctot + ((long)q * a.nch + cc) * 3;
      DD v; v.hi = o[0]; v.lo = o[1];
      e.p = dd_add(e.p, v);
      e.a = e.a + o[2];
    }
    e = pa_wave_sum(e);
    if (lane == 0) lpre[wid] = e;
  }
  __syncthreads();
  double x[SQ_PT];
#pragma unroll
  for (int m = 0; m < SQ_PT; ++m) x[m] = lt[tid * SQ_LDP + m];
  const int nmine = max(0, min(SQ_PT, len - tid * SQ_PT));   // my valid terms

  // thread totals -> exclusive prefix over the chunk, plus the earlier chunks
  PA ts = pa_zero();
  for (int m = 0; m < nmine; ++m) { ts.p = dd_add(ts.p, dd_from(x[m])); ts.a = ts.a + fabs(x[m]); }
  const PA ex = sq_scan_pa(ts, lthr);   // (its barrier also ends every read of lt above)
  PA p0 = lpre[0];
  for (int w = 1; w < SQ_NW; ++w) p0 = pa_add(p0, lpre[w]);
  DD P = dd_add(p0.p, ex.p);
  // an upper bound of the true prefix sum of |t|: the computed sums of positive terms are low by
  // at most a relative gamma_n, far inside the factor 2 of E below
  double A = p0.a + ex.a;
  unsigned long long* R = reinterpret_cast<unsigned long long*>(lt);
  for (int m = tid; m < SQ_CH; m += SQ_NT) R[m] = 0ull;

  // certification: key = sign * (exponent + 2048) of a certified binade, 0 = special term
  int key[SQ_PT];
  double r[SQ_PT];
#pragma unroll
  for (int m = 0; m < SQ_PT; ++m) {
    key[m] = 0;
    r[m] = 0.0;
    if (m < nmine) {
      P = dd_add(P, dd_from(x[m]));
      A = A + fabs(x[m]);
      const double k1 = (double)(base + tid * SQ_PT + m + 1);
      const double E = 2.0 * k1 * u * A;                         // >= gamma_k * sum |t| (margins)
      const double W = E + 4.0 * u * (fabs(P.hi) + E);           // + |P - P.hi| and the roundings
      const double lo = P.hi - W, hi = P.hi + W;
      if ((lo > 0.0 && hi > 0.0) || (lo < 0.0 && hi < 0.0)) {
        const int e1 = seq_expo(fabs(lo)), e2 = seq_expo(fabs(hi));
        if (e1 == e2 && e1 > -1000 && e1 < 1000) {
          const double Q = seq_pow2(e1 - 52);
          const double f = x[m] / Q;                               // exact (power of two)
          if (f - floor(f) != 0.5) {
            key[m] = (lo > 0.0 ? 1 : -1) * (e1 + 2048);
            r[m] = rint(f);                                        // |r| <= 2^54, an integer
          }
        }
      }
    }
  }
  // the neighbouring threads' boundary keys: lane shuffles inside a wavefront, LDS across
  int kprev = __shfl_up(key[SQ_PT - 1], 1, 64);
  int knext = __shfl_down(key[0], 1, 64);
  if (lane == 63) lkl[wid] = key[SQ_PT - 1];
  if (lane == 0) lkf[wid] = key[0];
  __syncthreads();
  if (lane == 0) kprev = wid > 0 ? lkl[wid - 1] : 0;
  if (lane == 63) knext = wid + 1 < SQ_NW ? lkf[wid + 1] : 0;
  // run starts: the chunk's first term, a special term or one after it, a change of binade/sign
  int start[SQ_PT];
  {
    int prev = kprev;
#pragma unroll
    for (int m = 0; m < SQ_PT; ++m) {
      const bool first = (tid == 0 && m == 0);
      start[m] = (m < nmine) && (first || key[m] == 0 || prev == 0 || key[m] != prev);
      prev = key[m];
    }
  }
  // does the next thread's first term start a run?  (it has terms only if all of mine are valid)
  const int klast = key[SQ_PT - 1];
  const int nstart0 = (knext == 0 || klast == 0 || knext != klast) ? 1 : 0;
  // run ends (runs of two or more terms emit their G there)
  int endr[SQ_PT];
#pragma unroll
  for (int m = 0; m < SQ_PT; ++m) {
    const int L = tid * SQ_PT + m;   // the term's index in the chunk
    int ns;                          // does the next term start a run (or the chunk end)?
    if (L + 1 >= len) ns = 1;
    else if (m + 1 < SQ_PT) ns = start[m + 1];
    else ns = nstart0;
    endr[m] = (L < len) && !start[m] && ns;
  }
  int nst = 0, nval = 0;
#pragma unroll
  for (int m = 0; m < SQ_PT; ++m) { nst += start[m]; nval += start[m] + endr[m]; }
  I2 tot2;
  const I2 ex2 = sq_scan_i2(I2{nst, nval}, lcnt, &tot2);   // (its barrier also orders R's zeroing)
  const int run = ex2.a - 1;
  int slot = ex2.b;
  // integer sums of the runs (exact; wrap-around is harmless, the totals fit): each thread adds
  // its share of every run it touches with one LDS atomic
  {
    int rr = run;
    long long acc = 0;
    bool have = false;
#pragma unroll
    for (int m = 0; m < SQ_PT; ++m) {
      if (start[m]) {
        if (have) atomicAdd(&R[rr], (unsigned long long)acc);
        ++rr;
        acc = 0;
        have = false;
      } else if (m < nmine) {
        acc += (long long)r[m];
        have = true;
      }
    }
    if (have) atomicAdd(&R[rr], (unsigned long long)acc);
  }
  __syncthreads();
  // values in order: a run start writes its term, a run end writes G = Q * R
  {
    int rr = run;
#pragma unroll
    for (int m = 0; m < SQ_PT; ++m) {
      rr += start[m];
      if (start[m]) t[base + slot++] = x[m];
      if (endr[m]) {
        const int e = (key[m] < 0 ? -key[m] : key[m]) - 2048;
        t[base + slot++] = seq_pow2(e - 52) * (double)(long long)R[rr];
      }
    }
  }
  if (tid == 0) a.cnt[(long)q * a.nch + c] = tot2.b;
}

// The serial fold (grid: nv, one chain per workgroup).
template <int CH>
__global__ void __launch_bounds__(SQ_NT) k_seq_fold(SeqArgs a, Skip sk, const int* __restrict__ fl) {
  if (seq_skip(sk, fl)) return;
  __shared__ int loff[SQ_WIN + 1];
  __shared__ double lbuf[2 * SQ_BUF];
  __shared__ I2 lfold[SQ_NW];
  sq_fold<CH>(a, blockIdx.x, loff, lbuf, lfold);
}

// The pipeline's three launches for chunks of CH terms (a.nch set for CH).
template <int CH>
inline void seq_launch(const SeqArgs& a, int nv, Skip skip, const int* fl, hipStream_t s) {
  const dim3 gc((unsigned)a.nch, (unsigned)nv);
  hipLaunchKernelGGL(k_seq_chunks<CH>, gc, dim3(SQ_NT), 0, s, a, skip, fl);
  hipLaunchKernelGGL(k_seq_compact<CH>, gc, dim3(SQ_NT), 0, s, a, skip, fl);
  hipLaunchKernelGGL(k_seq_fold<CH>, dim3((unsigned)nv), dim3(SQ_NT), 0, s, a, skip, fl);
}
// The pipeline for chunks of ch terms (512, 1024 or 4096); a.nch is set here.
inline void seq_launch_chunks(SeqArgs a, int ch, int nv, Skip skip, const int* fl, hipStream_t s) {
  a.nch = (int)((a.n + ch - 1) / ch);
  if (ch == SQ_CH_TINY) seq_launch<SQ_CH_TINY>(a, nv, skip, fl, s);
  else if (ch == SQ_CH_SHORT) seq_launch<SQ_CH_SHORT>(a, nv, skip, fl, s);
  else seq_launch<SQ_CH_LONG>(a, nv, skip, fl, s);
}

}  // namespace fugu
