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
// t[q n + k]), in chunks of SQ_CH terms:
//   k_seq_chunks   chunk totals: double-double sum of t and the sum of |t|;
//   k_seq_compact  each chunk certifies its terms and splits them into runs (consecutive terms
//                  of one certified binade, sign and no tie).  A run becomes two values, its first
//                  term t_a (added by a real FP add, which puts s in the binade) and
//                  G = Q * sum_{a<k<=b} rint(t_k / Q) (exact: |G| < 2^(e+1) is a multiple of Q);
//                  an uncertain term or a tie becomes one value, itself.  The values overwrite
//                  the chunk's terms in place, in order, with their count;
//   k_seq_fold     one wavefront lane per chain adds every chunk's values one by one from zero
//                  -- the same sequence of FP additions as the reference, with each run's
//                  middle collapsed into one exact addition.
// Adding G is exact, and adding +0.0 changes nothing (s is never -0.0: from +0.0, an FP sum is
// -0.0 only when both operands are).  Terms that are NaN or infinite fail certification and are
// added one by one, so they propagate as in the reference.
#pragma once
#include "fugu2d_common.hpp"

namespace fugu {

constexpr int SQ_CH = 4096;             // terms per chunk
constexpr int SQ_NT = 256;              // threads per chunk workgroup
constexpr int SQ_PT = SQ_CH / SQ_NT;    // consecutive terms per thread
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

// Workgroup (SQ_NT threads) exclusive scans; every thread gets its own prefix and the total.
__device__ inline DD seq_scan_dd(DD v, DD* lds, DD* total) {
  const int tid = threadIdx.x;
  lds[tid] = v;
  __syncthreads();
  for (int o = 1; o < SQ_NT; o <<= 1) {
    DD a = lds[tid];
    if (tid >= o) a = dd_add(lds[tid - o], a);
    __syncthreads();
    lds[tid] = a;
    __syncthreads();
  }
  const DD incl = lds[tid];
  *total = lds[SQ_NT - 1];
  const DD ex = tid ? lds[tid - 1] : dd_from(0.0);
  __syncthreads();
  (void)incl;
  return ex;
}
template <typename V>
__device__ inline V seq_scan(V v, V* lds, V* total) {
  const int tid = threadIdx.x;
  lds[tid] = v;
  __syncthreads();
  for (int o = 1; o < SQ_NT; o <<= 1) {
    V a = lds[tid];
    if (tid >= o) a = lds[tid - o] + a;
    __syncthreads();
    lds[tid] = a;
    __syncthreads();
  }
  *total = lds[SQ_NT - 1];
  const V ex = tid ? lds[tid - 1] : V(0);
  __syncthreads();
  return ex;
}

// Chunk totals (grid: nch x nv).
__global__ void __launch_bounds__(SQ_NT) k_seq_chunks(SeqArgs a, Skip sk, const int* __restrict__ fl) {
  if (seq_skip(sk, fl)) return;
  const int c = blockIdx.x, q = blockIdx.y, tid = threadIdx.x;
  const double* t = a.t + (long)q * a.n;
  DD s = dd_from(0.0);
  double A = 0.0;
#pragma unroll
  for (int m = 0; m < SQ_PT; ++m) {
    const long idx = (long)c * SQ_CH + tid + m * SQ_NT;
    if (idx < a.n) {
      const double x = t[idx];
      s = dd_add(s, dd_from(x));
      A = A + fabs(x);
    }
  }
  __shared__ DD ls[SQ_NT];
  __shared__ double la[SQ_NT];
  DD tot;
  double atot;
  (void)seq_scan_dd(s, ls, &tot);
  (void)seq_scan<double>(A, la, &atot);
  if (tid == 0) {
    double* o = a.ctot + ((long)q * a.nch + c) * 3;
    o[0] = tot.hi; o[1] = tot.lo; o[2] = atot;
  }
}

// Certify, split into runs and compact one chunk in place (grid: nch x nv).
__global__ void __launch_bounds__(SQ_NT) k_seq_compact(SeqArgs a, Skip sk, const int* __restrict__ fl) {
  if (seq_skip(sk, fl)) return;
  const int c = blockIdx.x, q = blockIdx.y, tid = threadIdx.x;
  double* t = a.t + (long)q * a.n;
  const long base = (long)c * SQ_CH;
  const int len = (int)min((long)SQ_CH, a.n - base);
  const double u = 0x1p-53;

  __shared__ double lt[SQ_CH];           // the chunk's terms, then the runs' integer sums
  __shared__ DD lsd[SQ_NT];
  __shared__ double lsa[SQ_NT];
  __shared__ int lsi[SQ_NT];
  __shared__ int lkey[SQ_NT + 1];
  __shared__ int lstart[SQ_NT + 1];
  __shared__ DD p0s;
  __shared__ double a0s;

  // prefix of the earlier chunks (any order: only an approximation of P is needed)
  {
    DD s = dd_from(0.0);
    double A = 0.0;
    for (int cc = tid; cc < c; cc += SQ_NT) {
      const double* o = a.ctot + ((long)q * a.nch + cc) * 3;
      DD v; v.hi = o[0]; v.lo = o[1];
      s = dd_add(s, v);
      A = A + o[2];
    }
    DD tot;
    double atot;
    (void)seq_scan_dd(s, lsd, &tot);
    (void)seq_scan<double>(A, lsa, &atot);
    if (tid == 0) { p0s = tot; a0s = atot; }
  }
  // coalesced load of the chunk into LDS, then SQ_PT consecutive terms per thread
  for (int m = tid; m < SQ_CH; m += SQ_NT) lt[m] = m < len ? t[base + m] : 0.0;
  __syncthreads();
  double x[SQ_PT];
#pragma unroll
  for (int m = 0; m < SQ_PT; ++m) x[m] = lt[tid * SQ_PT + m];
  const int nmine = max(0, min(SQ_PT, len - tid * SQ_PT));   // my valid terms

  // thread totals -> exclusive prefix over the chunk
  DD ts = dd_from(0.0);
  double ta = 0.0;
  for (int m = 0; m < nmine; ++m) { ts = dd_add(ts, dd_from(x[m])); ta = ta + fabs(x[m]); }
  DD dummy;
  double adummy;
  DD P = seq_scan_dd(ts, lsd, &dummy);
  double A = seq_scan<double>(ta, lsa, &adummy);
  P = dd_add(p0s, P);
  // an upper bound of the true prefix sum of |t|: the computed sums of positive terms are low by
  // at most a relative gamma_n, far inside the factor 2 of E below
  A = a0s + A;

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
  // run starts: the chunk's first term, a special term or one after it, a change of binade/sign
  lkey[tid + 1] = key[SQ_PT - 1];
  if (tid == 0) lkey[0] = 0;
  __syncthreads();
  int start[SQ_PT];
  {
    int prev = lkey[tid];
#pragma unroll
    for (int m = 0; m < SQ_PT; ++m) {
      const bool first = (tid == 0 && m == 0);
      start[m] = (m < nmine) && (first || key[m] == 0 || prev == 0 || key[m] != prev);
      prev = key[m];
    }
  }
  lstart[tid] = start[0];
  if (tid == 0) lstart[SQ_NT] = 1;
  __syncthreads();
  // run ends (runs of two or more terms emit their G there)
  int endr[SQ_PT];
#pragma unroll
  for (int m = 0; m < SQ_PT; ++m) {
    const int L = tid * SQ_PT + m;   // the term's index in the chunk
    int ns;                          // does the next term start a run (or the chunk end)?
    if (L + 1 >= len) ns = 1;
    else if (m + 1 < SQ_PT) ns = start[m + 1];
    else ns = lstart[tid + 1];
    endr[m] = (L < len) && !start[m] && ns;
  }
  int nst = 0, nval = 0;
#pragma unroll
  for (int m = 0; m < SQ_PT; ++m) { nst += start[m]; nval += start[m] + endr[m]; }
  int tot_runs, tot_vals;
  int run = seq_scan<int>(nst, lsi, &tot_runs) - 1;
  int slot = seq_scan<int>(nval, lsi, &tot_vals);
  // integer sums of the runs (exact; wrap-around is harmless, the totals fit)
  unsigned long long* R = reinterpret_cast<unsigned long long*>(lt);
  for (int m = tid; m < SQ_CH; m += SQ_NT) R[m] = 0ull;
  __syncthreads();
  {
    int rr = run;
#pragma unroll
    for (int m = 0; m < SQ_PT; ++m) {
      rr += start[m];
      if (m < nmine && !start[m]) atomicAdd(&R[rr], (unsigned long long)(long long)r[m]);
    }
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
  if (tid == 0) a.cnt[(long)q * a.nch + c] = tot_vals;
}

// Serial fold (grid: nv; one chain per workgroup): lane 0 adds every chunk's values in order
// while the other wavefronts stage the next piece in LDS.
__global__ void __launch_bounds__(SQ_NT) k_seq_fold(SeqArgs a, Skip sk, const int* __restrict__ fl) {
  if (seq_skip(sk, fl)) return;
  const int q = blockIdx.x, tid = threadIdx.x;
  const double* t = a.t + (long)q * a.n;
  const int* cnt = a.cnt + (long)q * a.nch;
  __shared__ int off[SQ_WIN + 1];
  __shared__ int lsi[SQ_NT];
  __shared__ double buf[2][SQ_BUF];
  double s = 0.0;
  for (int w0 = 0; w0 < a.nch; w0 += SQ_WIN) {
    const int wn = min(SQ_WIN, a.nch - w0);
    // offsets of the window's chunks in the virtual concatenation of their values
    int my[SQ_WIN / SQ_NT];
    int loc = 0;
#pragma unroll
    for (int m = 0; m < SQ_WIN / SQ_NT; ++m) {
      const int cc = tid * (SQ_WIN / SQ_NT) + m;
      my[m] = cc < wn ? cnt[w0 + cc] : 0;
      loc += my[m];
    }
    int T;
    int ex = seq_scan<int>(loc, lsi, &T);
#pragma unroll
    for (int m = 0; m < SQ_WIN / SQ_NT; ++m) {
      off[tid * (SQ_WIN / SQ_NT) + m] = ex;
      ex += my[m];
    }
    if (tid == 0) off[SQ_WIN] = T;
    __syncthreads();
    const int npc = (T + SQ_BUF - 1) / SQ_BUF;
    auto stage = [&](int pc, int b) {
      for (int j = pc * SQ_BUF + tid - 64; j < min(T, (pc + 1) * SQ_BUF); j += SQ_NT - 64) {
        int lo = 0, hi = wn;   // last chunk with off[cc] <= j (off is non-decreasing)
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (off[mid] <= j) lo = mid; else hi = mid;
        }
        buf[b][j - pc * SQ_BUF] = t[(long)(w0 + lo) * SQ_CH + (j - off[lo])];
      }
    };
    if (tid >= 64 && npc > 0) stage(0, 0);
    __syncthreads();
    for (int pc = 0; pc < npc; ++pc) {
      if (tid >= 64 && pc + 1 < npc) stage(pc + 1, (pc + 1) & 1);
      if (tid == 0) {
        const double* bb = buf[pc & 1];
        const int nn = min(SQ_BUF, T - pc * SQ_BUF);
        if (nn == SQ_BUF) {
#pragma unroll 32
          for (int m = 0; m < SQ_BUF; ++m) s = s + bb[m];
        } else {
          for (int m = 0; m < nn; ++m) s = s + bb[m];
        }
      }
      __syncthreads();
    }
  }
  if (tid == 0) a.tot[q] = s;
}

}  // namespace fugu
