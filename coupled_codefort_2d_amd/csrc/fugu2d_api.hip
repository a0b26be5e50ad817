// fugu2d_api.hip -- context, C ABI and host-side orchestration of the FUGU-2D hot path.
//
// One context = one device-resident domain on one GPU with one HIP stream.  Solver
// scalars live in device memory; the host reads back only the Krylov stop flags (one
// 12-byte copy per Krylov iteration) and the scalars a caller asks for.
#include "fugu2d.h"
#include "fugu2d_kernels.hpp"
#include "fugu2d_seqsum.hpp"
#include "fugu2d_io.hpp"
#include "fugu2d_mg.hpp"
#include "fugu2d_mgs.hpp"
#include "fugu2d_adjoint.hpp"
#include "fugu2d_iface_adj.hpp"

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

using namespace fugu;

namespace {

thread_local std::string g_err;

struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

#define HIPCHK(x)                                                                      \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) throw Error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCH(kern, grid, block, stream, ...)                  \
  do {                                                          \
    hipLaunchKernelGGL(kern, grid, block, 0, stream, __VA_ARGS__); \
    HIPCHK(hipGetLastError());                                  \
  } while (0)

// LAUNCH bracketed by timing events when the context's kernel timing is on.
// Timed launch: HIP events around the kernel on its stream, plus the launch's algorithmic
// bytes (unique fields read + written once) for the roofline.  With a launch log open, one
// line "name bytes 1" per launch (the HBM-counter pairing of hbm_traffic.py).
#define TLAUNCH(ctx, name, abytes, kern, grid, block, stream, ...)                \
  do {                                                                           \
    if ((ctx)->launch_log) std::fprintf((ctx)->launch_log, "%s %.0f 1\n", (name), (double)(abytes)); \
    KTimer* kt_ = (ctx)->kt_on ? kt_slot(ctx, name) : nullptr;                   \
    if (kt_) HIPCHK(hipEventRecord(kt_->ev[2 * kt_->used], stream));             \
    LAUNCH(kern, grid, block, stream, __VA_ARGS__);                              \
    if (kt_) {                                                                   \
      HIPCHK(hipEventRecord(kt_->ev[2 * kt_->used + 1], stream));                \
      kt_->bytes[kt_->used] = (double)(abytes);                                  \
      kt_->used++;                                                               \
    }                                                                            \
  } while (0)

template <typename T>
struct MgLevel {
  MgDims d;
  T *cx = nullptr, *cy = nullptr, *m = nullptr, *dinv = nullptr, *b = nullptr, *phi = nullptr;
  T* cinv = nullptr;   // coarsest level of at most MG_DENSE_MAX cells: its dense inverse
  T* phi2 = nullptr;   // the fused post-smoothing's output (the level's result when it ran)
  // a register-resident tail starting at this level: the quads' register images and the
  // coarsest level's LDS image (k_mg_tail_q_pack), valid for tq_nl levels until the next set-up
  TqQuad<T>* tq = nullptr;
  T* tqc = nullptr;
  bool tq_ok = false;
  size_t tq_nl = 0;
};

// A multigrid hierarchy from the finest extents down (mass: allocate the mass terms).
template <typename T>
std::vector<MgLevel<T>> mg_levels(fugu_ctx* c, MgDims d, bool mass, int max_levels = 0);

// Per-precision solver workspace.  Triplets are contiguous [u | v | p].
template <typename T>
struct Work {
  std::vector<MgLevel<T>> mg;       // pressure multigrid hierarchy (chosen fast forms)
  std::vector<MgLevel<T>> vmg[2];   // u and v velocity-block hierarchies (allocated on first use)
  size_t vmg_nl[2] = {0, 0};        // their levels in use (cut where the mass term dominates)
  int vmg_cyc[2] = {1, 1};          // V-cycles per solve of each block
  T *X = nullptr, *B = nullptr, *bh = nullptr, *r = nullptr, *rt = nullptr, *uu = nullptr;
  T *vv = nullptr, *ww = nullptr, *ss = nullptr, *tt = nullptr, *z = nullptr, *q = nullptr;
  // properties (aliases of the context's FP64 fields for T = double)
  T *rho_c = nullptr, *rho_u = nullptr, *rho_v = nullptr, *kap = nullptr, *zeta = nullptr, *eta = nullptr;
  T *uadv = nullptr, *vadv = nullptr, *penu = nullptr, *penv = nullptr, *mdu = nullptr, *mdv = nullptr;
  // preconditioner
  T *phi = nullptr, *pr = nullptr, *pz = nullptr, *pp0 = nullptr, *pp1 = nullptr, *Ap = nullptr, *dinv = nullptr;
  T *cx = nullptr, *cy = nullptr, *zuw = nullptr, *ju0 = nullptr, *ju1 = nullptr, *zvw = nullptr, *jv0 = nullptr, *jv1 = nullptr;
  bool ready = false;
};

}  // namespace

// Optional per-kernel device timing (HIP events around selected launches on the
// context stream); resolved after each synchronising call.
struct KTimer {
  std::vector<hipEvent_t> ev;   // start/stop pairs
  std::vector<double> bytes;    // algorithmic bytes of each pending launch
  int used = 0;
  double total_ms = 0.0, total_bytes = 0.0;
  long long count = 0;
};
constexpr int KT_POOL = 1024;

struct fugu_ctx {
  Grid g;
  fugu_params prm;
  hipStream_t s = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<void*> allocs;
  double* sc = nullptr;
  int* fl = nullptr;
  double* partials = nullptr;
  size_t partials_n = 0;   // its length in doubles
  // Summation order of the reference-algorithm reductions (coupled_dot, the PCG's dots and
  // means): true (default) = the reference's sequential order, bit for bit; false = the
  // double-double tree (fast).  The chosen fast forms and the adjoint always use the tree
  // (tree_depth > 0 inside them).  terms: the per-term buffer of the reference order,
  // 5 chains x (owned u + v + p points), with the chunk totals, value counts and chain sums of
  // the sequential-sum pipeline; allocated on first use.
  bool sum_ref = true;
  int tree_depth = 0;
  double* terms = nullptr;
  double* seq_ctot = nullptr;
  int* seq_cnt = nullptr;
  double* seq_tot = nullptr;
  int vof_limiter = FUGU_LIMITER_VAN_LEER;   // the VOF face limiter (van Leer: the reference's)
  double* dclamp = nullptr;                  // per-cell clamp correction of the step's VOF update
  double* h_sc = nullptr;
  int* h_fl = nullptr;
  int* d_hfl = nullptr;   // h_fl as the device addresses it (mapped pinned memory)
  size_t ntrip = 0;
  // resident state: (u, v, p) is one contiguous triplet
  double *X = nullptr, *u = nullptr, *v = nullptr, *p = nullptr;
  double *c = nullptr, *u_prev = nullptr, *v_prev = nullptr, *unm1 = nullptr, *vnm1 = nullptr;
  double *rho_c = nullptr, *rho_u = nullptr, *rho_v = nullptr, *kap = nullptr, *zeta = nullptr, *eta = nullptr;
  double *kcurv = nullptr, *fu = nullptr, *fv = nullptr, *phis_u = nullptr, *phis_v = nullptr;
  bool has_solid = false;
  // step temporaries
  double *uadv = nullptr, *vadv = nullptr, *cnew = nullptr, *csA = nullptr, *csB = nullptr, *cs = nullptr;
  double *nxu = nullptr, *nyv = nullptr, *fx = nullptr, *fy = nullptr, *mdu = nullptr, *mdv = nullptr;
  double *penu = nullptr, *penv = nullptr, *Bst = nullptr;
  double* p_prev = nullptr;   // p^(n-1) for FUGU_SOLVE_EXTRAPOLATE
  bool have_pprev = false;
  // fine-API staging
  double *su[4] = {}, *sv[4] = {}, *sp[4] = {};
  Work<double> wd;
  Work<float> wf;
  float *f_penu = nullptr, *f_penv = nullptr, *f_mdu = nullptr, *f_mdv = nullptr;
  std::vector<double*> kv, kz;   // outer FGMRES basis and preconditioned directions (FP64 triplets)
  std::vector<float*> kv32, kz32; // the same in FP32 for the mixed form's Arnoldi cycles
  double* fg = nullptr;          // outer FGMRES Hessenberg / Givens state (device)
  hipEvent_t fg_ev[2] = {nullptr, nullptr};
  int pcg_calls = 0;
  bool use_mg = false;   // pressure block by multigrid V-cycles instead of the reference PCG
  bool use_vmg = false;  // velocity blocks by multigrid V-cycles instead of three Jacobi sweeps
  int mg_cycles = 1;
  int mg_tail_cells = -1;    // coarse-tail start: < 0 the LDS-pool fit, > 0 a cell count, 0 never
  bool mg_fused = true;       // fused (temporal-blocking) level kernels instead of one launch per half-sweep
  bool mg_tail_stage = true; // stage the coarse tail in LDS when it fits
  bool mg_tail_quad = true;  // the coarse tail with its coefficients in registers where it applies
  // register-strip level kernels (fugu2d_mgs.hpp): 1 on the large and the small levels
  // (mg_strip_shape), 0 never, 2 / 3 the large / small shapes on every fused level (bit-identical
  // either way)
  int mg_strip = 1;
  // the shape rule's bounds (log2 of a level's cells): large strips from 2^mgs_large_log2 cells
  // (half that with three or more sweeps), small strips up to 2^mgs_small_log2
  int mgs_large_log2 = 21, mgs_small_log2 = 17;
  // terms per chunk of the reference-order sums: 0 by the sum's length (sq_chunk), else 512, 1024
  // or 4096 (the same sums bit for bit)
  int seq_chunk = 0;
  int fg_last_steps = 0;   // FGMRES steps of the last refined step solve (the default restart length)
  // red-black sweeps before and after each coarse correction: pressure and velocity hierarchies
  // (three on the velocity levels: fewer FGMRES steps, -5 % solve time on viscous-dominated blocks)
  int mg_nu_p = 2, mg_nu_v = 3;
  // matvec from LDS tiles of this many rows (4 or 8; 0: straight from the arrays); 8 rows:
  // -3.4 % FP64 and -1 % mixed solve time at 2048^2, no change at 512x1024
  int mv_tile = 8;
  bool gs_vec = true;  // Gram-Schmidt multi-dot with 16-byte loads
  bool pcd_cross = true;     // the PCD's eta cross term from the solved v block into the u block
  int vmg_levels = 0;        // velocity hierarchies: at most this many levels (0: coarsen fully)
  // velocity hierarchies cut where diffusion / mass < vmg_theta (0: never), the cut coarsest level
  // smoothed by vmg_coarse_nu red-black sweeps (2, 3 or 4).  Measured: 10 / 4 keeps the FGMRES
  // step counts of full depth on the 512 x 1024 bubble (2 levels instead of 3 at 3 / 2) and on
  // the 2048^2 inclusions (6-7 of 10 levels), with fewer launches per step.
  double vmg_theta = 10.0;
  double vmg_theta2 = 40.0;   // the cut for blocks solved by two or more V-cycles (measured at 2048^2:
                              // the step counts of theta 10, 3-4 % less time)
  int vmg_coarse_nu = 4;
  double fg_est = 0.9;       // FGMRES ends a cycle when its residual estimate is below fg_est * tol
  // second Gram-Schmidt pass when ||w'||^2 < reorth_frac ||w||^2: only on severe cancellation
  // (more than 99 % of ||w||^2 removed by the first pass).  With the PCD preconditioner the
  // Kahan-Parlett 1/2 fires on almost every step without changing a single iteration count or
  // final residual on the rising bubble (512x1024) or the 2048^2 inclusions, FP64 or mixed.
  double reorth_frac = 0.01;
  double pcd_visc = 0.0; // viscous term of the Schur approximation (0 = the reference PCD)
  double pcd_visc_coef = 1.0;   // its coefficient when FUGU_PRECOND_VISC is set
  // centre z_p before the multigrid pressure block (walls).  Off (zp_centre false) only for the
  // PCD applied by FGMRES directly to its basis vectors (fgmres_pcd): their pressure parts have
  // zero mean up to rounding (the pressure rows of B are zero and the discrete divergence of A X
  // sums to zero with walls), and the coarsest level projects the constant out anyway; measured:
  // the same step counts on both workloads, 1-2 % less time.  Inner Krylov vectors are centred.
  bool zp_centre = false;
  bool fgmres_pcd = false;
  bool pcd_fused_glue = true;   // the v store, its halo and the cross term in one launch (FUGU_PCD_FUSED_GLUE=0: three)
  // V-cycles per velocity block solve: vmg_cycles > 0 fixes it; 0 (default) takes 2 where the
  // block's finest diffusion-to-mass ratio exceeds vmg_cycle_ratio, else 1.  Measured: on the
  // 2048^2 inclusions (ratio ~1e4, viscous-dominated inclusions) a second cycle cuts the FGMRES
  // steps per solve from ~172 to ~95 (-20 % solve time, FP64 and mixed); on the 512 x 1024
  // bubble (ratio ~30) it costs more than it saves.
  int vmg_cycles = 0;
  double vmg_cycle_ratio = 1000.0;
  // FGMRES steps as HIP graphs (FUGU_GRAPHS=1): step k of a cycle (preconditioner, matvec,
  // Gram-Schmidt, the Hessenberg update, the next basis vector) is captured once and replayed,
  // so the ~40 launches of a step cost one graph launch on the host.  Every launch of a step
  // reads its per-solve scalars from device slots, and its buffers (basis k, direction k, the
  // multigrid levels, the work triplets) are fixed, so a replay repeats the captured work
  // exactly.  The graphs are rebuilt when anything they bake in changes (fg_key).  Off by
  // default: on the 512 x 1024 bubble the host keeps ahead of the device without them, and the
  // per-kernel cost (a few microseconds between dependent kernels) is on the device side.
  bool graphs = false;
  bool svof_blocked = true;   // SVOF sweeps temporally blocked (FUGU_SVOF_BLOCKED=0: one per launch)
  int svof_pred = 0;          // the last blocked SVOF's iteration count (its next call's first batch)
  struct FgKey {
    int prec = 0, mvflags = 0, nl0 = 0, nl1 = 0, flags = 0, mg_cycles = 0, cyc0 = 0, cyc1 = 0;
    double dt = 0.0, pcd_visc = 0.0;
    bool operator==(const FgKey& o) const {
      return prec == o.prec && mvflags == o.mvflags && nl0 == o.nl0 && nl1 == o.nl1 && flags == o.flags &&
             cyc0 == o.cyc0 && cyc1 == o.cyc1 &&
             mg_cycles == o.mg_cycles && dt == o.dt && pcd_visc == o.pcd_visc;
    }
  } fg_key;
  std::vector<hipGraphExec_t> fg_graph;
  // The reference Krylov solvers' iteration as a HIP graph (krylov_graphs; off unless
  // FUGU_KRYLOV_GRAPHS=1): the body of one BiCGStab / BiCGStab(2) iteration -- its matvecs, PCDs with
  // their inner PCG iterations, dots and updates, every launch predicated on device flags and reading
  // its iteration number from F_KITER -- is captured once per key and replayed, so the thousands of
  // dependent launches of an iteration cost one graph launch on the host.  Bit-identical to the
  // eager launches (the same kernels with the same arguments).  Off by default: the solve is
  // bound by the GPU's dependent-launch chain, not by host launch cost (1-7 % at 256^2), and a
  // process under rocprofv3 kernel tracing crashed inside the runtime at the first re-capture.
  bool krylov_graphs = false;
  struct KryKey {
    int order = 0, consistent = 0, mvflags = 0, max_iter = 0, seq = 0, tsize = 0, mv_tile = 0;
    double dt = 0.0, pcd_visc = 0.0;
    const void* work = nullptr;
    bool operator==(const KryKey& o) const {
      return order == o.order && consistent == o.consistent && mvflags == o.mvflags && max_iter == o.max_iter &&
             seq == o.seq && tsize == o.tsize && mv_tile == o.mv_tile && dt == o.dt && pcd_visc == o.pcd_visc &&
             work == o.work;
    }
  } kry_key;
  hipGraphExec_t kry_graph = nullptr;
  int kry_graph_pcg = 0;   // PCG solves one replay stands for
  std::vector<int> fg_graph_pcg;   // pressure-block calls one replay stands for
  // adjoint: solve_fgmres applies A_eff^T and the transposed PCD instead of A and the PCD
  bool adjoint = false;
  struct AdjWork {
    double *y = nullptr, *hu0 = nullptr, *hu1 = nullptr, *zuwb = nullptr, *rvb = nullptr, *hv0 = nullptr;
    double *hv1 = nullptr, *zvwb = nullptr, *rpb = nullptr, *phib = nullptr, *phi = nullptr;
    double *mu = nullptr, *rho = nullptr, *bmu = nullptr, *bmv = nullptr, *t2 = nullptr, *ga = nullptr;
    double *G = nullptr, *lam[3] = {nullptr, nullptr, nullptr}, *h = nullptr;
    double *grc = nullptr, *gru = nullptr, *grv = nullptr, *gk = nullptr, *gz = nullptr, *ge = nullptr;
    double *gmu = nullptr, *grho = nullptr;
    bool ready = false;
  } aw;
  // the transposed PCD's scratch, per storage type of the Krylov vectors
  template <typename T>
  struct PcdTBuf {
    T *y = nullptr, *hu0 = nullptr, *hu1 = nullptr, *zuwb = nullptr, *rvb = nullptr, *hv0 = nullptr;
    T *hv1 = nullptr, *zvwb = nullptr, *rpb = nullptr, *phib = nullptr, *phi = nullptr;
  };
  PcdTBuf<double> ptd;
  PcdTBuf<float> ptf;
  // scratch of the interface-chain transposes (VOF, SVOF, curvature, CSF)
  struct IfaceWork {
    double *cb = nullptr, *tb = nullptr, *kb = nullptr, *sb0 = nullptr, *sb1 = nullptr, *acc = nullptr;
    double *csb = nullptr, *cbar = nullptr, *cbar2 = nullptr, *kbar = nullptr;
    double *gux = nullptr, *guy = nullptr, *ubar = nullptr, *fub = nullptr, *dx3[3] = {nullptr, nullptr, nullptr};
    double *gvx = nullptr, *gvy = nullptr, *vbar = nullptr, *fvb = nullptr, *dy3[3] = {nullptr, nullptr, nullptr};
    bool ready = false;
  } iw;
  std::vector<double*> traj_u, traj_v, traj_fu, traj_fv;   // forward trajectory (or window) of the adjoint horizon
  std::vector<double*> traj_c;                             // ... and its volume fractions (moving interface)
  struct Checkpoint {
    double *u1 = nullptr, *u2 = nullptr, *v1 = nullptr, *v2 = nullptr, *p = nullptr, *c = nullptr;
  };
  std::vector<Checkpoint> cp;   // restart states of the checkpointed adjoint horizon
  int traj_n = 0;
  bool kt_on = false;
  std::map<std::string, KTimer> kt;
  bool debug = std::getenv("FUGU_DEBUG") != nullptr;
  // FUGU_LAUNCH_LOG=<path>: one line "kernel algorithmic_bytes timed" per launch of the
  // multi-vector kernels, in launch order, to pair with per-dispatch hardware counters
  FILE* launch_log = nullptr;
};

namespace {

KTimer* kt_slot(fugu_ctx* c, const char* name) {
  KTimer& t = c->kt[name];
  if (t.ev.empty()) {
    t.ev.resize(2 * KT_POOL);
    t.bytes.resize(KT_POOL);
    for (auto& e : t.ev) HIPCHK(hipEventCreate(&e));
  }
  return t.used < KT_POOL ? &t : nullptr;
}
void kt_resolve(fugu_ctx* c) {
  for (auto& kv : c->kt) {
    KTimer& t = kv.second;
    for (int k = 0; k < t.used; ++k) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, t.ev[2 * k], t.ev[2 * k + 1]));
      t.total_ms += ms;
      t.total_bytes += t.bytes[k];
      t.count++;
    }
    t.used = 0;
  }
}

template <typename T>
T* dalloc(fugu_ctx* c, size_t n) {
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, n * sizeof(T) + 256));
  HIPCHK(hipMemsetAsync(p, 0, n * sizeof(T) + 256, c->s));
  c->allocs.push_back(p);
  return (T*)p;
}

// Reductions in the reference's order unless a fast form or the adjoint is running.
bool seq_sums(const fugu_ctx* c) { return c->sum_ref && c->tree_depth == 0; }
// Sink of a reduction with `nterms` terms per chain: the term buffer (reference order) or the
// per-workgroup partials (tree).
Sink sink(fugu_ctx* c, long nterms) {
  if (!seq_sums(c)) return Sink{c->partials, 0, nullptr};
  if (!c->terms) {
    // at most 5 chains (the fused GCR(2) dots) x every owned point of a triplet
    const Grid& g = c->g;
    const size_t nmax = (size_t)3 * g.nx * g.ny;
    const size_t nch = (nmax + SQ_CH_MIN - 1) / SQ_CH_MIN;
    c->terms = dalloc<double>(c, 5 * nmax);
    c->seq_ctot = dalloc<double>(c, 5 * 3 * nch);
    c->seq_cnt = dalloc<int>(c, 5 * nch);
    c->seq_tot = dalloc<double>(c, 8);
  }
  return Sink{c->terms, nterms, c->seq_tot};
}
// The sequential-sum pipeline of nv chains after a pass-1 kernel wrote them into `sk`
// (reference order only; nothing in the tree order).  skip: the pass-1 kernel's skip flag.
void seq_reduce(fugu_ctx* c, const Sink& sk, int nv, Skip skip) {
  if (!sk.n) return;
  SeqArgs a;
  a.t = sk.p; a.n = sk.n;
  a.ctot = c->seq_ctot; a.cnt = c->seq_cnt; a.tot = c->seq_tot;
  const int ch = (c->seq_chunk == SQ_CH_TINY || c->seq_chunk == SQ_CH_SHORT || c->seq_chunk == SQ_CH_LONG)
                     ? c->seq_chunk
                     : sq_chunk(sk.n);
  seq_launch_chunks(a, ch, nv, skip, (const int*)c->fl, c->s);
  HIPCHK(hipGetLastError());
}
long cdot_terms(const Grid& g) {
  const long iu = g.per_x ? g.nx : g.nx - 1, jv = g.per_y ? g.ny : g.ny - 1;
  return (long)g.ny * iu + jv * (long)g.nx + (long)g.nx * g.ny;
}
long cell_terms(const Grid& g) { return (long)g.nx * g.ny; }
// Preconditioner switches of one solve (multigrid pressure / velocity blocks, the viscous Schur
// term, V-cycles per application): set for the solve, restored to the reference PCD when the
// scope ends, also when the solve throws, so later drop-in calls never inherit them.
struct PrecondScope {
  fugu_ctx* c;
  explicit PrecondScope(fugu_ctx* ctx) : c(ctx) { reset(c); }
  ~PrecondScope() { reset(c); }
  static void reset(fugu_ctx* c) {
    c->use_mg = false; c->use_vmg = false; c->pcd_visc = 0.0; c->mg_cycles = 1;
  }
  PrecondScope(const PrecondScope&) = delete;
  PrecondScope& operator=(const PrecondScope&) = delete;
};
// The chosen fast forms and the adjoint sum in the tree order (not parity-bound).
struct TreeSums {
  fugu_ctx* c;
  explicit TreeSums(fugu_ctx* ctx) : c(ctx) { c->tree_depth++; }
  ~TreeSums() { c->tree_depth--; }
  TreeSums(const TreeSums&) = delete;
  TreeSums& operator=(const TreeSums&) = delete;
};

inline unsigned nblk(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// (one-thread kernels: no host-memory copy, no stream synchronisation)
void set_slot(fugu_ctx* c, int slot, double v) { LAUNCH(k_set_slot, dim3(1), dim3(64), c->s, c->sc, slot, v); }
void reset_flags(fugu_ctx* c, bool relative_guards = false) {
  LAUNCH(k_reset_flags, dim3(1), dim3(64), c->s, c->fl, relative_guards ? 1 : 0, c->sc);
}
void read_flags(fugu_ctx* c) {
  HIPCHK(hipMemcpyAsync(c->h_fl, c->fl, F_COUNT * sizeof(int), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
}
void read_slots(fugu_ctx* c) {
  HIPCHK(hipMemcpyAsync(c->h_sc, c->sc, S_COUNT * sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
}

// host/device <-> device staging
void in(fugu_ctx* c, double* dst, const double* src, size_t n) {
  if (src) HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDefault, c->s));
}
void out(fugu_ctx* c, double* dst, const double* src, size_t n) {
  if (dst) HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDefault, c->s));
}
void sync(fugu_ctx* c) { HIPCHK(hipStreamSynchronize(c->s)); }

// ---------------------------------------------------------------------------------
template <typename T>
void halo(fugu_ctx* c, T* a, int kind, const int* skip = nullptr) {
  if (!c->g.per_x && !c->g.per_y) return;
  LAUNCH((k_halo<T>), grid_union(c->g), dim3(TX, TY), c->s, c->g, a, kind, skip);
}
template <typename T>
void halo_trip(fugu_ctx* c, T* X, const int* skip = nullptr) {
  halo(c, X, 0, skip);
  halo(c, X + c->g.nu, 1, skip);
  halo(c, X + c->g.nu + c->g.nv, 2, skip);
}
template <typename T>
void bcs(fugu_ctx* c, T* u, T* v, double lid, const int* skip = nullptr) {
  if (c->g.per_x && c->g.per_y) return;
  LAUNCH((k_velocity_bcs<T>), grid_union(c->g), dim3(TX, TY), c->s, c->g, u, v, lid, skip);
}
template <typename T>
void bcs_trip(fugu_ctx* c, T* X, double lid = 0.0, const int* skip = nullptr) {
  bcs(c, X, X + c->g.nu, lid, skip);
}

template <typename T>
void matvec(fugu_ctx* c, Work<T>& w, T* X, T* Y, double dt, int flags, const int* skip) {
  const Grid& g = c->g;
  halo_trip(c, X, skip);
  MatvecArgs<T> a;
  a.Xu = X; a.Xv = X + g.nu; a.Xp = X + g.nu + g.nv;
  a.Yu = Y; a.Yv = Y + g.nu; a.Yp = Y + g.nu + g.nv;
  a.uadv = w.uadv; a.vadv = w.vadv; a.rho_c = w.rho_c; a.rho_u = w.rho_u; a.rho_v = w.rho_v;
  a.kap = w.kap; a.zeta = w.zeta; a.eta = w.eta;
  a.penu = w.penu; a.penv = w.penv; a.mdu = w.mdu; a.mdv = w.mdv;
  a.dt = dt; a.flags = flags;
  // 13 fields per cell (X_u, X_v, X_p, rho_u, rho_v, rho_c, kappa, vertex mu, u_adv, v_adv -> Y_u, Y_v,
  // Y_p); AMP swaps rho_c, u_adv, v_adv for mdot_u, mdot_v (-1); the penalty adds pen_u, pen_v
  const double mv_fields = 13.0 - ((flags & MV_CONS) ? 1.0 : 0.0) + ((flags & MV_PEN) ? 2.0 : 0.0);
  const char* nm = sizeof(T) == 8 ? "matvec_f64" : "matvec_f32";
  const double ab = mv_fields * sizeof(T) * g.nx * g.ny;
  if (c->mv_tile == 8)
    TLAUNCH(c, nm, ab, (k_matvec_tile<T, 8>), dim3((unsigned)((g.ldw + TX - 1) / TX), (unsigned)((g.rw + 7) / 8)),
            dim3(TX, 8), c->s, g, a, skip);
  else if (c->mv_tile == 4)
    TLAUNCH(c, nm, ab, (k_matvec_tile<T, 4>), dim3((unsigned)((g.ldw + TX - 1) / TX), (unsigned)((g.rw + 3) / 4)),
            dim3(TX, 4), c->s, g, a, skip);
  else
    TLAUNCH(c, nm, ab, (k_matvec<T>), grid_union(g), dim3(TX, TY), c->s, g, a, skip);
}

template <typename T>
void cdot(fugu_ctx* c, int np, const T* const* a, const T* const* b, int phase, int k, int max_iter) {
  DotArgs<T> d;
  for (int q = 0; q < 5; ++q) { d.a[q] = q < np ? a[q] : nullptr; d.b[q] = q < np ? b[q] : nullptr; }
  d.np = np; d.phase = phase; d.k = k; d.max_iter = max_iter;
  const dim3 gr = grid_union_red(c->g), bl(TX, TY);
  const int nb = (int)nblocks(gr);
  const int* fl = c->fl;
  const Sink sk = sink(c, cdot_terms(c->g));
  switch (np) {
    case 1:
      LAUNCH((k_cdot<T, 1>), gr, bl, c->s, c->g, d, sk, fl);
      seq_reduce(c, sk, 1, Skip{F_BRK, 1});
      LAUNCH((k_fold_krylov<1>), dim3(1), dim3(NT), c->s, nb, phase, k, max_iter, sk, c->sc, c->fl);
      break;
    case 2:
      LAUNCH((k_cdot<T, 2>), gr, bl, c->s, c->g, d, sk, fl);
      seq_reduce(c, sk, 2, Skip{F_BRK, 1});
      LAUNCH((k_fold_krylov<2>), dim3(1), dim3(NT), c->s, nb, phase, k, max_iter, sk, c->sc, c->fl);
      break;
    case 5:
      LAUNCH((k_cdot<T, 5>), gr, bl, c->s, c->g, d, sk, fl);
      seq_reduce(c, sk, 5, Skip{F_BRK, 1});
      LAUNCH((k_fold_krylov<5>), dim3(1), dim3(NT), c->s, nb, phase, k, max_iter, sk, c->sc, c->fl);
      break;
    default: throw Error("cdot: unsupported pair count");
  }
}

// sc[slot] = interior sum of a cell field / (nx ny) (two passes); skip_mode as k_pcg_sum.
template <typename T>
void cell_mean(fugu_ctx* c, const T* a, int slot, int skip_mode) {
  const dim3 gr = grid_cells_red(c->g);
  const Sink sk = sink(c, cell_terms(c->g));
  LAUNCH((k_pcg_sum<T>), gr, dim3(TX, TY), c->s, c->g, a, sk, skip_mode, (const int*)c->fl);
  seq_reduce(c, sk, 1, skip_mode == 1 ? Skip{F_BRK, 1} : skip_mode == 2 ? Skip{F_P_DONE, 2} : Skip{-1, 0});
  LAUNCH(k_fold_sum, dim3(1), dim3(NT), c->s, c->g, (int)nblocks(gr), slot, skip_mode, sk, c->sc,
         (const int*)c->fl);
}
template <typename T>
void cdot1(fugu_ctx* c, const T* a, const T* b, int phase, int k = 0, int mi = 0) {
  const T* A[1] = {a};
  const T* Bp[1] = {b};
  cdot<T>(c, 1, A, Bp, phase, k, mi);
}

template <typename T>
void lin(fugu_ctx* c, int op, T* o, const T* x0, const T* x1 = nullptr, const T* x2 = nullptr,
         const T* x3 = nullptr, int c0 = -1, int c1 = -1, int c2 = -1, double s0 = 1.0, bool skip = true) {
  LinArgs<T> a;
  a.out = o; a.x0 = x0; a.x1 = x1; a.x2 = x2; a.x3 = x3; a.c0 = c0; a.c1 = c1; a.c2 = c2;
  a.s0 = s0; a.op = op; a.n = c->ntrip; a.skip_on = skip ? 1 : 0;
  LAUNCH((k_lin<T, T>), dim3(nblk(a.n)), dim3(256), c->s, a, c->sc, c->fl);
}

// ---------------------------------------------------------------------------------
// Workspaces
template <typename T>
void work_alloc(fugu_ctx* c, Work<T>& w, bool own_props) {
  if (w.ready) return;
  const Grid& g = c->g;
  const size_t n3 = c->ntrip;
  T** trips[] = {&w.X, &w.B, &w.bh, &w.r, &w.rt, &w.uu, &w.vv, &w.ww, &w.ss, &w.tt, &w.z, &w.q};
  for (T** t : trips) *t = dalloc<T>(c, n3);
  if (own_props) {
    w.rho_c = dalloc<T>(c, g.np); w.rho_u = dalloc<T>(c, g.nu); w.rho_v = dalloc<T>(c, g.nv);
    w.kap = dalloc<T>(c, g.np); w.zeta = dalloc<T>(c, g.nw); w.eta = dalloc<T>(c, g.nw);
    w.uadv = dalloc<T>(c, g.nu); w.vadv = dalloc<T>(c, g.nv);
  }
  w.phi = dalloc<T>(c, g.np); w.pr = dalloc<T>(c, g.np); w.pz = dalloc<T>(c, g.np);
  w.pp0 = dalloc<T>(c, g.np); w.pp1 = dalloc<T>(c, g.np); w.Ap = dalloc<T>(c, g.np); w.dinv = dalloc<T>(c, g.np);
  w.cx = dalloc<T>(c, g.nu); w.cy = dalloc<T>(c, g.nv);
  w.zuw = dalloc<T>(c, g.nu); w.ju0 = dalloc<T>(c, g.nu); w.ju1 = dalloc<T>(c, g.nu);
  w.zvw = dalloc<T>(c, g.nv); w.jv0 = dalloc<T>(c, g.nv); w.jv1 = dalloc<T>(c, g.nv);
  w.mg = mg_levels<T>(c, MgDims{g.nx, g.ny, g.per_x, g.per_y}, false);
  w.ready = true;
}

// FP64 workspace views the context's resident property fields.
void bind_props_d(fugu_ctx* c, bool pen, bool cons) {
  Work<double>& w = c->wd;
  w.rho_c = c->rho_c; w.rho_u = c->rho_u; w.rho_v = c->rho_v;
  w.kap = c->kap; w.zeta = c->zeta; w.eta = c->eta;
  w.uadv = c->uadv; w.vadv = c->vadv;
  w.penu = pen ? c->penu : nullptr; w.penv = pen ? c->penv : nullptr;
  w.mdu = cons ? c->mdu : nullptr; w.mdv = cons ? c->mdv : nullptr;
}

// FP32 copies of the properties for the mixed inner solve.
void bind_props_f(fugu_ctx* c, bool pen, bool cons) {
  Work<float>& w = c->wf;
  const Grid& g = c->g;
  auto cv = [&](float* o, const double* i, size_t n) {
    LAUNCH((k_convert<float, double>), dim3(nblk(n)), dim3(256), c->s, o, i, n, c->sc, -1, 1.0, 0);
  };
  cv(w.rho_c, c->rho_c, g.np); cv(w.rho_u, c->rho_u, g.nu); cv(w.rho_v, c->rho_v, g.nv);
  cv(w.kap, c->kap, g.np); cv(w.zeta, c->zeta, g.nw); cv(w.eta, c->eta, g.nw);
  cv(w.uadv, c->uadv, g.nu); cv(w.vadv, c->vadv, g.nv);
  w.penu = w.penv = nullptr;
  w.mdu = w.mdv = nullptr;
  if (pen) {
    if (!c->f_penu) { c->f_penu = dalloc<float>(c, g.nu); c->f_penv = dalloc<float>(c, g.nv); }
    w.penu = c->f_penu; w.penv = c->f_penv;
    cv(w.penu, c->penu, g.nu); cv(w.penv, c->penv, g.nv);
  }
  if (cons) {
    if (!c->f_mdu) { c->f_mdu = dalloc<float>(c, g.nu); c->f_mdv = dalloc<float>(c, g.nv); }
    w.mdu = c->f_mdu; w.mdv = c->f_mdv;
    cv(w.mdu, c->mdu, g.nu); cv(w.mdv, c->mdv, g.nv);
  }
}

// ---------------------------------------------------------------------------------
// PCG pressure solve.
template <typename T>
void pcg_setup(fugu_ctx* c, Work<T>& w) {
  LAUNCH((k_pcg_coef<T>), grid_union(c->g), dim3(TX, TY), c->s, c->g, w.rho_u, w.rho_v, w.cx, w.cy, w.dinv);
}

template <typename T>
void pcg(fugu_ctx* c, Work<T>& w, T* phi, const T* rhs, bool phi_zero, double tol, int max_iter,
         bool force_proj = false) {
  const Grid& g = c->g;
  // zero-mean projection of the rhs: both directions walls (the reference), or forced
  const int walls_both = (force_proj || (!g.per_x && !g.per_y)) ? 1 : 0;
  const dim3 gc = grid_cells_red(g), bl(TX, TY), b1(NT);
  const int nb = (int)nblocks(gc);
  const Sink sk = sink(c, cell_terms(g));
  const double* sc = c->sc;
  const int* fl = c->fl;
  c->pcg_calls++;
  if (walls_both) cell_mean<T>(c, rhs, (int)S_P_MEAN, 1);
  LAUNCH((k_pcg_init<T>), gc, bl, c->s, g, rhs, phi, w.cx, w.cy, w.dinv, w.pr, w.pz, w.pp0, walls_both,
         phi_zero ? 1 : 0, sk, sc, fl);
  seq_reduce(c, sk, 2, Skip{F_BRK, 1});
  LAUNCH(k_fold_pcg_init, dim3(1), b1, c->s, nb, sk, c->sc, c->fl);
  T* P[2] = {w.pp0, w.pp1};
  for (int it = 1; it <= max_iter; ++it) {
    // z, p_old, c_x, c_y read; p_new, Ap written
    TLAUNCH(c, sizeof(T) == 8 ? "pcg_iter1_f64" : "pcg_iter1_f32", 6.0 * sizeof(T) * g.nx * g.ny, (k_pcg_iter1<T>), gc, bl, c->s, g, it,
            w.pz, P[(it - 1) & 1], P[it & 1], w.Ap, w.cx, w.cy, sk, sc, fl);
    seq_reduce(c, sk, 1, Skip{F_P_DONE, 1});
    LAUNCH(k_fold_pcg_iter1, dim3(1), b1, c->s, nb, it, sk, c->sc, c->fl);
    // phi, p, r, Ap, D^-1 read; phi, r, z written
    TLAUNCH(c, sizeof(T) == 8 ? "pcg_iter2_f64" : "pcg_iter2_f32", 8.0 * sizeof(T) * g.nx * g.ny, (k_pcg_iter2<T>), gc, bl, c->s, g, phi,
            P[it & 1], w.pr, w.Ap, w.pz, w.dinv, sk, sc, fl);
    seq_reduce(c, sk, 2, Skip{F_P_DONE, 1});
    LAUNCH(k_fold_pcg_iter2, dim3(1), b1, c->s, nb, it, max_iter, sk, c->sc, c->fl);
  }
  (void)tol;
}

// ---------------------------------------------------------------------------------
// Pressure multigrid (chosen fast forms).
inline dim3 g2(int nx, int ny, int bx = 64, int by = 4) {
  return dim3((unsigned)((nx + bx - 1) / bx), (unsigned)((ny + by - 1) / by));
}
// grid of a fused level launch with tiles of 64 x th cells
inline dim3 mgt_grid(const MgDims& d, int th) {
  return dim3((unsigned)((d.nx + MGT_W - 1) / MGT_W), (unsigned)((d.ny + th - 1) / th));
}

// Levels of at most MG_TAIL_BLOCK_CELLS cells (the only ones whose tail can fit the LDS pool)
// keep c_x, c_y (, m), D^-1, b and phi back to back in one allocation, level after level, in
// the coarse tail's LDS pool order: staging a tail is then one contiguous copy.
constexpr size_t MG_TAIL_BLOCK_CELLS = 8192;
template <typename T>
std::vector<MgLevel<T>> mg_levels(fugu_ctx* c, MgDims d, bool mass, int max_levels) {
  // the fused level kernels address a level's arrays with 32-bit byte offsets and 24-bit
  // index products
  if ((size_t)(d.nx + 1) * (size_t)(d.ny + 1) * sizeof(T) >= ((size_t)1 << 32) || d.nx + 1 >= (1 << 24) ||
      d.ny + 1 >= (1 << 24))
    throw Error("multigrid: a level's arrays must stay below 4 GiB (" + std::to_string(d.nx) + " x " +
                std::to_string(d.ny) + " cells)");
  std::vector<MgDims> ds;
  for (;;) {
    ds.push_back(d);
    if (!mg_can_coarsen(d) || (max_levels > 0 && (int)ds.size() >= max_levels)) break;
    d = mg_coarse_dims(d);
  }
  size_t first_blk = ds.size(), blk_n = 0;
  for (size_t q = 0; q < ds.size(); ++q) {
    const size_t n = (size_t)ds[q].nx * ds[q].ny;
    if (n > MG_TAIL_BLOCK_CELLS) continue;
    first_blk = std::min(first_blk, q);
    blk_n += (size_t)(ds[q].nx + 1) * ds[q].ny + (size_t)ds[q].nx * (ds[q].ny + 1) + (mass ? 4 : 3) * n;
  }
  T* blk = blk_n ? dalloc<T>(c, blk_n) : nullptr;
  std::vector<MgLevel<T>> lv;
  for (size_t q = 0; q < ds.size(); ++q) {
    MgLevel<T> L;
    L.d = ds[q];
    const size_t n = (size_t)L.d.nx * L.d.ny;
    const size_t ncx = (size_t)(L.d.nx + 1) * L.d.ny, ncy = (size_t)L.d.nx * (L.d.ny + 1);
    auto take = [&](size_t k) -> T* {
      if (q < first_blk) return dalloc<T>(c, k);
      T* p = blk;
      blk += k;
      return p;
    };
    L.cx = take(ncx);
    L.cy = take(ncy);
    if (mass) L.m = take(n);
    L.dinv = take(n);
    L.b = take(n);
    L.phi = take(n);
    L.phi2 = dalloc<T>(c, n);
    if (q + 1 == ds.size() && n <= (size_t)MG_DENSE_MAX) L.cinv = dalloc<T>(c, n * n);
    lv.push_back(L);
  }
  return lv;
}

template <typename T>
size_t mg_tail_start(const fugu_ctx* c, const std::vector<MgLevel<T>>& lv, size_t nl);
template <typename T>
bool mg_tail_quad_ok(const fugu_ctx* c, const std::vector<MgLevel<T>>& lv, size_t l, size_t nl);
template <typename T>
void mg_tail_pack(fugu_ctx* c, std::vector<MgLevel<T>>& lv, size_t l, size_t nl);

// Coarse operators and inverse diagonals below a level 0 whose c_x, c_y (and m) are set; then
// the register images of a register-resident coarse tail.
template <typename T>
void mg_galerkin(fugu_ctx* c, std::vector<MgLevel<T>>& lv, size_t nl) {
  const dim3 bl(64, 4);
  for (auto& L : lv) L.tq_ok = false;   // the tails' images are rebuilt from the new operators
  for (size_t l = 0; l + 1 < nl; ++l) {
    const MgDims& f = lv[l].d;
    const MgDims& C = lv[l + 1].d;
    LAUNCH((k_mg_coarsen<T>), g2(C.nx + 1, C.ny + 1), bl, c->s, f, (const T*)lv[l].cx, (const T*)lv[l].cy,
           lv[l + 1].cx, lv[l + 1].cy);
    if (lv[l].m) LAUNCH((k_mg_coarsen_m<T>), g2(C.nx, C.ny), bl, c->s, f, (const T*)lv[l].m, lv[l + 1].m);
  }
  for (size_t l = 0; l < nl; ++l) {
    MgLevel<T>& L = lv[l];
    LAUNCH((k_mg_dinv<T>), g2(L.d.nx, L.d.ny), bl, c->s, L.d, (const T*)L.cx, (const T*)L.cy, (const T*)L.m,
           L.dinv);
  }
  const MgLevel<T>& C = lv[nl - 1];
  if (C.cinv)   // singular (pressure) block: no mass term
    LAUNCH((k_mg_coarse_inv<T>), dim3(1), dim3(1024), c->s, C.d, (const T*)C.cx, (const T*)C.cy, (const T*)C.m,
           C.m ? 0 : 1, C.cinv);
  const size_t ts = mg_tail_start(c, lv, nl);
  if (ts < nl && mg_tail_quad_ok(c, lv, ts, nl)) mg_tail_pack(c, lv, ts, nl);
}

template <typename T>
void mg_setup(fugu_ctx* c, Work<T>& w) {
  const Grid& g = c->g;
  if (mg_odd_periodic(w.mg[0].d))
    throw Error("multigrid preconditioner: a periodic direction needs an even cell count");
  LAUNCH((k_mg_coef0<T>), g2(g.nx + 1, g.ny + 1), dim3(64, 4), c->s, g, w.rho_u, w.rho_v, w.mg[0].cx, w.mg[0].cy);
  mg_galerkin(c, w.mg, w.mg.size());
}

// Velocity-block hierarchies from the current properties and dt (allocated on first use).
// Depth: a coarse level's diffusion-to-mass ratio is a quarter of its parent's (conductances
// c_H ~ c, mass m_H = 4m), so once the finest ratio r0 = max (sum c) / m has dropped below
// vmg_theta the level is mass-dominated and a few red-black sweeps solve it well enough: the
// hierarchy is cut at 1 + ceil(log4(r0 / vmg_theta)) levels (one host read per set-up, so the
// depth is a function of the current properties alone and recomputed steps repeat bit for bit).
template <typename T>
void vmg_setup(fugu_ctx* c, Work<T>& w, T dt) {
  const Grid& g = c->g;
  for (int comp = 0; comp < 2; ++comp) {
    if (mg_odd_periodic(vmg_dims(g, comp)))
      throw Error("velocity multigrid: a periodic direction needs an even cell count");
    auto& lv = w.vmg[comp];
    if (lv.empty()) lv = mg_levels<T>(c, vmg_dims(g, comp), true, c->vmg_levels);
    MgLevel<T>& L = lv[0];
    LAUNCH((k_vmg_coef0<T>), g2(L.d.nx + 1, L.d.ny + 1), dim3(64, 4), c->s, g, L.d, comp,
           comp == 0 ? (const T*)w.rho_u : (const T*)w.rho_v, (const T*)w.kap, (const T*)w.zeta, (const T*)w.eta,
           dt, comp == 0 ? (const T*)w.penu : (const T*)w.penv, L.cx, L.cy, L.m);
  }
  double r0[2] = {1.0e300, 1.0e300};
  if (c->vmg_theta > 0.0 || c->vmg_cycles <= 0) {
    const MgLevel<T>& U = w.vmg[0][0];
    const MgLevel<T>& V = w.vmg[1][0];
    const unsigned nb = std::min<unsigned>(nblk(std::max((size_t)U.d.nx * U.d.ny, (size_t)V.d.nx * V.d.ny)),
                                           (unsigned)RED_MAXB);
    LAUNCH((k_vmg_ratio<T>), dim3(nb), dim3(256), c->s, U.d, (const T*)U.cx, (const T*)U.cy, (const T*)U.m, V.d,
           (const T*)V.cx, (const T*)V.cy, (const T*)V.m, c->partials);
    LAUNCH(k_fold_absmax, dim3(1), dim3(NT), c->s, (int)nb, (const double*)c->partials, c->sc);
    read_slots(c);
    r0[0] = c->h_sc[S_DOT0];
    r0[1] = c->h_sc[S_DOT1];
  }
  for (int comp = 0; comp < 2; ++comp) {
    auto& lv = w.vmg[comp];
    w.vmg_cyc[comp] = c->vmg_cycles > 0 ? c->vmg_cycles : (r0[comp] > c->vmg_cycle_ratio ? 2 : 1);
    const double theta = (w.vmg_cyc[comp] >= 2 && c->vmg_theta > 0.0) ? c->vmg_theta2 : c->vmg_theta;
    size_t nl = 1;
    for (double r = r0[comp]; r > theta && nl < lv.size(); r *= 0.25) ++nl;
    w.vmg_nl[comp] = nl;
    if (c->debug)
      std::fprintf(stderr, "[fugu] velocity block %d: ratio %.3g, %zu of %zu levels, %d cycle(s)\n", comp, r0[comp], nl,
                   lv.size(), w.vmg_cyc[comp]);
    mg_galerkin(c, lv, nl);
  }
}

template <typename T>
void mg_smooth(fugu_ctx* c, MgLevel<T>& L, int colour, int zero_other) {
  const int half = (L.d.nx + 1) / 2;
  // one colour: b, D^-1 and phi (written) of half the cells, phi of the other half (neighbours),
  // all of c_x and c_y
  TLAUNCH(c, sizeof(T) == 8 ? "mg_smooth_f64" : "mg_smooth_f32", 4.0 * sizeof(T) * L.d.nx * L.d.ny, (k_mg_rbgs<T>), g2(half, L.d.ny), dim3(64, 4),
          c->s, L.d, colour, (const T*)L.cx, (const T*)L.cy, (const T*)L.dinv, (const T*)L.b, L.phi, zero_other);
}

// LDS elements the tail needs from level l down to nl-1 (coefficients, mass, D^-1, rhs, iterate).
template <typename T>
size_t mg_tail_footprint(const std::vector<MgLevel<T>>& lv, size_t l, size_t nl) {
  size_t e = 0;
  for (size_t q = l; q < nl; ++q) {
    const size_t nx = lv[q].d.nx, ny = lv[q].d.ny;
    e += (nx + 1) * ny + nx * (ny + 1) + (lv[q].m ? 4 : 3) * nx * ny;
  }
  return e;
}

// First level of the coarse tail (lv.size(): none).  mg_tail_cells < 0 (default): the highest
// level of at most MG_TAIL_AUTO_CELLS cells whose tail fits the LDS pool (a single workgroup is
// slower than the fused level kernels above that size); > 0: the first level of at most that
// many cells; 0: no tail.  At most MG_TAIL_LEVELS levels, and the coarsest level must fit the
// LDS solve.
constexpr int MG_TAIL_AUTO_CELLS = 2048;
template <typename T>
size_t mg_tail_start(const fugu_ctx* c, const std::vector<MgLevel<T>>& lv, size_t nl) {
  const MgDims& dc = lv[nl - 1].d;
  if (c->mg_tail_cells == 0 || (size_t)dc.nx * dc.ny > (size_t)MG_COARSE_MAX) return nl;
  size_t ts = 0;
  if (c->mg_tail_cells > 0) {
    while (ts < nl && (size_t)lv[ts].d.nx * lv[ts].d.ny > (size_t)c->mg_tail_cells) ++ts;
  } else {
    while (ts + 1 < nl && (mg_tail_footprint(lv, ts, nl) * sizeof(T) > (size_t)MG_TAIL_POOL ||
                           (size_t)lv[ts].d.nx * lv[ts].d.ny > (size_t)MG_TAIL_AUTO_CELLS))
      ++ts;
  }
  if (ts + MG_TAIL_LEVELS < nl) ts = nl - MG_TAIL_LEVELS;
  return ts;
}

// Timing names of the fused level kernels by sweep count (two: the plain name).
inline const char* mg_kname(bool post, int nu, bool f64) {
  static const char* const names[2][2][5] = {
      {{"", "mg_pre1_f32", "mg_pre_f32", "mg_pre3_f32", "mg_pre4_f32"},
       {"", "mg_pre1_f64", "mg_pre_f64", "mg_pre3_f64", "mg_pre4_f64"}},
      {{"", "mg_post1_f32", "mg_post_f32", "mg_post3_f32", "mg_post4_f32"},
       {"", "mg_post1_f64", "mg_post_f64", "mg_post3_f64", "mg_post4_f64"}}};
  return names[post ? 1 : 0][f64 ? 1 : 0][(nu >= 1 && nu <= 4) ? nu : 2];
}

// Register-strip level kernels (fugu2d_mgs.hpp) on the large and on the small levels, the
// LDS-tile kernels (fugu2d_mg.hpp) in between.  Measured against the tile kernels (same results
// bit for bit):
//  - large (at least 2^21 cells, or 2^20 with three or more sweeps): 8-row strips (FP64 two-sweep
//    levels: 4 rows); at 2048^2 the FP64 three-sweep velocity levels -25 % (pre 76 -> 57 us, post
//    72 -> 54 us), FP32 -7 to -27 %;
//  - small (at most 2^17 cells): 2-row strips (4 rows for the pre-smoothing with three or more
//    sweeps, whose halo needs a taller tile), each wavefront's chain of instructions a quarter
//    as long: -24 to -41 % for the two-sweep pressure levels, -15 to -35 % for the velocity
//    levels (64 x 128 to 256 x 512);
//  - in between (the bubble's 512 x 1024 levels), the tile kernels are faster.
enum class MgsShape { none, large, small };
inline MgsShape mg_strip_shape(const fugu_ctx* c, const MgDims& d, int nu) {
  if (c->mg_strip == 0) return MgsShape::none;
  if (c->mg_strip == 2) return MgsShape::large;   // every fused level (tests)
  if (c->mg_strip == 3) return MgsShape::small;
  const double n = (double)d.nx * d.ny;
  const int lg = c->mgs_large_log2;
  if (n >= std::ldexp(1.0, lg) || (n >= std::ldexp(1.0, lg - 1) && nu >= 3)) return MgsShape::large;
  if (n <= std::ldexp(1.0, c->mgs_small_log2)) return MgsShape::small;
  return MgsShape::none;
}
inline const char* mgs_kname(bool post, int nu, bool f64) {
  static const char* const names[2][2][5] = {
      {{"", "mgs_pre1_f32", "mgs_pre_f32", "mgs_pre3_f32", "mgs_pre4_f32"},
       {"", "mgs_pre1_f64", "mgs_pre_f64", "mgs_pre3_f64", "mgs_pre4_f64"}},
      {{"", "mgs_post1_f32", "mgs_post_f32", "mgs_post3_f32", "mgs_post4_f32"},
       {"", "mgs_post1_f64", "mgs_post_f64", "mgs_post3_f64", "mgs_post4_f64"}}};
  return names[post ? 1 : 0][f64 ? 1 : 0][(nu >= 1 && nu <= 4) ? nu : 2];
}
// Strip shape (rows per wavefront, wavefronts) of a launch: one table for the dispatch and for
// the partials count.
struct MgsCfg {
  int R, NW;
};
inline MgsCfg mgs_cfg(bool post, int nu, MgsShape sh, bool f64) {
  if (sh == MgsShape::small) {
    if (post) return nu == 4 ? MgsCfg{4, 8} : MgsCfg{2, 8};
    return nu >= 3 ? MgsCfg{4, 8} : MgsCfg{2, 8};
  }
  if (nu == 2 && f64) return post ? MgsCfg{4, 8} : MgsCfg{4, 16};
  return MgsCfg{8, 8};
}
// Workgroups of a level's fused post-smoothing launch (one double-double partial each).
inline size_t mg_post_groups(const fugu_ctx* c, const MgDims& d, int nu, bool f64) {
  const MgsShape sh = c->mg_fused ? mg_strip_shape(c, d, nu) : MgsShape::none;
  if (sh == MgsShape::none) {
    const dim3 g = mgt_grid(d, MGT_TH);
    return (size_t)g.x * g.y;
  }
  const MgsCfg k = mgs_cfg(true, nu, sh, f64);
  const int hp = 2 * nu, tw = MGS_EW - 2 * hp, th = k.NW * k.R - 2 * hp;
  return (size_t)((d.nx + tw - 1) / tw) * (size_t)((d.ny + th - 1) / th);
}
template <typename T, int NU, int R, int NW>
void mgs_pre_launch(fugu_ctx* c, const char* nm, double ab, const MgLevel<T>& L, T* out, T* bc, T* phic,
                    const T* g_in) {
  const dim3 g = mgs_grid_pre<NU, R, NW>(L.d);
  if (L.m)
    TLAUNCH(c, nm, ab, (k_mgs_pre<T, NU, R, NW, true>), g, dim3(64 * NW), c->s, L.d, (const T*)L.cx, (const T*)L.cy,
            (const T*)L.m, (const T*)L.b, out, bc, phic, g_in);
  else
    TLAUNCH(c, nm, ab, (k_mgs_pre<T, NU, R, NW, false>), g, dim3(64 * NW), c->s, L.d, (const T*)L.cx, (const T*)L.cy,
            (const T*)nullptr, (const T*)L.b, out, bc, phic, g_in);
}
// Pre-smoothing (+ restriction when bc is set) by a strip kernel of the given shape.
template <typename T>
void mgs_pre(fugu_ctx* c, const char* nm, double ab, const MgLevel<T>& L, int nu, MgsShape sh, T* out, T* bc,
             T* phic, const T* g_in) {
  const MgsCfg k = mgs_cfg(false, nu, sh, sizeof(T) == 8);
#define FUGU_MGS_PRE(NUV, RV, NWV)                                            \
  if (nu == NUV && k.R == RV && k.NW == NWV) {                                \
    mgs_pre_launch<T, NUV, RV, NWV>(c, nm, ab, L, out, bc, phic, g_in);       \
    return;                                                                   \
  }
  FUGU_MGS_PRE(1, 2, 8) FUGU_MGS_PRE(2, 2, 8) FUGU_MGS_PRE(3, 4, 8) FUGU_MGS_PRE(4, 4, 8)
  FUGU_MGS_PRE(1, 8, 8) FUGU_MGS_PRE(2, 8, 8) FUGU_MGS_PRE(3, 8, 8) FUGU_MGS_PRE(4, 8, 8) FUGU_MGS_PRE(2, 4, 16)
#undef FUGU_MGS_PRE
  throw Error("multigrid: no strip kernel for this shape");
}
template <typename T, int NU, int R, int NW>
int mgs_post_launch(fugu_ctx* c, const char* nm, double ab, const MgLevel<T>& L, const T* phi, const T* phic,
                    T* out, double* psum) {
  const dim3 g = mgs_grid_post<NU, R, NW>(L.d);
  if (L.m)
    TLAUNCH(c, nm, ab, (k_mgs_post<T, NU, R, NW, true>), g, dim3(64 * NW), c->s, L.d, (const T*)L.cx,
            (const T*)L.cy, (const T*)L.m, (const T*)L.b, phi, phic, out, psum);
  else
    TLAUNCH(c, nm, ab, (k_mgs_post<T, NU, R, NW, false>), g, dim3(64 * NW), c->s, L.d, (const T*)L.cx,
            (const T*)L.cy, (const T*)nullptr, (const T*)L.b, phi, phic, out, psum);
  return (int)(g.x * g.y);
}
// Prolongation + post-smoothing by a strip kernel of the given shape; returns its workgroup
// count (the psum partials).
template <typename T>
int mgs_post(fugu_ctx* c, const char* nm, double ab, const MgLevel<T>& L, int nu, MgsShape sh, const T* phi,
             const T* phic, T* out, double* psum) {
  const MgsCfg k = mgs_cfg(true, nu, sh, sizeof(T) == 8);
#define FUGU_MGS_POST(NUV, RV, NWV) \
  if (nu == NUV && k.R == RV && k.NW == NWV) return mgs_post_launch<T, NUV, RV, NWV>(c, nm, ab, L, phi, phic, out, psum);
  FUGU_MGS_POST(1, 2, 8) FUGU_MGS_POST(2, 2, 8) FUGU_MGS_POST(3, 2, 8) FUGU_MGS_POST(4, 4, 8)
  FUGU_MGS_POST(1, 8, 8) FUGU_MGS_POST(2, 8, 8) FUGU_MGS_POST(3, 8, 8) FUGU_MGS_POST(4, 8, 8) FUGU_MGS_POST(2, 4, 8)
#undef FUGU_MGS_POST
  throw Error("multigrid: no strip kernel for this shape");
}

// V-cycle on level l; `zero_guess`: phi of level l is (treated as) zero on entry; `project`:
// the operator is singular (no mass term, no Dirichlet face) and the coarsest rhs is centred.
// nl: the levels in use (a velocity hierarchy may be cut short, its last one mass-dominated).
// Returns the buffer holding the level's result: phi2 after a fused post-smoothing (whose
// tiles cannot update in place), phi otherwise -- a function of the path alone, so the same
// launches always touch the same buffers (a captured cycle replays correctly).
// psum (level 0 only): where a fused post-smoothing writes the result, its per-tile double-double
// sums go there and *psum_nb = their count (left 0 when no such launch ran or they do not fit).
// guess (not zero_guess): the buffer holding the initial iterate (nullptr: phi).  The fused
// kernels read it where it lies and write the other buffer; the in-place paths get it copied
// into phi first.
// The coarse tail's level arguments (levels l .. nl-1) with the LDS-pool offsets of the
// staged general tail; nu, coarse_sweeps, project and zero_guess are the caller's.
template <typename T>
MgTailArgs<T> mg_tail_args(const fugu_ctx* c, const std::vector<MgLevel<T>>& lv, size_t l, size_t nl) {
  MgTailArgs<T> ta{};
  ta.nlev = (int)(nl - l);
  int off = 0;
  for (size_t q = l; q < nl; ++q) {
    MgTailLevel<T> t{lv[q].d, lv[q].cx, lv[q].cy, lv[q].m, lv[q].dinv, lv[q].cinv, lv[q].b, lv[q].phi};
    const int nx = lv[q].d.nx, ny = lv[q].d.ny, n = nx * ny;
    t.ocx = off; off += (nx + 1) * ny;
    t.ocy = off; off += nx * (ny + 1);
    t.om = lv[q].m ? off : -1; off += lv[q].m ? n : 0;
    t.odinv = off; off += n;
    t.ob = off; off += n;
    t.ophi = off; off += n;
    ta.L[q - l] = t;
  }
  ta.blk = lv[l].cx;
  ta.nblk = off;
  return ta;
}
// The register-resident tail applies: even smoothing levels whose quads fit the workgroup, at
// most MG_TAILQ_LEVELS of them, b and phi of every level in its LDS pool, the coarsest
// operator's image in its LDS.
template <typename T>
bool mg_tail_quad_ok(const fugu_ctx* c, const std::vector<MgLevel<T>>& lv, size_t l, size_t nl) {
  const int nsm = (int)(nl - l) - 1;
  bool quad = c->mg_tail_quad && nsm >= 1 && nsm <= MG_TAILQ_LEVELS;
  size_t qpool = 0;
  for (size_t q = l; q < nl; ++q) {
    const MgDims& d = lv[q].d;
    qpool += 2 * (size_t)d.nx * d.ny;
    if (q + 1 < nl) quad = quad && (d.nx % 2 == 0) && (d.ny % 2 == 0) && (d.nx / 2) * (d.ny / 2) <= MG_TAILQ_NTH;
  }
  return quad && qpool <= (size_t)MG_TAILQ_POOL &&
         mg_tailq_coarse_elems(lv[nl - 1].d, lv[nl - 1].cinv != nullptr) <= (size_t)MG_TAILQ_COARSE;
}
// The register images of the tail starting at level l (k_mg_tail_q_pack), after a set-up of the
// levels' operators.
template <typename T>
void mg_tail_pack(fugu_ctx* c, std::vector<MgLevel<T>>& lv, size_t l, size_t nl) {
  MgLevel<T>& Lt = lv[l];
  if (!Lt.tq) {
    Lt.tq = dalloc<TqQuad<T>>(c, (size_t)MG_TAILQ_LEVELS * MG_TAILQ_NTH);
    Lt.tqc = dalloc<T>(c, MG_TAILQ_COARSE);
  }
  const MgTailArgs<T> ta = mg_tail_args(c, lv, l, nl);
  const int nsm = ta.nlev - 1;
  if (nsm == 1) LAUNCH((k_mg_tail_q_pack<T, 1>), dim3(1), dim3(MG_TAILQ_NTH), c->s, ta, Lt.tq, Lt.tqc);
  else if (nsm == 2) LAUNCH((k_mg_tail_q_pack<T, 2>), dim3(1), dim3(MG_TAILQ_NTH), c->s, ta, Lt.tq, Lt.tqc);
  else LAUNCH((k_mg_tail_q_pack<T, 3>), dim3(1), dim3(MG_TAILQ_NTH), c->s, ta, Lt.tq, Lt.tqc);
  Lt.tq_ok = true;
  Lt.tq_nl = nl;
}

template <typename T>
T* mg_vcycle(fugu_ctx* c, std::vector<MgLevel<T>>& lv, size_t l, bool zero_guess, int project, size_t nl,
             double* psum = nullptr, int* psum_nb = nullptr, const T* guess = nullptr) {
  MgLevel<T>& L = lv[l];
  const int nu = L.m ? c->mg_nu_v : c->mg_nu_p;   // red-black sweeps before and after the coarse correction
  const T* g_in = zero_guess ? nullptr : (guess ? guess : L.phi);
  auto guess_to_phi = [&]() {
    if (g_in && g_in != L.phi)
      HIPCHK(hipMemcpyAsync(L.phi, g_in, (size_t)L.d.nx * L.d.ny * sizeof(T), hipMemcpyDeviceToDevice, c->s));
  };
  if (l == mg_tail_start(c, lv, nl)) {   // the coarse tail in one launch
    guess_to_phi();
    MgTailArgs<T> ta = mg_tail_args(c, lv, l, nl);
    // a cut velocity hierarchy's coarsest level gets the cut treatment's sweeps, whatever its size
    const int cnu = (c->vmg_coarse_nu == 3 || c->vmg_coarse_nu == 4) ? c->vmg_coarse_nu : 2;
    ta.nu = nu; ta.coarse_sweeps = (nl < lv.size() && lv[nl - 1].m) ? cnu : 60;
    ta.project = project; ta.zero_guess = zero_guess ? 1 : 0;
    // staged: the levels' arrays must lie in the pool order from lv[l].cx on (the phi of the
    // levels below l excepted: the tail writes them before it reads them)
    bool packed = true;
    for (size_t q = l; q < nl && packed; ++q) {
      const MgTailLevel<T>& t = ta.L[q - l];
      const T* base = lv[l].cx;
      packed = lv[q].cx == base + t.ocx && lv[q].cy == base + t.ocy && (t.om < 0 || lv[q].m == base + t.om) &&
               lv[q].dinv == base + t.odinv && lv[q].b == base + t.ob;
    }
    const int off = ta.nblk;
    const int nsm = ta.nlev - 1;
    if (mg_tail_quad_ok(c, lv, l, nl)) {
      MgLevel<T>& Lt = lv[l];
      if (!Lt.tq_ok || Lt.tq_nl != nl) mg_tail_pack(c, lv, l, nl);   // (the set-up normally did it)
      ta.tq = Lt.tq;
      ta.tqc = Lt.tqc;
      if (nsm == 1) LAUNCH((k_mg_tail_q<T, 1>), dim3(1), dim3(MG_TAILQ_NTH), c->s, ta);
      else if (nsm == 2) LAUNCH((k_mg_tail_q<T, 2>), dim3(1), dim3(MG_TAILQ_NTH), c->s, ta);
      else LAUNCH((k_mg_tail_q<T, 3>), dim3(1), dim3(MG_TAILQ_NTH), c->s, ta);
    } else if (c->mg_tail_stage && packed && (size_t)off * sizeof(T) <= (size_t)MG_TAIL_POOL)
      LAUNCH((k_mg_tail<T, true, MG_TAIL_NTH>), dim3(1), dim3(MG_TAIL_NTH), c->s, ta);
    else
      LAUNCH((k_mg_tail<T, false, MG_TAIL_NTH>), dim3(1), dim3(MG_TAIL_NTH), c->s, ta);
    return L.phi;
  }
  if (l + 1 == nl) {
    if (L.m && ((size_t)L.d.nx * L.d.ny > (size_t)MG_COARSE_MAX || nl < lv.size())) {
      // truncated velocity hierarchy: the mass term dominates its coarsest level, which gets
      // vmg_coarse_nu red-black sweeps (one fused launch)
      const int cnu = c->vmg_coarse_nu;
      if (c->mg_fused) {
        const dim3 gt = mgt_grid(L.d, MGT_TH);
        const double ab = (g_in ? 6.0 : 5.0) * sizeof(T) * L.d.nx * L.d.ny;
        const char* nm = sizeof(T) == 8 ? "mg_cut_f64" : "mg_cut_f32";
        T* out = (g_in == L.phi) ? L.phi2 : L.phi;
        const int snu = (cnu == 3 || cnu == 4) ? cnu : 2;
        const MgsShape sh = mg_strip_shape(c, L.d, snu);
        if (sh != MgsShape::none) {
          mgs_pre<T>(c, sizeof(T) == 8 ? "mgs_cut_f64" : "mgs_cut_f32", ab, L, snu, sh, out, (T*)nullptr,
                     (T*)nullptr, g_in);
          return out;
        }
#define FUGU_MG_CUT(NUV)                                                                                     \
  TLAUNCH(c, nm, ab, (k_mg_pre<T, NUV, MGT_TH>), gt, dim3(16 * MGT_TH), c->s, L.d, (const T*)L.cx,             \
          (const T*)L.cy, (const T*)L.m, (const T*)L.b, out, (T*)nullptr, (T*)nullptr, g_in)
        if (cnu == 3) FUGU_MG_CUT(3);
        else if (cnu == 4) FUGU_MG_CUT(4);
        else FUGU_MG_CUT(2);
#undef FUGU_MG_CUT
        return out;
      } else {
        guess_to_phi();
        for (int s = 0; s < (cnu == 3 || cnu == 4 ? cnu : 2); ++s) {
          mg_smooth(c, L, 0, (s == 0 && zero_guess) ? 1 : 0);
          mg_smooth(c, L, 1, 0);
        }
      }
    } else if ((size_t)L.d.nx * L.d.ny <= (size_t)MG_COARSE_MAX) {
      LAUNCH((k_mg_coarse<T>), dim3(1), dim3(MG_COARSE_MAX), c->s, L.d, (const T*)L.cx, (const T*)L.cy,
             (const T*)L.dinv, (const T*)L.cinv, (const T*)L.b, L.phi, 60, project);
    } else {
      guess_to_phi();
      for (int s = 0; s < 40; ++s) {
        mg_smooth(c, L, 0, (s == 0 && zero_guess) ? 1 : 0);
        mg_smooth(c, L, 1, 0);
      }
    }
    return L.phi;
  }
  MgLevel<T>& C = lv[l + 1];
  const double ncell = (double)L.d.nx * L.d.ny;
  // FP64 levels of at least 2^21 cells: 64-row tiles (1024 threads) for the pre-smoothing with 3
  // sweeps.  Its 162-VGPR 32-row form fits one workgroup per CU; the 64-row form holds 6 cells
  // per thread instead of 8, fits 128 VGPRs (four waves per SIMD) and has 1.5x instead of 1.75x
  // halo overhead: 18 % faster at 2048^2, bit-identical (the tile shape changes no arithmetic),
  // at 1.19x its algorithmic HBM bytes instead of 1.07x.  With 2 sweeps the 32-row form fits
  // 124 VGPRs, two workgroups per CU, and is the faster one (44 against 57 us at 2048^2).  The
  // post-smoothing keeps the 32-row tiles: as fast, and 1.06x its bytes instead of 1.66x (its
  // larger grown tiles no longer share their halos in an XCD's L2).  Smaller levels keep more
  // workgroups than CUs.
  const bool th64 = sizeof(T) == 8 && ncell >= (double)(1 << 21);
  const bool pre64 = th64 && nu == 3;
  const dim3 gt = mgt_grid(L.d, pre64 ? 64 : MGT_TH);
  T* P = (g_in == L.phi) ? L.phi2 : L.phi;   // the fused pre-smoothing's output
  const MgsShape shape = c->mg_fused ? mg_strip_shape(c, L.d, nu) : MgsShape::none;
  if (shape != MgsShape::none) {
    const double ab = ((L.m ? 5.5 : 4.5) + (g_in ? 1.0 : 0.0)) * sizeof(T) * ncell;
    mgs_pre<T>(c, mgs_kname(false, nu, sizeof(T) == 8), ab, L, nu, shape, P, C.b, C.phi, g_in);
  } else if (c->mg_fused) {
    // b, c_x, c_y (, m) (, the initial iterate) read; phi written; coarse b and phi written (a
    // quarter each)
    const char* nm = mg_kname(false, nu, sizeof(T) == 8);
    const double ab = ((L.m ? 5.5 : 4.5) + (g_in ? 1.0 : 0.0)) * sizeof(T) * ncell;
#define FUGU_MG_PRE(NUV, THV)                                                                                \
  TLAUNCH(c, nm, ab, (k_mg_pre<T, NUV, THV>), gt, dim3(16 * THV), c->s, L.d, (const T*)L.cx,                   \
          (const T*)L.cy, (const T*)L.m, (const T*)L.b, P, C.b, C.phi, g_in)
    if constexpr (sizeof(T) == 8) {
      if (pre64) FUGU_MG_PRE(3, 64);
      else if (nu == 1) FUGU_MG_PRE(1, MGT_TH);
      else if (nu == 3) FUGU_MG_PRE(3, MGT_TH);
      else if (nu == 4) FUGU_MG_PRE(4, MGT_TH);
      else FUGU_MG_PRE(2, MGT_TH);
    } else {
      if (nu == 1) FUGU_MG_PRE(1, MGT_TH);
      else if (nu == 3) FUGU_MG_PRE(3, MGT_TH);
      else if (nu == 4) FUGU_MG_PRE(4, MGT_TH);
      else FUGU_MG_PRE(2, MGT_TH);
    }
#undef FUGU_MG_PRE
  } else {
    guess_to_phi();
    for (int s = 0; s < nu; ++s) {
      mg_smooth(c, L, 0, (s == 0 && zero_guess) ? 1 : 0);
      mg_smooth(c, L, 1, 0);
    }
    // fine c_x, c_y, (m,) b, phi read; coarse b, phi written (a quarter of the cells each)
    TLAUNCH(c, sizeof(T) == 8 ? "mg_resid_f64" : "mg_resid_f32", (L.m ? 5.5 : 4.5) * sizeof(T) * ncell,
            (k_mg_restrict<T>), g2(C.d.nx, C.d.ny), dim3(64, 4), c->s, L.d, (const T*)L.cx, (const T*)L.cy,
            (const T*)L.m, (const T*)L.b, (const T*)L.phi, C.b, C.phi);
  }
  const T* cres = mg_vcycle(c, lv, l + 1, true, project, nl);
  if (shape != MgsShape::none) {
    T* Q = (P == L.phi) ? L.phi2 : L.phi;
    const double pbytes = (L.m ? 6.25 : 5.25) * sizeof(T) * ncell;
    const int nb = mgs_post<T>(c, mgs_kname(true, nu, sizeof(T) == 8), pbytes, L, nu, shape, (const T*)P, cres, Q,
                               psum ? psum : (double*)nullptr);
    if (psum && psum_nb) *psum_nb = nb;
    return Q;
  }
  if (c->mg_fused) {
    // phi, coarse phi (a quarter), b, c_x, c_y (, m) read; the other buffer written
    T* Q = (P == L.phi) ? L.phi2 : L.phi;
    const dim3 gp = mgt_grid(L.d, MGT_TH);
    const char* pname = mg_kname(true, nu, sizeof(T) == 8);
    const double pbytes = (L.m ? 6.25 : 5.25) * sizeof(T) * ncell;
#define FUGU_MG_POST(NUV, THV)                                                                                      \
  TLAUNCH(c, pname, pbytes, (k_mg_post<T, NUV, THV>), gp, dim3(16 * THV), c->s, L.d, (const T*)L.cx, (const T*)L.cy, \
          (const T*)L.m, (const T*)L.b, (const T*)P, cres, Q, psum ? psum : (double*)nullptr)
    if (nu == 1) FUGU_MG_POST(1, MGT_TH);
    else if (nu == 3) FUGU_MG_POST(3, MGT_TH);
    else if (nu == 4) FUGU_MG_POST(4, MGT_TH);
    else FUGU_MG_POST(2, MGT_TH);
#undef FUGU_MG_POST
    if (psum && psum_nb) *psum_nb = (int)(gp.x * gp.y);
    return Q;
  }
  LAUNCH((k_mg_prolong<T>), g2(L.d.nx, L.d.ny), dim3(64, 4), c->s, L.d, cres, L.phi);
  for (int s = 0; s < nu; ++s) {
    mg_smooth(c, L, 1, 0);
    mg_smooth(c, L, 0, 0);
  }
  return L.phi;
}

// phi (padded, interior written) ~= A_p^-1 (z_p - mean) by mg_cycles V-cycles from zero.
// Leaves S_P_MEAN = mean(phi) (both directions walls) and F_P_DONE = 0 for k_pcd_rp.
template <typename T>
void mg_pressure(fugu_ctx* c, Work<T>& w, T* phi, const T* Zp, bool force_proj = false) {
  const Grid& g = c->g;
  const int walls_both = (force_proj || (!g.per_x && !g.per_y)) ? 1 : 0;
  c->pcg_calls++;
  if (walls_both) cell_mean<T>(c, Zp, (int)S_P_MEAN, 0);
  LAUNCH((k_mg_load<T>), g2(g.nx, g.ny), dim3(64, 4), c->s, g, Zp, w.mg[0].b, walls_both, (const double*)c->sc,
         (int)S_P_MEAN, c->fl, (int)F_P_DONE);
  MgLevel<T>& L0 = w.mg[0];
  const T* res = nullptr;
  for (int k = 0; k < c->mg_cycles; ++k)   // later cycles smooth from the last result
    res = mg_vcycle(c, w.mg, 0, k == 0, 1, w.mg.size(), nullptr, nullptr, res);
  LAUNCH((k_mg_store<T>), g2(g.nx, g.ny), dim3(64, 4), c->s, g, res, phi);
  if (walls_both) cell_mean<T>(c, (const T*)phi, (int)S_P_MEAN, 0);
}

// The PCD's pressure block in the chosen fast forms: returns the compact level-0 result
// phi ~= A_p^-1 (z_p - mean) (mg_cycles V-cycles from zero), and with walls in both directions
// leaves S_P_MEAN = mean(phi), from the last post-smoothing's per-tile sums where the fused
// kernel wrote the result (double-double sums: the same mean whichever kernel summed), and
// F_P_DONE = 0.  No padded copy: k_pcd_rp reads the compact array.
template <typename T>
const T* mg_pressure_compact(fugu_ctx* c, Work<T>& w, const T* Zp) {
  const Grid& g = c->g;
  const int walls_both = (!g.per_x && !g.per_y) ? 1 : 0;
  c->pcg_calls++;
  const int centre = walls_both && (c->zp_centre || !c->fgmres_pcd);
  if (centre) cell_mean<T>(c, Zp, (int)S_P_MEAN, 0);
  MgLevel<T>& L0 = w.mg[0];
  LAUNCH((k_mg_load<T>), g2(g.nx, g.ny), dim3(64, 4), c->s, g, Zp, L0.b, centre, (const double*)c->sc,
         (int)S_P_MEAN, c->fl, (int)F_P_DONE);
  // the level-0 post-smoothing's workgroups (one partial each), when it is a fused launch
  const size_t tiles = mg_post_groups(c, L0.d, c->mg_nu_p, sizeof(T) == 8);
  double* psum = (walls_both && 2 * tiles <= c->partials_n) ? c->partials : nullptr;
  const T* res = nullptr;
  int nb = 0;
  for (int k = 0; k < c->mg_cycles; ++k) {   // later cycles smooth from the last result
    nb = 0;
    res = mg_vcycle(c, w.mg, 0, k == 0, 1, w.mg.size(), k + 1 == c->mg_cycles ? psum : nullptr, &nb, res);
  }
  if (walls_both) {
    if (nb == 0) {
      nb = (int)std::min<unsigned>(nblk((size_t)L0.d.nx * L0.d.ny), (unsigned)RED_MAXB);
      LAUNCH((k_mg_sum<T>), dim3((unsigned)nb), dim3(256), c->s, L0.d, res, c->partials);
    }
    LAUNCH(k_fold_sum, dim3(1), dim3(NT), c->s, g, nb, (int)S_P_MEAN, 0, Sink{c->partials, 0, nullptr}, c->sc,
           (const int*)c->fl);
  }
  return res;
}

// out (padded u or v array) ~= A_comp^-1 rhs by one V-cycle from zero on the velocity block
// (rows of the block; wall faces copy rhs, ghosts zero, as the Jacobi blocks leave them).
// loaded: level 0's b already holds the rhs rows (written by the kernel that formed them).
// out == nullptr: no store (the caller fuses it); returns the compact result either way.
template <typename T>
const T* vmg_solve(fugu_ctx* c, Work<T>& w, int comp, const T* rhs, T* out, bool loaded = false) {
  const Grid& g = c->g;
  auto& lv = w.vmg[comp];
  MgLevel<T>& L = lv[0];
  if (!loaded) LAUNCH((k_vmg_load<T>), g2(L.d.nx, L.d.ny), dim3(64, 4), c->s, g, L.d, rhs, comp, L.b);
  const T* res = nullptr;
  for (int k = 0; k < w.vmg_cyc[comp]; ++k)   // later cycles smooth from the last result
    res = mg_vcycle(c, lv, 0, k == 0, 0, w.vmg_nl[comp], nullptr, nullptr, res);
  if (out) LAUNCH((k_vmg_store<T>), grid_union(g), dim3(TX, TY), c->s, g, comp, res, rhs, out);
  return res;
}

// block PCD preconditioner R = P^-1 Z (triplets).  bcs_after: the caller applies the homogeneous
// velocity BCs to R next; then (multigrid velocity blocks, fused glue) the u store applies them
// itself and the return value is true (the caller skips its own).
template <typename T>
bool pcd(fugu_ctx* c, Work<T>& w, const T* Z, T* R, T dt, bool bcs_after = false) {
  const Grid& g = c->g;
  const int walls_both = (!g.per_x && !g.per_y) ? 1 : 0;
  const T *Zu = Z, *Zv = Z + g.nu, *Zp = Z + g.nu + g.nv;
  T *Ru = R, *Rv = R + g.nu, *Rp = R + g.nu + g.nv;
  const dim3 gu = grid_union(g), gc = grid_cells(g), bl(TX, TY);
  const T* phc = nullptr;
  if (c->use_mg) {
    phc = mg_pressure_compact<T>(c, w, Zp);
  } else {
    pcg<T>(c, w, w.phi, Zp, true, 1.0e-5, 150);
    if (walls_both)   // mean of phi, unless the zero-rhs early exit fired (skip when F_P_DONE >= 2)
      cell_mean<T>(c, (const T*)w.phi, (int)S_P_MEAN, 2);
  }
  LAUNCH((k_pcd_rp<T>), gu, bl, c->s, g, w.phi, w.uadv, w.vadv, dt, walls_both, Rp, c->sc, c->fl, Zp, w.kap,
         (T)c->pcd_visc, phc);
  halo(c, Rp, 2);
  LAUNCH((k_pcd_zwork<T>), gu, bl, c->s, g, Zu, Zv, (const T*)Rp, w.zuw, w.zvw,
         c->use_vmg ? w.vmg[1][0].b : (T*)nullptr);
  // v block, then the u block's rhs with the cross term: store, halo and cross fused in one
  // launch when the velocity blocks are multigrid solves (pcd_fused_glue)
  const bool fuse_v = c->use_vmg && c->pcd_cross && c->pcd_fused_glue;
  if (c->use_vmg) {
    const T* resv = vmg_solve<T>(c, w, 1, w.zvw, fuse_v ? (T*)nullptr : Rv, true);
    if (fuse_v)
      LAUNCH((k_vmg_store_cross<T>), gu, bl, c->s, g, resv, (const T*)w.zvw, Rv, (const T*)w.eta, (const T*)w.zuw,
             w.vmg[0][0].b);
  } else {
    LAUNCH((k_jacobi_v<T>), gu, bl, c->s, g, (const T*)nullptr, w.jv0, w.zvw, w.rho_v, w.kap, w.zeta, w.eta, dt, w.penv);
    LAUNCH((k_jacobi_v<T>), gu, bl, c->s, g, (const T*)w.jv0, w.jv1, w.zvw, w.rho_v, w.kap, w.zeta, w.eta, dt, w.penv);
    LAUNCH((k_jacobi_v<T>), gu, bl, c->s, g, (const T*)w.jv1, Rv, w.zvw, w.rho_v, w.kap, w.zeta, w.eta, dt, w.penv);
  }
  if (!fuse_v) {
    halo(c, Rv, 1);
    if (c->pcd_cross || !c->use_vmg)
      LAUNCH((k_pcd_cross<T>), gc, bl, c->s, g, (const T*)Rv, w.eta, w.zuw, c->use_vmg ? w.vmg[0][0].b : (T*)nullptr);
  }
  const bool fuse_u = c->use_vmg && c->pcd_fused_glue && bcs_after;
  if (c->use_vmg) {
    const T* resu = vmg_solve<T>(c, w, 0, w.zuw, fuse_u ? (T*)nullptr : Ru, c->pcd_cross);
    if (fuse_u) LAUNCH((k_vmg_store_bcs<T>), gu, bl, c->s, g, resu, (const T*)w.zuw, Ru, Rv);
  } else {
    LAUNCH((k_jacobi_u<T>), gu, bl, c->s, g, (const T*)nullptr, w.ju0, w.zuw, w.rho_u, w.kap, w.zeta, w.eta, dt, w.penu);
    LAUNCH((k_jacobi_u<T>), gu, bl, c->s, g, (const T*)w.ju0, w.ju1, w.zuw, w.rho_u, w.kap, w.zeta, w.eta, dt, w.penu);
    LAUNCH((k_jacobi_u<T>), gu, bl, c->s, g, (const T*)w.ju1, Ru, w.zuw, w.rho_u, w.kap, w.zeta, w.eta, dt, w.penu);
  }
  if (!fuse_u) halo(c, Ru, 0);
  return fuse_u;
}

// ---------------------------------------------------------------------------------
// One iteration of the reference Krylov solvers, from its HIP graph or eagerly.  The graph holds
// the iteration's launches exactly as `emit` enqueues them; it is captured on the first iteration
// of a key and replayed after that.  Only the reference PCD qualifies (the multigrid forms read
// set-up results on the host), and neither kernel timing nor the launch log may be on.
void kry_graph_clear(fugu_ctx* c) {
  if (c->kry_graph) (void)hipGraphExecDestroy(c->kry_graph);
  c->kry_graph = nullptr;
  c->kry_graph_pcg = 0;
}
template <typename T, typename F>
void kry_iter(fugu_ctx* c, Work<T>& w, int order, bool consistent, int mvflags, int max_iter, double dt, F&& emit) {
  const bool graphed = c->krylov_graphs && !c->use_mg && !c->use_vmg && !c->kt_on && !c->launch_log;
  if (!graphed) {
    emit();
    return;
  }
  fugu_ctx::KryKey key;
  key.order = order; key.consistent = consistent ? 1 : 0; key.mvflags = mvflags; key.max_iter = max_iter;
  key.seq = seq_sums(c) ? 1 : 0; key.tsize = (int)sizeof(T); key.mv_tile = c->mv_tile;
  key.dt = dt; key.pcd_visc = c->pcd_visc; key.work = (const void*)&w;
  if (!c->kry_graph || !(key == c->kry_key)) {
    kry_graph_clear(c);
    const int pc0 = c->pcg_calls;
    hipGraph_t gr = nullptr;
    HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    try {
      emit();
    } catch (...) {
      (void)hipStreamEndCapture(c->s, &gr);
      if (gr) (void)hipGraphDestroy(gr);
      throw;
    }
    HIPCHK(hipStreamEndCapture(c->s, &gr));
    hipGraphExec_t ge = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    HIPCHK(e);
    c->kry_graph = ge;
    c->kry_graph_pcg = c->pcg_calls - pc0;
    c->kry_key = key;
  } else {
    c->pcg_calls += c->kry_graph_pcg;
  }
  HIPCHK(hipGraphLaunch(c->kry_graph, c->s));
}

// BiCGStab (src/mod_bicgstab2.f90:77-264), exact reference recurrences.
template <typename T>
void bicgstab1(fugu_ctx* c, Work<T>& w, double dt, double tol, int max_iter, int mvflags, bool has_lid,
               double lid, int* iters, double* res) {
  const int* brk = c->fl + F_BRK;
  T *X = w.X, *r = w.r, *r0 = w.rt, *p = w.uu, *v = w.vv, *s = w.ss, *t = w.tt, *z = w.z;
  reset_flags(c);
  set_slot(c, S_TOL, tol);
  set_slot(c, S_P_TOL, 1.0e-5);
  pcg_setup(c, w);
  if (c->use_mg) mg_setup(c, w);
  if (c->use_vmg) vmg_setup(c, w, (T)dt);
  bcs_trip(c, X);
  pcd(c, w, w.B, w.bh, (T)dt);
  cdot1<T>(c, w.bh, w.bh, PH_NORM_B);
  matvec(c, w, X, z, dt, mvflags, nullptr);
  pcd(c, w, z, v, (T)dt);
  lin<T>(c, L_DIFF, r, w.bh, v);
  lin<T>(c, L_COPY, r0, r);
  lin<T>(c, L_COPY, p, r);
  {
    const T* A[2] = {r0, r};
    const T* Bq[2] = {r, r};
    cdot<T>(c, 2, A, Bq, PH_B1_INIT, 0, max_iter);
  }
  read_flags(c);
  int k = 1;
  if (!c->h_fl[F_DONE]) {
    // the iteration number k reaches the phases through the device counter F_KITER (kk = -1),
    // so every iteration enqueues the same launches
    const int kk = -1;
    auto iter = [&]() {
      LAUNCH(k_iter_begin, dim3(1), dim3(64), c->s, c->fl);
      bcs_trip(c, p, 0.0, brk);
      matvec(c, w, p, z, dt, mvflags, brk);
      pcd(c, w, z, v, (T)dt);
      bcs_trip(c, v, 0.0, brk);
      cdot1<T>(c, r0, v, PH_B1_RV, kk, max_iter);
      lin<T>(c, L_SUB1, s, r, v, nullptr, nullptr, S_ALPHA);
      bcs_trip(c, s, 0.0, brk);
      matvec(c, w, s, z, dt, mvflags, brk);
      pcd(c, w, z, t, (T)dt);
      bcs_trip(c, t, 0.0, brk);
      {
        const T* A[2] = {t, t};
        const T* Bq[2] = {t, s};
        cdot<T>(c, 2, A, Bq, PH_B1_TT_TS, kk, max_iter);
      }
      lin<T>(c, L_ADD2, X, X, p, s, nullptr, S_ALPHA, S_OMEGA);
      lin<T>(c, L_SUB1, r, s, t, nullptr, nullptr, S_OMEGA);
      {
        const T* A[2] = {r, r0};
        const T* Bq[2] = {r, r};
        cdot<T>(c, 2, A, Bq, PH_B1_RES, kk, max_iter);
      }
      lin<T>(c, L_PUPD, p, r, p, v, nullptr, S_BETA, S_OMEGA);
      bcs_trip(c, X, 0.0, brk);
    };
    for (k = 1; k <= max_iter; ++k) {
      kry_iter(c, w, 1, false, mvflags, max_iter, dt, iter);
      read_flags(c);
      if (c->h_fl[F_DONE]) break;
    }
  }
  if (has_lid) bcs_trip(c, X, lid);
  read_flags(c);
  read_slots(c);
  *iters = c->h_fl[F_DONE] ? c->h_fl[F_ITER] : max_iter;
  *res = c->h_sc[S_REL_RES];
}

// BiCGStab(2) (src/mod_bicgstab2.f90:270-570).  consistent = BCs on r and s before
// their matvecs (not in the reference; see fugu2d.h FUGU_B2_CONSISTENT_BCS).
template <typename T>
void bicgstab2(fugu_ctx* c, Work<T>& w, double dt, double tol, int max_iter, int mvflags, bool has_lid,
               double lid, bool consistent, int* iters, double* res, bool relative_guards = false) {
  const int* brk = c->fl + F_BRK;
  T *X = w.X, *r = w.r, *rt = w.rt, *uu = w.uu, *vv = w.vv, *ww = w.ww, *ss = w.ss, *tt = w.tt, *z = w.z;
  reset_flags(c, relative_guards);
  set_slot(c, S_TOL, tol);
  set_slot(c, S_P_TOL, 1.0e-5);
  pcg_setup(c, w);
  if (c->use_mg) mg_setup(c, w);
  if (c->use_vmg) vmg_setup(c, w, (T)dt);
  bcs_trip(c, X);
  pcd(c, w, w.B, w.bh, (T)dt);
  cdot1<T>(c, w.bh, w.bh, PH_NORM_B);
  matvec(c, w, X, z, dt, mvflags, nullptr);
  pcd(c, w, z, w.q, (T)dt);
  lin<T>(c, L_DIFF, r, w.bh, w.q);
  lin<T>(c, L_COPY, rt, r);
  HIPCHK(hipMemsetAsync(uu, 0, c->ntrip * sizeof(T), c->s));
  HIPCHK(hipMemsetAsync(vv, 0, c->ntrip * sizeof(T), c->s));
  cdot1<T>(c, r, r, PH_B2_INIT, 0, max_iter);
  read_flags(c);
  if (!c->h_fl[F_DONE]) {
    const int kk = -1;   // the iteration number comes from F_KITER, as in bicgstab1
    auto iter = [&]() {
      LAUNCH(k_iter_begin, dim3(1), dim3(64), c->s, c->fl);
      // even step
      cdot1<T>(c, rt, r, PH_B2_XI_EVEN, kk, max_iter);
      lin<T>(c, L_SUB1, uu, r, uu, nullptr, nullptr, S_BETA);
      bcs_trip(c, uu, 0.0, brk);
      matvec(c, w, uu, z, dt, mvflags, brk);
      pcd(c, w, z, vv, (T)dt);
      bcs_trip(c, vv, 0.0, brk);
      cdot1<T>(c, rt, vv, PH_B2_GAMMA, kk, max_iter);
      lin<T>(c, L_SUB1, r, r, vv, nullptr, nullptr, S_ALPHA);
      if (consistent) bcs_trip(c, r, 0.0, brk);
      matvec(c, w, r, z, dt, mvflags, brk);
      pcd(c, w, z, ss, (T)dt);
      bcs_trip(c, ss, 0.0, brk);
      lin<T>(c, L_ADD1, X, X, uu, nullptr, nullptr, S_ALPHA);
      // odd step
      cdot1<T>(c, rt, ss, PH_B2_XI_ODD, kk, max_iter);
      lin<T>(c, L_SUB1, vv, ss, vv, nullptr, nullptr, S_BETA);
      bcs_trip(c, vv, 0.0, brk);
      matvec(c, w, vv, z, dt, mvflags, brk);
      pcd(c, w, z, ww, (T)dt);
      bcs_trip(c, ww, 0.0, brk);
      cdot1<T>(c, rt, ww, PH_B2_GAMMA, kk, max_iter);
      lin<T>(c, L_SUB1, uu, r, uu, nullptr, nullptr, S_BETA);
      lin<T>(c, L_SUB1, r, r, vv, nullptr, nullptr, S_ALPHA);
      lin<T>(c, L_SUB1, ss, ss, ww, nullptr, nullptr, S_ALPHA);
      if (consistent) bcs_trip(c, ss, 0.0, brk);
      matvec(c, w, ss, z, dt, mvflags, brk);
      pcd(c, w, z, tt, (T)dt);
      bcs_trip(c, tt, 0.0, brk);
      {
        const T* A[5] = {r, ss, ss, tt, r};
        const T* Bq[5] = {ss, ss, tt, tt, tt};
        cdot<T>(c, 5, A, Bq, PH_B2_GCR, kk, max_iter);
      }
      lin<T>(c, L_ADD3, X, X, r, ss, uu, S_OM1, S_OM2, S_ALPHA);
      lin<T>(c, L_SUB2, r, r, ss, tt, nullptr, S_OM1, S_OM2);
      lin<T>(c, L_SUB2, uu, uu, vv, ww, nullptr, S_OM1, S_OM2);
      cdot1<T>(c, r, r, PH_B2_RES, kk, max_iter);
      bcs_trip(c, X, 0.0, brk);
    };
    for (int k = 1; k <= max_iter; ++k) {
      kry_iter(c, w, 2, consistent, mvflags, max_iter, dt, iter);
      read_flags(c);
      if (c->h_fl[F_DONE]) break;
    }
  }
  if (has_lid) bcs_trip(c, X, lid);
  read_flags(c);
  read_slots(c);
  *iters = c->h_fl[F_DONE] ? c->h_fl[F_ITER] : max_iter;
  *res = c->h_sc[S_REL_RES];
}

// ---------------------------------------------------------------------------------
// True relative residual ||B - A X|| / ||B|| (coupled_dot norm, FP64) of the resident
// X triplet against the FP64 workspace B.  Uses wd.z (A X) and wd.q (residual).
double true_residual(fugu_ctx* c, const double* X, const double* B, int mvflags, double dt) {
  TreeSums tree(c);   // a diagnostic, not a reference quantity
  Work<double>& w = c->wd;
  HIPCHK(hipMemcpyAsync(w.X, X, c->ntrip * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  bcs_trip(c, w.X);
  matvec(c, w, w.X, w.z, dt, mvflags, nullptr);
  lin<double>(c, L_DIFF, w.q, B, w.z, nullptr, nullptr, -1, -1, -1, 1.0, false);
  {
    const double* A[2] = {w.q, B};
    const double* Bq[2] = {w.q, B};
    cdot<double>(c, 2, A, Bq, PH_STORE, 0, 0);
  }
  read_slots(c);
  const double nb = c->h_sc[S_DOT1];
  return nb > 0 ? sqrt(c->h_sc[S_DOT0] / nb) : sqrt(c->h_sc[S_DOT0]);
}

// ---------------------------------------------------------------------------------
// Adjoint building blocks (no counterpart in the reference).
void adj_alloc(fugu_ctx* c) {
  auto& a = c->aw;
  if (a.ready) return;
  const Grid& g = c->g;
  a.y = dalloc<double>(c, c->ntrip);
  a.hu0 = dalloc<double>(c, g.nu); a.hu1 = dalloc<double>(c, g.nu); a.zuwb = dalloc<double>(c, g.nu);
  a.rvb = dalloc<double>(c, g.nv); a.hv0 = dalloc<double>(c, g.nv); a.hv1 = dalloc<double>(c, g.nv);
  a.zvwb = dalloc<double>(c, g.nv); a.rpb = dalloc<double>(c, g.np); a.phib = dalloc<double>(c, g.np);
  a.phi = dalloc<double>(c, g.np);
  a.mu = dalloc<double>(c, g.np); a.rho = dalloc<double>(c, g.np);
  a.bmu = dalloc<double>(c, g.nu); a.bmv = dalloc<double>(c, g.nv);
  a.t2 = dalloc<double>(c, c->ntrip); a.ga = dalloc<double>(c, c->ntrip); a.G = dalloc<double>(c, c->ntrip);
  for (auto& l : a.lam) l = dalloc<double>(c, c->ntrip);
  a.h = dalloc<double>(c, g.nu + g.nv);
  a.grc = dalloc<double>(c, g.np); a.gru = dalloc<double>(c, g.nu); a.grv = dalloc<double>(c, g.nv);
  a.gk = dalloc<double>(c, g.np); a.gz = dalloc<double>(c, g.nw); a.ge = dalloc<double>(c, g.nw);
  a.gmu = dalloc<double>(c, g.np); a.grho = dalloc<double>(c, g.np);
  a.ready = true;
}

void zero(fugu_ctx* c, double* a, size_t n) {
  LAUNCH(k_axpby, dim3(nblk(n)), dim3(256), c->s, a, (const double*)nullptr, 0.0, (const double*)nullptr, 0.0, n);
}
void axpby(fugu_ctx* c, double* o, const double* x, double a, const double* y, double b, size_t n) {
  LAUNCH(k_axpby, dim3(nblk(n)), dim3(256), c->s, o, x, a, y, b, n);
}
void acc_prod(fugu_ctx* c, double* o, double s, const double* w, const double* x, const double* y, size_t n) {
  LAUNCH(k_acc_prod, dim3(nblk(n)), dim3(256), c->s, o, s, w, x, y, n);
}

// H^T of one padded field (y pass undone first, then x).
template <typename T>
void halo_T(fugu_ctx* c, T* a, int kind) {
  const Grid& g = c->g;
  if (g.per_y) {
    const int ni = (kind == 0) ? g.nx + 1 + 2 * g.nh : g.nx + 2 * g.nh;
    LAUNCH((k_halo_T_y<T>), dim3(nblk((size_t)ni * g.nh)), dim3(256), c->s, g, a, kind);
  }
  if (g.per_x) {
    const int nj = (kind == 1) ? g.ny + 1 + 2 * g.nh : g.ny + 2 * g.nh;
    LAUNCH((k_halo_T_x<T>), dim3(nblk((size_t)nj * g.nh)), dim3(256), c->s, g, a, kind);
  }
}
// E^T of a triplet: H^T, then B^T (y walls, then x walls), then the owned-set mask.
template <typename T>
void extend_T(fugu_ctx* c, T* X) {
  const Grid& g = c->g;
  T *u = X, *v = X + g.nu, *p = X + g.nu + g.nv;
  halo_T(c, u, 0); halo_T(c, v, 1); halo_T(c, p, 2);
  if (!g.per_y) LAUNCH((k_bcs_T_y<T>), dim3(nblk((size_t)g.nx + 1 + 2 * g.nh)), dim3(256), c->s, g, u, v);
  if (!g.per_x) LAUNCH((k_bcs_T_x<T>), dim3(nblk((size_t)g.ny + 1 + 2 * g.nh)), dim3(256), c->s, g, u, v);
  LAUNCH((k_mask_trip<T>), grid_union(g), dim3(TX, TY), c->s, g, u, v, p);
}
// ---------------------------------------------------------------------------------
// Transposes of the interface chain (no counterpart in the reference): the reverse mode of
// vof(), svof() at a given sweep count, curvature() and csf() at a forward state.
void iw_alloc(fugu_ctx* c) {
  auto& w = c->iw;
  if (w.ready) return;
  const Grid& g = c->g;
  for (double** a : {&w.cb, &w.tb, &w.kb, &w.sb0, &w.sb1, &w.acc, &w.csb, &w.cbar, &w.cbar2, &w.kbar})
    *a = dalloc<double>(c, g.np);
  for (double** a : {&w.gux, &w.guy, &w.ubar, &w.fub, &w.dx3[0], &w.dx3[1], &w.dx3[2]}) *a = dalloc<double>(c, g.nu);
  for (double** a : {&w.gvx, &w.gvy, &w.vbar, &w.fvb, &w.dy3[0], &w.dy3[1], &w.dy3[2]}) *a = dalloc<double>(c, g.nv);
  w.ready = true;
}
// cbar_in = (dVOF/dc_in)^T cbar_out; ubar, vbar = (dVOF/du)^T cbar_out, (dVOF/dv)^T cbar_out
// (all three overwritten; cbar_out is not modified).
void vof_T(fugu_ctx* c, const double* cin, const double* u, const double* v, double dt, const double* cbar_out,
           double* cbar_in, double* ubar, double* vbar) {
  iw_alloc(c);
  auto& w = c->iw;
  const Grid& g = c->g;
  const dim3 gu = grid_union(g), bl(TX, TY);
  HIPCHK(hipMemcpyAsync(w.cb, cbar_out, g.np * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  halo_T(c, w.cb, 2);
  LAUNCH(k_vof_T_cells, gu, bl, c->s, g, cin, u, v, dt, c->vof_limiter, (const double*)w.cb, w.tb, cbar_in);
  VofTFaces f;
  f.ubar = ubar; f.vbar = vbar;
  for (int q = 0; q < 3; ++q) { f.dx3[q] = w.dx3[q]; f.dy3[q] = w.dy3[q]; }
  LAUNCH(k_vof_T_faces, gu, bl, c->s, g, cin, u, v, dt, c->vof_limiter, (const double*)w.tb, f);
  LAUNCH(k_vof_T_gather, gu, bl, c->s, g, u, v, f, cbar_in);
}
// crawbar (interior) += (dSVOF/dc_raw)^T csbar at `iters` sweeps.
void svof_T(fugu_ctx* c, const double* csbar, double delta, int iters, double* crawbar) {
  iw_alloc(c);
  auto& w = c->iw;
  const Grid& g = c->g;
  const dim3 gu = grid_union(g), bl(TX, TY);
  const double cx = (delta / g.dx) * (delta / g.dx);
  const double cy = (delta / g.dy) * (delta / g.dy);
  const double diag = 1.0 + 2.0 * (cx + cy);
  LAUNCH(k_fill_T_p, gu, bl, c->s, g, csbar, w.sb0);
  LAUNCH(k_axpby, dim3(nblk(g.np)), dim3(256), c->s, w.acc, (const double*)nullptr, 0.0, (const double*)nullptr, 0.0,
         (size_t)g.np);
  double *a = w.sb0, *b = w.sb1;
  for (int it = iters; it >= 1; --it) {
    LAUNCH(k_svof_T_sweep, gu, bl, c->s, g, (const double*)a, b, w.acc, cx, cy, diag);
    std::swap(a, b);
  }
  LAUNCH(k_svof_T_final, gu, bl, c->s, g, (const double*)w.acc, (const double*)a, crawbar);
}
// csbar += (dcurvature/dcs)^T kbar.
void curv_T(fugu_ctx* c, const double* cs, const double* kbar, double* csbar) {
  iw_alloc(c);
  auto& w = c->iw;
  const Grid& g = c->g;
  const dim3 gu = grid_union(g), bl(TX, TY);
  HIPCHK(hipMemcpyAsync(w.kb, kbar, g.np * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  halo_T(c, w.kb, 2);
  LAUNCH(k_curv_T_faces, gu, bl, c->s, g, cs, (const double*)w.kb, w.gux, w.guy, w.gvx, w.gvy);
  LAUNCH(k_curv_T_cells, gu, bl, c->s, g, (const double*)w.gux, (const double*)w.guy, (const double*)w.gvx,
         (const double*)w.gvy, csbar);
}
// cbar += (dCSF/dc_raw)^T fbar; kbar = (dCSF/dk)^T fbar.
void csf_T(fugu_ctx* c, const double* craw, const double* kc, double sigma, const double* fub, const double* fvb,
           double* cbar, double* kbar) {
  LAUNCH(k_csf_T, grid_union(c->g), dim3(TX, TY), c->s, c->g, craw, kc, sigma, fub, fvb, cbar, kbar);
}

template <typename T>
void mask_trip(fugu_ctx* c, T* X) {
  const Grid& g = c->g;
  LAUNCH((k_mask_trip<T>), grid_union(g), dim3(TX, TY), c->s, g, X, X + g.nu, X + g.nu + g.nv);
}
template <typename T>
void zero_t(fugu_ctx* c, T* a, size_t n) {
  HIPCHK(hipMemsetAsync(a, 0, n * sizeof(T), c->s));
}

// O = A_eff^T L with w's operator fields (storage T, FP64 arithmetic).
template <typename T>
void matvec_T_w(fugu_ctx* c, Work<T>& w, const T* L, T* O, double dt, int adv) {
  const Grid& g = c->g;
  MatvecTArgs<T> a;
  a.Lu = L; a.Lv = L + g.nu; a.Lp = L + g.nu + g.nv;
  a.Ou = O; a.Ov = O + g.nu; a.Op = O + g.nu + g.nv;
  a.uadv = w.uadv; a.vadv = w.vadv; a.rho_c = w.rho_c; a.rho_u = w.rho_u; a.rho_v = w.rho_v;
  a.kap = w.kap; a.zeta = w.zeta; a.eta = w.eta; a.dt = dt; a.adv = adv;
  TLAUNCH(c, sizeof(T) == 8 ? "matvec_T" : "matvec_T_f32", 13.0 * sizeof(T) * g.nx * g.ny, (k_matvec_T<T>),
          grid_union(g), dim3(TX, TY), c->s, g, a);
  extend_T(c, O);
}
// O = A_eff^T L with the context's FP64 operator fields (c->wd bindings).
void matvec_T(fugu_ctx* c, const double* L, double* O, double dt, int adv) { matvec_T_w<double>(c, c->wd, L, O, dt, adv); }

// The transposed PCD's scratch in the storage type of its Krylov vectors (FP64: the adjoint
// work arrays; FP32: allocated on first use).
template <typename T>
fugu_ctx::PcdTBuf<T>& pcd_t_buf(fugu_ctx* c);
template <>
fugu_ctx::PcdTBuf<double>& pcd_t_buf<double>(fugu_ctx* c) {
  auto& b = c->ptd;
  auto& a = c->aw;
  b.y = a.y; b.hu0 = a.hu0; b.hu1 = a.hu1; b.zuwb = a.zuwb; b.rvb = a.rvb; b.hv0 = a.hv0; b.hv1 = a.hv1;
  b.zvwb = a.zvwb; b.rpb = a.rpb; b.phib = a.phib; b.phi = a.phi;
  return b;
}
template <>
fugu_ctx::PcdTBuf<float>& pcd_t_buf<float>(fugu_ctx* c) {
  auto& b = c->ptf;
  if (!b.y) {
    const Grid& g = c->g;
    b.y = dalloc<float>(c, c->ntrip);
    b.hu0 = dalloc<float>(c, g.nu); b.hu1 = dalloc<float>(c, g.nu); b.zuwb = dalloc<float>(c, g.nu);
    b.rvb = dalloc<float>(c, g.nv); b.hv0 = dalloc<float>(c, g.nv); b.hv1 = dalloc<float>(c, g.nv);
    b.zvwb = dalloc<float>(c, g.nv); b.rpb = dalloc<float>(c, g.np); b.phib = dalloc<float>(c, g.np);
    b.phi = dalloc<float>(c, g.np);
  }
  return b;
}

// Z = P_L^T Y: the transposed linearised PCD (w's operator fields and hierarchies, storage T).
// Pressure block: ptol > 0 -> projected PCG to ptol (pmax iterations at most); ptol <= 0 ->
// c->mg_cycles multigrid V-cycles (needs mg_setup).
template <typename T>
void pcd_T_w(fugu_ctx* c, Work<T>& w, const T* Y, T* Z, double dt, double ptol, int pmax, double visc) {
  const Grid& g = c->g;
  auto& a = pcd_t_buf<T>(c);
  const dim3 gu = grid_union(g), bl(TX, TY);
  HIPCHK(hipMemcpyAsync(a.y, Y, c->ntrip * sizeof(T), hipMemcpyDeviceToDevice, c->s));
  mask_trip(c, a.y);
  const T *yu = a.y, *yv = a.y + g.nu;
  T *Zu = Z, *Zv = Z + g.nu, *Zp = Z + g.nu + g.nv;
  // r_u = J_u^3(z_uw): three transposed sweeps, z_uw,b accumulated (velocity multigrid: the
  // symmetric V-cycle of the symmetric u block stands for its own transpose)
  if (c->use_vmg) {
    vmg_solve<T>(c, w, 0, yu, a.zuwb);
  } else {
    zero_t(c, a.zuwb, g.nu);
    LAUNCH((k_jacobi_u_T<T>), gu, bl, c->s, g, yu, a.hu0, a.zuwb, w.rho_u, w.kap, w.zeta, w.eta, dt);
    LAUNCH((k_jacobi_u_T<T>), gu, bl, c->s, g, (const T*)a.hu0, a.hu1, a.zuwb, w.rho_u, w.kap, w.zeta, w.eta, dt);
    LAUNCH((k_jacobi_u_T<T>), gu, bl, c->s, g, (const T*)a.hu1, a.hu0, a.zuwb, w.rho_u, w.kap, w.zeta, w.eta, dt);
  }
  // r_v from the output and from the u block's eta cross term
  HIPCHK(hipMemcpyAsync(a.rvb, yv, g.nv * sizeof(T), hipMemcpyDeviceToDevice, c->s));
  LAUNCH((k_cross_T<T>), gu, bl, c->s, g, (const T*)a.zuwb, w.eta, a.rvb);
  if (c->use_vmg) {
    vmg_solve<T>(c, w, 1, (const T*)a.rvb, a.zvwb);
  } else {
    zero_t(c, a.zvwb, g.nv);
    LAUNCH((k_jacobi_v_T<T>), gu, bl, c->s, g, (const T*)a.rvb, a.hv0, a.zvwb, w.rho_v, w.kap, w.zeta, w.eta, dt);
    LAUNCH((k_jacobi_v_T<T>), gu, bl, c->s, g, (const T*)a.hv0, a.hv1, a.zvwb, w.rho_v, w.kap, w.zeta, w.eta, dt);
    LAUNCH((k_jacobi_v_T<T>), gu, bl, c->s, g, (const T*)a.hv1, a.hv0, a.zvwb, w.rho_v, w.kap, w.zeta, w.eta, dt);
  }
  // z_u, z_v and r_p,b from z_uw,b and z_vw,b
  LAUNCH((k_zwork_T<T>), gu, bl, c->s, g, (const T*)a.zuwb, (const T*)a.zvwb, (const T*)(a.y + g.nu + g.nv), Zu,
         Zv, a.rpb);
  zero_t(c, Zp, g.np);
  LAUNCH((k_rp_T<T>), gu, bl, c->s, g, (const T*)a.rpb, w.uadv, w.vadv, dt, w.kap, visc, a.phib, Zp);
  // z_p += Pi A_p^+ Pi phi_b (symmetric), centred
  zero_t(c, a.phi, g.np);
  if (ptol > 0.0) {
    reset_flags(c);
    set_slot(c, S_P_TOL, ptol);
    pcg_setup(c, w);
    pcg<T>(c, w, a.phi, (const T*)a.phib, true, ptol, pmax, true);
  } else {   // c->mg_cycles V-cycles from zero on the projected rhs (mg_setup done by the caller)
    mg_pressure<T>(c, w, a.phi, (const T*)a.phib, true);
  }
  cell_mean<T>(c, (const T*)a.phi, (int)S_P_MEAN, 0);
  LAUNCH((k_add_centred<T>), gu, bl, c->s, g, Zp, (const T*)a.phi, (const double*)c->sc, (int)S_P_MEAN);
  mask_trip(c, Z);
}
void pcd_T(fugu_ctx* c, const double* Y, double* Z, double dt, double ptol, int pmax, double visc) {
  pcd_T_w<double>(c, c->wd, Y, Z, dt, ptol, pmax, visc);
}

void fg_graphs_clear(fugu_ctx* c) {
  for (auto& e : c->fg_graph)
    if (e) (void)hipGraphExecDestroy(e);
  c->fg_graph.clear();
  c->fg_graph_pcg.clear();
}

// Run FGMRES step k (`emit` enqueues its launches) from its graph, capturing it on first use.
template <typename F>
void fg_graph_run(fugu_ctx* c, const fugu_ctx::FgKey& key, int k, F&& emit) {
  if (!(key == c->fg_key)) {
    fg_graphs_clear(c);
    c->fg_key = key;
  }
  if ((int)c->fg_graph.size() <= k) {
    c->fg_graph.resize(k + 1, nullptr);
    c->fg_graph_pcg.resize(k + 1, 0);
  }
  if (!c->fg_graph[k]) {
    const int pc0 = c->pcg_calls;
    hipGraph_t gr = nullptr;
    HIPCHK(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    try {
      emit();
    } catch (...) {
      (void)hipStreamEndCapture(c->s, &gr);
      if (gr) (void)hipGraphDestroy(gr);
      throw;
    }
    HIPCHK(hipStreamEndCapture(c->s, &gr));
    hipGraphExec_t ge = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    HIPCHK(e);
    c->fg_graph[k] = ge;
    c->fg_graph_pcg[k] = c->pcg_calls - pc0;
  } else {
    c->pcg_calls += c->fg_graph_pcg[k];
  }
  HIPCHK(hipGraphLaunch(c->fg_graph[k], c->s));
}

// the chosen fast solve: restarted flexible GMRES(m) on the coupled system (coupled_dot
// norm), each restart recomputing the true residual B - A X and updating X in FP64, so the
// stopping test ||B - A X|| / ||B|| <= tol is met in FP64.  Preconditioner (max_inner > 0):
// an inner BiCGStab(2) + PCD solve of A z = v in precision T with a capped iteration count;
// (max_inner == 0) the PCD itself, in precision T.  With T = float and the PCD preconditioner
// the whole Arnoldi cycle runs in FP32 storage (GMRES-based iterative refinement): basis,
// directions, matvec operands and the preconditioner are FP32 triplets, the Gram-Schmidt
// products, sums and the Hessenberg/Givens scalars FP64; the cycle's correction is added to
// the FP64 X and the next restart's residual is FP64 again.  Flexibility absorbs the
// preconditioner's inexactness (FP32 rounding, non-converged inner recurrences).  The adjoint
// (c->adjoint) runs the same cycles on the transposed operator and the transposed PCD.
template <typename T>
void solve_fgmres(fugu_ctx* c, Work<T>& w, double* X, double dt, double tol, double inner_tol, int max_steps,
                  int max_inner, int m, int mvflags, bool has_lid, double lid, int* iters, int* outer,
                  double* res) {
  TreeSums tree(c);
  Work<double>& wd = c->wd;
  const size_t n3 = c->ntrip;
  const bool ir32 = sizeof(T) == sizeof(float) && max_inner == 0;
  if (m > FG_M - 1) m = FG_M - 1;
  if (m < 1) m = 1;
  while ((int)c->kv.size() < m + 1) c->kv.push_back(dalloc<double>(c, n3));
  if (ir32) {
    while ((int)c->kv32.size() < m + 1) c->kv32.push_back(dalloc<float>(c, n3));
    while ((int)c->kz32.size() < m) c->kz32.push_back(dalloc<float>(c, n3));
  } else {
    while ((int)c->kz.size() < m) c->kz.push_back(dalloc<double>(c, n3));
  }
  if (!c->fg) c->fg = dalloc<double>(c, FG_SIZE);
  double* B = c->Bst;
  HIPCHK(hipMemcpyAsync(B, wd.B, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  bcs_trip(c, X);
  reset_flags(c);
  if (max_inner == 0) {   // the PCD itself is the flexible preconditioner
    set_slot(c, S_P_TOL, 1.0e-5);
    pcg_setup(c, w);
    if (c->use_mg) mg_setup(c, w);
    if (c->use_vmg) vmg_setup(c, w, (T)dt);
  }
  cdot1<double>(c, B, B, PH_STORE);
  read_slots(c);
  const double bn = sqrt(c->h_sc[S_DOT0]) > 0 ? sqrt(c->h_sc[S_DOT0]) : 1.0;
  set_slot(c, S_FG_TOL, c->fg_est * tol * bn);
  // steps as graphs: the PCD forms only (an inner Krylov solve reads flags on the host), not
  // while kernels are timed or logged one by one
  const bool graphed = c->graphs && max_inner == 0 && !c->adjoint && c->use_mg && !c->kt_on && !c->launch_log;
  fugu_ctx::FgKey key;
  if (graphed) {
    key.prec = ir32 ? 4 : (int)sizeof(T);
    key.mvflags = mvflags;
    key.nl0 = c->use_vmg ? (int)w.vmg_nl[0] : 0;
    key.nl1 = c->use_vmg ? (int)w.vmg_nl[1] : 0;
    key.cyc0 = c->use_vmg ? w.vmg_cyc[0] : 0;
    key.cyc1 = c->use_vmg ? w.vmg_cyc[1] : 0;
    key.flags = (c->use_vmg ? 1 : 0) | (c->pcd_cross ? 2 : 0) | (c->mg_fused ? 4 : 0) | (c->mg_tail_stage ? 8 : 0) |
                (c->mg_tail_quad ? 16 : 0);
    key.mg_cycles = c->mg_cycles;
    key.dt = dt;
    key.pcd_visc = c->pcd_visc;
  }
  int total_inner = 0, steps = 0;
  double rr = 1.0;
  struct FgPcdFlag {   // PCD calls made by this solve directly on its basis vectors
    fugu_ctx* c;
    bool old;
    FgPcdFlag(fugu_ctx* cc, bool on) : c(cc), old(cc->fgmres_pcd) { c->fgmres_pcd = on; }
    ~FgPcdFlag() { c->fgmres_pcd = old; }
  } fg_pcd_flag(c, max_inner == 0 && !c->adjoint);
  // o = x0 + sgn sum_q c_q xs[q].  gs_mode 0: plain; 1: first Gram-Schmidt pass (norm and the
  // re-orthogonalisation test); 2: second pass, predicated on F_REORTH (norm when it runs).
  // fg_k >= 0 (gs_mode 2): the FGMRES step k follows the second pass's norm when that pass runs
  auto maxpy = [&](auto* o, const auto* x0, auto* const* xs, int nq, double sgn, int c_off, int gs_mode,
                   int fg_k = -1) {
    using O = std::remove_cv_t<std::remove_pointer_t<decltype(o)>>;
    using V = std::remove_cv_t<std::remove_pointer_t<std::remove_cv_t<std::remove_pointer_t<decltype(xs)>>>>;
    MaxpyArgs<O, V> ma{};
    ma.out = o; ma.x0 = x0; ma.nq = nq; ma.c_off = c_off; ma.sgn = sgn; ma.n = n3;
    for (int q = 0; q < nq; ++q) ma.x[q] = xs[q];
    unsigned nb = nblk(n3);
    if (nb > (unsigned)RED_MAXB) nb = RED_MAXB;
    double* pa = gs_mode ? c->partials : (double*)nullptr;
    const double ab = (2.0 * sizeof(O) + nq * sizeof(V)) * n3;
    const char* tname = sizeof(O) == 4 ? "maxpy_f32" : (sizeof(V) == 4 ? "maxpy_update" : "maxpy_f64");
    if (gs_mode == 2 && c->launch_log) std::fprintf(c->launch_log, "%s %.0f 0\n", tname, ab);
    if (gs_mode == 2)   // predicated: not timed (its bytes depend on a device-side flag)
      LAUNCH((k_maxpy<O, V>), dim3(nb), dim3(256), c->s, ma, (const double*)c->sc, pa, (const int*)c->fl,
             (int)F_REORTH);
    else
      TLAUNCH(c, tname, ab, (k_maxpy<O, V>), dim3(nb), dim3(256), c->s, ma, (const double*)c->sc, pa,
              (const int*)c->fl, -1);
    if (gs_mode)
      LAUNCH(k_fold_gs_norm, dim3(1), dim3(NT), c->s, (int)nb, gs_mode, c->g.vol, (const double*)c->partials,
             c->sc, c->fl, fg_k, c->fg, fg_k >= 0 ? c->d_hfl + F_COUNT + (fg_k & 1) : (int*)nullptr);
  };
  // W -= V h1 (coefficients at S_HBUF) fused with <W', V_i> -> S_HBUF + 64 + i when the
  // re-orthogonalisation test fires (zero otherwise), ||W'||^2 -> S_DOT0, the test -> F_REORTH;
  // when it does not fire, the FGMRES step fg_k follows in the fold
  auto maxpy_dot = [&](auto* W, auto* const* basis, int nq, int fg_k) {
    using V = std::remove_cv_t<std::remove_pointer_t<decltype(W)>>;
    MaxpyArgs<V, V> ma{};
    ma.out = W; ma.x0 = W; ma.nq = nq; ma.c_off = 0; ma.sgn = -1.0; ma.n = n3;
    for (int q = 0; q < MAXPY_MAX; ++q) ma.x[q] = q < nq ? basis[q] : W;   // padding slots: x0, coefficient 0
    unsigned nb = nblk(n3);
    if (nb > (unsigned)RED_MAXB) nb = RED_MAXB;
    const double ab = (nq + 2.0) * sizeof(V) * n3;
    const char* tname = sizeof(V) == 4 ? "maxpy_dot_f32" : "maxpy_dot_f64";
    int stride = 0;
#define FUGU_MAXPY_DOT(NQ)                                                                                  \
    if (nq <= NQ && !stride) {                                                                             \
      stride = NQ + 1;                                                                                      \
      TLAUNCH(c, tname, ab, (k_maxpy_dot<V, NQ>), dim3(nb), dim3(256), c->s, ma, (const double*)c->sc,     \
              c->partials);                                                                                 \
    }
    FUGU_MAXPY_DOT(8) FUGU_MAXPY_DOT(16) FUGU_MAXPY_DOT(32) FUGU_MAXPY_DOT(48)
#undef FUGU_MAXPY_DOT
    LAUNCH(k_fold_maxpy_dot, dim3(1), dim3(1024), c->s, (int)nb, nq, stride, c->g.vol, c->reorth_frac,
           (const double*)c->partials, c->sc, c->fl, fg_k, c->fg,
           fg_k >= 0 ? c->d_hfl + F_COUNT + (fg_k & 1) : (int*)nullptr);
  };
  // <W, V_i>, i = 0..k -> S_HBUF + off + i (and vol <W, W> -> S_DOT1); pass 2 predicated on F_REORTH
  auto dots_into = [&](const auto* W, auto* const* basis, int k, int off, bool second) {
    using V = std::remove_cv_t<std::remove_pointer_t<decltype(W)>>;
    MdotArgs<V> md{};
    md.w = W; md.nv = k + 1; md.n = n3;
    for (int q = 0; q <= k; ++q) md.v[q] = basis[q];
    const unsigned nch = (unsigned)((k + 1 + MD_CH - 1) / MD_CH);
    const int flag = second ? (int)F_REORTH : -1;
    const double ab = (k + 2.0) * sizeof(V) * n3;
    int vec = c->gs_vec ? 1 : 0;   // 16-byte loads when every array allows them
    if (((uintptr_t)W & 15) != 0) vec = 0;
    for (int q = 0; q <= k; ++q)
      if (((uintptr_t)basis[q] & 15) != 0) vec = 0;
    if (second && c->launch_log)
      std::fprintf(c->launch_log, "%s %.0f 0\n", sizeof(V) == 4 ? "mdot_f32" : "mdot_f64", ab);
    if (second) {   // predicated: not timed
      LAUNCH((k_mdot<V>), dim3(MD_BX, nch), dim3(256), c->s, md, c->partials, (const int*)c->fl, flag, vec);
    } else {
      TLAUNCH(c, sizeof(V) == 4 ? "mdot_f32" : "mdot_f64", ab, (k_mdot<V>), dim3(MD_BX, nch), dim3(256), c->s,
              md, c->partials, (const int*)c->fl, flag, vec);
    }
    LAUNCH(k_mdot_fold, dim3(1), dim3(1024), c->s, k + 1, off, c->g.vol, (const double*)c->partials, c->sc,
           (const int*)c->fl, flag);
  };
  while (true) {
    reset_flags(c);
    HIPCHK(hipMemcpyAsync(wd.X, X, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    if (c->adjoint) matvec_T(c, wd.X, wd.z, dt, (mvflags & MV_ADV) ? 1 : 0);
    else matvec(c, wd, wd.X, wd.z, dt, mvflags, nullptr);
    lin<double>(c, L_DIFF, c->kv[0], B, wd.z, nullptr, nullptr, -1, -1, -1, 1.0, false);
    cdot1<double>(c, c->kv[0], c->kv[0], PH_STORE);
    read_slots(c);
    const double beta = sqrt(c->h_sc[S_DOT0]);
    rr = beta / bn;
    if (c->debug) std::fprintf(stderr, "[fugu] fgmres restart: true res %.3e after %d steps\n", rr, steps);
    if (rr <= tol || beta == 0.0 || steps >= max_steps) break;
    LAUNCH(k_fgmres_init, dim3(1), dim3(1), c->s, c->sc, c->fg, c->fl);
    if (ir32)   // v_0 = r / beta rounded to FP32
      LAUNCH((k_convert<float, double>), dim3(nblk(n3)), dim3(256), c->s, c->kv32[0], (const double*)c->kv[0], n3,
             c->sc, (int)S_INVH, 1.0, 0);
    else
      lin<double>(c, L_SCALE, c->kv[0], c->kv[0], nullptr, nullptr, nullptr, S_INVH, -1, -1, 1.0, false);
    for (int k = 0; k < m && steps < max_steps; ++k) {
      // preconditioner z_k ~ A^-1 v_k in precision T
      int it = 1;
      auto step_k = [&]() {
      if (ir32) {
        float* z = c->kz32[k];
        if (c->adjoint) {
          // the transposed PCD and operator on FP32 storage (FP64 arithmetic in the kernels)
          pcd_T_w<float>(c, c->wf, (const float*)c->kv32[k], z, dt, -1.0, c->mg_cycles, c->pcd_visc);
          bcs_trip(c, z);
          HIPCHK(hipMemcpyAsync(c->wf.X, z, n3 * sizeof(float), hipMemcpyDeviceToDevice, c->s));
          matvec_T_w<float>(c, c->wf, c->wf.X, c->wf.z, dt, (mvflags & MV_ADV) ? 1 : 0);
        } else {
          if (!pcd<float>(c, c->wf, (const float*)c->kv32[k], z, (float)dt, true)) bcs_trip(c, z);
          // W = A z_k (FP32 operands; the matvec's halo exchange only rewrites z_k's periodic
          // ghosts with the images the PCD already put there), then Gram-Schmidt on FP32 storage
          matvec(c, c->wf, z, c->wf.z, dt, mvflags, nullptr);
        }
        float* W = c->wf.z;
        dots_into(W, c->kv32.data(), k, 0, false);
        maxpy_dot(W, c->kv32.data(), k + 1, k);
        maxpy(W, W, c->kv32.data(), k + 1, -1.0, 64, 2, k);   // the step k runs in one of the two folds
        lin<float>(c, L_SCALE, c->kv32[k + 1], W, nullptr, nullptr, nullptr, S_INVH, -1, -1, 1.0, false);
      } else {
        bool bcs_done = false;
        if (max_inner > 0) {
          LAUNCH((k_convert<T, double>), dim3(nblk(n3)), dim3(256), c->s, w.B, (const double*)c->kv[k], n3, c->sc,
                 -1, 1.0, 0);
          HIPCHK(hipMemsetAsync(w.X, 0, n3 * sizeof(T), c->s));
          double ires = 0.0;
          bicgstab2<T>(c, w, dt, inner_tol, max_inner, mvflags, false, 0.0, true, &it, &ires, true);
          LAUNCH((k_convert<double, T>), dim3(nblk(n3)), dim3(256), c->s, c->kz[k], (const T*)w.X, n3, c->sc, -1,
                 1.0, 0);
          reset_flags(c);
        } else if (c->adjoint) {
          pcd_T(c, (const double*)c->kv[k], c->kz[k], dt, -1.0, c->mg_cycles, c->pcd_visc);
        } else if (sizeof(T) == sizeof(double)) {
          bcs_done = pcd(c, w, (const T*)c->kv[k], (T*)c->kz[k], (T)dt, true);
        }
        if (!bcs_done) bcs_trip(c, c->kz[k]);
        // W = A z_k, then classical Gram-Schmidt twice, all coefficients device-resident
        if (c->adjoint) {
          HIPCHK(hipMemcpyAsync(wd.X, c->kz[k], n3 * sizeof(double), hipMemcpyDeviceToDevice, c->s));
          matvec_T(c, wd.X, wd.z, dt, (mvflags & MV_ADV) ? 1 : 0);
        } else {
          matvec(c, wd, c->kz[k], wd.z, dt, mvflags, nullptr);
        }
        double* W = wd.z;
        // classical Gram-Schmidt, second pass only when the first lost more than half of ||W||^2;
        // the second pass's products come out of the first pass's sweep
        dots_into(W, c->kv.data(), k, 0, false);
        maxpy_dot(W, c->kv.data(), k + 1, k);
        maxpy(W, W, c->kv.data(), k + 1, -1.0, 64, 2, k);   // the step k runs in one of the two folds
        lin<double>(c, L_SCALE, c->kv[k + 1], W, nullptr, nullptr, nullptr, S_INVH, -1, -1, 1.0, false);
      }
      };
      if (graphed) fg_graph_run(c, key, k, step_k);
      else step_k();
      total_inner += it;
      ++steps;
      // convergence is read one step behind, so the device always has the next step queued; the
      // fold that ran step k wrote its flag into h_fl[F_COUNT + (k & 1)] (mapped pinned memory)
      HIPCHK(hipEventRecord(c->fg_ev[k & 1], c->s));
      if (k >= 1) {
        HIPCHK(hipEventSynchronize(c->fg_ev[(k - 1) & 1]));
        if (c->h_fl[F_COUNT + ((k - 1) & 1)]) break;
      }
    }
    LAUNCH(k_fgmres_solve, dim3(1), dim3(1), c->s, c->sc, (const double*)c->fg);
    read_slots(c);
    const int kk = (int)c->h_sc[S_HBUF + 127];
    if (c->debug) std::fprintf(stderr, "[fugu] fgmres cycle: %d directions\n", kk);
    if (kk > 0) {
      if (ir32) maxpy(X, (const double*)X, c->kz32.data(), kk, 1.0, 128, 0);
      else maxpy(X, (const double*)X, c->kz.data(), kk, 1.0, 128, 0);
    }
    bcs_trip(c, X);
  }
  if (has_lid) bcs_trip(c, X, lid);
  *iters = total_inner;
  *outer = steps;
  *res = rr;
}

// ---------------------------------------------------------------------------------
// SVOF, curvature, CSF, mixture and VOF helpers on resident buffers
void svof(fugu_ctx* c, const double* craw, double* cs_out, double delta, int max_iter, double tol, int* it_out) {
  const Grid& g = c->g;
  const double cx = (delta / g.dx) * (delta / g.dx);
  const double cy = (delta / g.dy) * (delta / g.dy);
  const double diag = 1.0 + 2.0 * (cx + cy);
  HIPCHK(hipMemcpyAsync(c->csA, craw, g.np * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  HIPCHK(hipMemsetAsync(c->fl + F_S_DONE, 0, 2 * sizeof(int), c->s));
  double* B[2] = {c->csA, c->csB};
  const dim3 tb((unsigned)((g.nx + SVOF_TW - 1) / SVOF_TW), (unsigned)((g.ny + SVOF_TH - 1) / SVOF_TH));
  const size_t ntb = (size_t)tb.x * tb.y;
  const double* last = c->csA;
  if (c->svof_blocked && ntb * SVOF_S <= c->partials_n) {
    // SVOF_S iterations per launch; block b reads B[b & 1] and writes B[(b + 1) & 1]; the
    // launch after the last block only recomputes a block that converged before its end
    const int nbk = (max_iter + SVOF_S - 1) / SVOF_S;
    auto block = [&](int b) {
      const int it0 = b * SVOF_S + 1, cnt = b < nbk ? std::min(SVOF_S, max_iter - b * SVOF_S) : 0;
      LAUNCH(k_svof_block, tb, dim3(SVOF_NT), c->s, g, craw, (const double*)B[b & 1], B[(b + 1) & 1], cx, cy, diag, b,
             cnt, c->partials, (const int*)c->fl);
      if (cnt > 0)
        LAUNCH(k_fold_svof_blk, dim3(1), dim3(NT), c->s, (int)ntb, b, it0, cnt, tol, (const double*)c->partials,
               c->sc, c->fl);
    };
    // The blocks after convergence launch and exit at once.  A first batch sized from the last
    // call's iteration count (one block of margin) leaves them out; the flags read after it
    // decide whether the rest is needed, or only the launch that recomputes a block that
    // converged before its end (the same launches as the full sequence up to that point).
    const int first = c->svof_pred > 0 ? std::min(nbk, (c->svof_pred + SVOF_S - 1) / SVOF_S + 1) : nbk;
    for (int b = 0; b <= first; ++b) block(b);
    read_flags(c);
    if (first < nbk) {
      if (!c->h_fl[F_S_DONE]) {
        for (int b = first + 1; b <= nbk; ++b) block(b);
        read_flags(c);
      } else if (c->h_fl[F_S_RBLK] == first) {
        block(first + 1);
      }
    }
    const int itf = c->h_fl[F_S_ITER];
    c->svof_pred = c->h_fl[F_S_DONE] ? itf : 0;
    if (it_out) *it_out = itf;
    if (max_iter >= 1) last = B[((itf - 1) / SVOF_S + 1) & 1];   // the output of the block holding itf
  } else {
    const dim3 gr = grid_cells_red(g);
    for (int it = 1; it <= max_iter; ++it) {
      LAUNCH(k_svof_sweep, gr, dim3(TX, TY), c->s, g, craw, (const double*)B[(it - 1) & 1], B[it & 1], cx, cy,
             diag, c->partials, (const int*)c->fl);
      LAUNCH(k_fold_svof, dim3(1), dim3(NT), c->s, (int)nblocks(gr), it, tol, (const double*)c->partials, c->sc,
             c->fl);
    }
    // the last iterate is in B[iter & 1]
    read_flags(c);
    const int itf = c->h_fl[F_S_ITER];
    if (it_out) *it_out = itf;
    if (max_iter >= 1) last = B[itf & 1];
  }
  // ghosts: C_smooth starts as a copy of C_raw (ghosts included), then exchange + mirror
  HIPCHK(hipMemcpyAsync(cs_out, craw, g.np * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  {
    // interior of the last iterate into cs_out
    LinArgs<double> a;  // not used: copy interior via 2D memcpy below
    (void)a;
    const size_t rowb = (size_t)g.nx * sizeof(double);
    HIPCHK(hipMemcpy2DAsync(cs_out + IP(g, 1, 1), g.ldp * sizeof(double), last + IP(g, 1, 1),
                            g.ldp * sizeof(double), rowb, g.ny, hipMemcpyDeviceToDevice, c->s));
  }
  halo(c, cs_out, 2);
  if (!g.per_x || !g.per_y) LAUNCH(k_mirror_p<double>, grid_union(g), dim3(TX, TY), c->s, g, cs_out);
}

void curvature(fugu_ctx* c, const double* cs, double* kc) {
  const Grid& g = c->g;
  LAUNCH(k_normals, grid_union(g), dim3(TX, TY), c->s, g, cs, c->nxu, c->nyv);
  LAUNCH(k_curv, grid_union(g), dim3(TX, TY), c->s, g, (const double*)c->nxu, (const double*)c->nyv, kc);
  halo(c, kc, 2);
}

void csf(fugu_ctx* c, const double* craw, const double* kc, double* fu, double* fv) {
  LAUNCH(k_csf, grid_union(c->g), dim3(TX, TY), c->s, c->g, craw, kc, c->prm.sigma, fu, fv);
}

void mixture(fugu_ctx* c, const double* cc) {
  LAUNCH(k_mixture, grid_union(c->g), dim3(TX, TY), c->s, c->g, cc, c->prm.rho1, c->prm.rho2, c->prm.mu1,
         c->prm.mu2, c->rho_c, c->rho_u, c->rho_v, c->kap, c->zeta, c->eta);
}

void vof(fugu_ctx* c, const double* cin, double* cout, const double* u, const double* v, double dt,
         double* fx, double* fy, double* dclamp = nullptr) {
  LAUNCH(k_vof, grid_union(c->g), dim3(TX, TY), c->s, c->g, cin, cout, u, v, dt, fx, fy, c->vof_limiter, dclamp);
  if (fx) halo(c, fx, 0);
  if (fy) halo(c, fy, 1);
  halo(c, cout, 2);
}

// ---------------------------------------------------------------------------------
// one advance_coupled_step on the resident state.
int step_impl(fugu_ctx* c, double dt, int step_idx, const fugu_step_opts* o, fugu_step_stats* st) {
  const Grid& g = c->g;
  const fugu_params& P = c->prm;
  fugu_step_opts def = {1, 0, 0.0, 1.0e-6, 500, 0, 0.0, 0.0, 0u, 0.0, 0, 0, 0};
  if (!o) o = &def;
  const bool bdf2 = (P.time_order == 2 && step_idx >= 2);
  const double dt_eff = bdf2 ? dt / 1.5 : dt;
  const double lid = o->has_lid ? o->u_lid : 0.0;
  const bool pen = c->has_solid;
  const bool amp = o->use_amp != 0;
  HIPCHK(hipEventRecord(c->ev[0], c->s));
  // 1. u_adv, u_prev (src/mod_solver_coupled.f90:209-222): unm1 <- u_prev, u_prev <- u, u_adv <- 2u - unm1
  HIPCHK(hipMemcpyAsync(c->unm1, c->u_prev, g.nu * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  HIPCHK(hipMemcpyAsync(c->vnm1, c->v_prev, g.nv * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  LAUNCH(k_ab2, dim3(nblk(g.nu)), dim3(256), c->s, c->uadv, c->u_prev, (const double*)c->u, g.nu, step_idx <= 1 ? 1 : 0);
  LAUNCH(k_ab2, dim3(nblk(g.nv)), dim3(256), c->s, c->vadv, c->v_prev, (const double*)c->v, g.nv, step_idx <= 1 ? 1 : 0);
  // 2. VOF (:224-235) with u^n
  vof(c, c->c, c->cnew, c->u, c->v, dt, amp ? c->fx : nullptr, amp ? c->fy : nullptr, c->dclamp);
  {   // volume budget of the update: sum C before and after, and the clamp's part (S_VOL_*)
    const dim3 gr = grid_cells_red(g);
    LAUNCH(k_vol_sums, gr, dim3(TX, TY), c->s, g, (const double*)c->c, (const double*)c->cnew,
           (const double*)c->dclamp, c->partials);
    LAUNCH(k_fold_vol, dim3(1), dim3(NT), c->s, (int)nblocks(gr), g.vol, (const double*)c->partials, c->sc);
  }
  if (amp) {
    LAUNCH(k_amp, dim3(nblk(g.nu)), dim3(256), c->s, c->mdu, (const double*)c->fx, (const double*)c->uadv, g.nu,
           P.rho1 - P.rho2, P.rho2);
    LAUNCH(k_amp, dim3(nblk(g.nv)), dim3(256), c->s, c->mdv, (const double*)c->fy, (const double*)c->vadv, g.nv,
           P.rho1 - P.rho2, P.rho2);
  }
  std::swap(c->c, c->cnew);
  // 3. mixture (:237-239)
  mixture(c, c->c);
  // 4. SVOF + curvature + CSF (:241-252)
  if (P.sigma > 1.0e-14) {
    const double delta = 1.5 * (g.dx < g.dy ? g.dx : g.dy);
    svof(c, c->c, c->cs, delta, 150, 1.0e-7, nullptr);
    curvature(c, c->cs, c->kcurv);
    csf(c, c->c, c->kcurv, c->fu, c->fv);
  } else {
    HIPCHK(hipMemsetAsync(c->kcurv, 0, g.np * sizeof(double), c->s));
    HIPCHK(hipMemsetAsync(c->fu, 0, g.nu * sizeof(double), c->s));
    HIPCHK(hipMemsetAsync(c->fv, 0, g.nv * sizeof(double), c->s));
  }
  // 5. penalty (:254-259)
  if (pen) {
    LAUNCH(k_scale, dim3(nblk(g.nu)), dim3(256), c->s, c->penu, (const double*)c->phis_u, g.nu, P.alpha_penalty);
    LAUNCH(k_scale, dim3(nblk(g.nv)), dim3(256), c->s, c->penv, (const double*)c->phis_v, g.nv, P.alpha_penalty);
  }
  // 6. RHS (:261-303) into wd.B
  Work<double>& w = c->wd;
  {
    RhsArgs a;
    a.un = c->u; a.unm1 = c->unm1; a.vn = c->v; a.vnm1 = c->vnm1; a.rho_u = c->rho_u; a.rho_v = c->rho_v;
    a.fu = c->fu; a.fv = c->fv; a.penu = pen ? c->penu : nullptr; a.penv = pen ? c->penv : nullptr;
    a.zeta = c->zeta; a.eta = c->eta;
    a.Bu = w.B; a.Bv = w.B + g.nu; a.Bp = w.B + g.nu + g.nv;
    a.dt = dt; a.gx = P.gx; a.gy = P.gy; a.u_solid = o->u_solid; a.v_solid = o->v_solid; a.u_lid = lid;
    a.bdf2 = bdf2 ? 1 : 0; a.has_lid = o->has_lid ? 1 : 0;
    LAUNCH(k_rhs, grid_union(g), dim3(TX, TY), c->s, g, a);
  }
  // 7. solve (:305-366); initial guess X^n, or 2 X^n - X^(n-1) (u_adv, v_adv are 2u^n - u^(n-1))
  if (o->solver_flags & FUGU_SOLVE_EXTRAPOLATE) {
    if (!c->p_prev) c->p_prev = dalloc<double>(c, g.np);
    if (step_idx >= 2 && c->have_pprev) {
      HIPCHK(hipMemcpyAsync(c->u, c->uadv, g.nu * sizeof(double), hipMemcpyDeviceToDevice, c->s));
      HIPCHK(hipMemcpyAsync(c->v, c->vadv, g.nv * sizeof(double), hipMemcpyDeviceToDevice, c->s));
      LAUNCH(k_extrap, dim3(nblk(g.np)), dim3(256), c->s, c->p, c->p_prev, g.np);
    } else {
      HIPCHK(hipMemcpyAsync(c->p_prev, c->p, g.np * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    }
    c->have_pprev = true;
  } else {
    c->have_pprev = false;
  }
  bind_props_d(c, pen, amp);
  const int mvflags = MV_ADV | (pen ? MV_PEN : 0) | (amp ? MV_CONS : 0);
  HIPCHK(hipEventRecord(c->ev[1], c->s));
  int iters = 0, outer = 0;
  double res = 0.0;
  c->pcg_calls = 0;
  PrecondScope pscope(c);
  c->use_mg = (o->solver_flags & FUGU_PRECOND_MG) != 0;
  c->use_vmg = (o->solver_flags & FUGU_PRECOND_VMG) != 0;
  c->pcd_visc = (o->solver_flags & FUGU_PRECOND_VISC) ? c->pcd_visc_coef : 0.0;
  // pressure V-cycles per PCD: the caller's, else 2 from 2^22 cells (fewer FGMRES steps pay for the
  // second cycle there: 2048^2 5.33 -> 5.48 FP64 solves/s, 111/84/84 -> 96/77/78 steps), else 1
  // (512 x 1024: 174 against 145 solves/s with 2)
  c->mg_cycles = o->mg_cycles > 0 ? o->mg_cycles : ((double)g.nx * g.ny >= (double)(1 << 22) ? 2 : 1);
  const double itol = o->inner_tol > 0 ? o->inner_tol : 1.0e-3;
  const int imax = o->inner_max_iter > 0 ? o->inner_max_iter : (o->inner_max_iter < 0 ? 0 : 25);
  // default restart (step k's Gram-Schmidt sweeps read k + 2 basis vectors twice; a restart costs
  // one true residual and a host round trip): 10 after a refined solve of 16-30 FGMRES steps on
  // this context, otherwise 20 up to 2^21 cells and 30 above.  Measured (whole steps): the
  // 512x1024 bubble (20 steps per solve) 153 -> 163-165 FP64 solves/s with 10 (same step count;
  // 8: 147, 12: 156, 15: 154); the 128^2 cavity (12-15 steps) 534 time steps/s with 20, 478 with
  // 10; the 256^2 droplet (7-12) alike with either; 2048^2 (84-111 steps at 1e-8) 5.30 / 8.36
  // solves/s (FP64 / mixed) with 30, 5.50 / 8.49 with 20, 4.24 / 6.39 with 10 -- but 20 stalls a
  // 1e-13 solve there (2e-11 after 600 steps), so 30 stays.
  const int rm = o->restart > 0 ? o->restart
                                : ((c->fg_last_steps >= 16 && c->fg_last_steps <= 30)
                                       ? 10
                                       : ((long)c->g.nx * c->g.ny <= (1L << 21) ? 20 : 30));
  if (o->solver_flags & FUGU_SOLVE_REFINED) {
    if (o->solver_flags & FUGU_SOLVE_MIXED) {
      work_alloc<float>(c, c->wf, true);
      bind_props_f(c, pen, amp);
      solve_fgmres<float>(c, c->wf, c->X, dt_eff, o->tol, itol, o->max_iter, imax, rm, mvflags, true, lid,
                          &iters, &outer, &res);
    } else {
      solve_fgmres<double>(c, c->wd, c->X, dt_eff, o->tol, itol, o->max_iter, imax, rm, mvflags, true, lid,
                           &iters, &outer, &res);
    }
    c->fg_last_steps = outer;
  } else {
    HIPCHK(hipMemcpyAsync(w.X, c->X, c->ntrip * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    if (o->use_bicgstab2)
      bicgstab2<double>(c, w, dt_eff, o->tol, o->max_iter, mvflags, true, lid,
                        (o->solver_flags & FUGU_B2_CONSISTENT_BCS) != 0, &iters, &res);
    else
      bicgstab1<double>(c, w, dt_eff, o->tol, o->max_iter, mvflags, true, lid, &iters, &res);
    HIPCHK(hipMemcpyAsync(c->X, w.X, c->ntrip * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  }
  HIPCHK(hipEventRecord(c->ev[2], c->s));
  // 8. final halos + BCs (:368-372)
  halo(c, c->u, 0);
  halo(c, c->v, 1);
  halo(c, c->p, 2);
  bcs(c, c->u, c->v, lid);
  HIPCHK(hipEventRecord(c->ev[3], c->s));
  sync(c);
  kt_resolve(c);
  if (st) {
    float ms_solve = 0.f, ms_step = 0.f;
    HIPCHK(hipEventElapsedTime(&ms_solve, c->ev[1], c->ev[2]));
    HIPCHK(hipEventElapsedTime(&ms_step, c->ev[0], c->ev[3]));
    st->iters = iters;
    st->outer = outer;
    st->res = res;
    st->true_res = -1.0;
    st->solve_seconds = ms_solve * 1e-3;
    st->step_seconds = ms_step * 1e-3;
    st->pcg_calls = c->pcg_calls;
    read_slots(c);
    st->vol_before = c->h_sc[S_VOL_OLD];
    st->vol_change = c->h_sc[S_VOL_NEW] - c->h_sc[S_VOL_OLD];
    st->vol_clamp = c->h_sc[S_VOL_CLAMP];
  }
  int rc = 0;
  if (!(res == res)) rc |= FUGU_W_NONFINITE;
  if (o->solver_flags & FUGU_SOLVE_REFINED) {
    if (res > o->tol) rc |= FUGU_W_MAX_ITER;   // stopped at max_iter FGMRES steps short of tol
  } else if (iters >= o->max_iter) {
    rc |= FUGU_W_MAX_ITER;
  }
  return rc;
}

// ---------------------------------------------------------------------------------
// Multi-step adjoint of the frozen-interface BDF2 horizon (no counterpart in the reference).
//   forward: properties from the per-cell fields (fixed over the horizon); per step n
//     u_adv = u^{n-1} (n = 1) or 2u^{n-1} - u^{n-2}; C^n = VOF(C^{n-1}, u^{n-1}); unit-sigma CSF
//     force fhat^n from C^n (recorded); B = (rho/dt) hist + rho g + sigma fhat; A_eff(dt_eff) X^n
//     = B by FP64 FGMRES + PCD (multigrid pressure block, viscous Schur term) to opts->tol;
//     u^n = E X^n.  J = 1/2 sum_band rho_face u^2 vol at step N.
//   reverse: A_n^T lam_n = dJ/dX^n + couplings of steps n+1, n+2 (BDF history and the AB2
//     advecting velocity), by FGMRES on A_eff^T with the transposed PCD; gradients
//     -lam^T dA/dcoef X + lam^T dB/dcoef chained through the per-cell property map.
int adjoint_run(fugu_ctx* c, const fugu_adjoint_opts* o, const double* mu_in, const double* rho_in, double sigma,
                const double* c0, const double* u0, const double* v0, double* g_mu, double* g_rho,
                double* g_sigma, fugu_adjoint_stats* st) {
  TreeSums tree(c);
  const Grid& g = c->g;
  const fugu_params& P = c->prm;
  const int N = o->nsteps;
  const double dt = o->dt;
  if (N < 1 || !(dt > 0.0)) throw Error("adjoint: nsteps >= 1 and dt > 0 required");
  const bool reuse = (o->flags & FUGU_ADJ_REUSE_INTERFACE) != 0;
  const bool fwd_only = (o->flags & FUGU_ADJ_FORWARD_ONLY) != 0;
  const bool mixed = (o->flags & FUGU_ADJ_MIXED_FORWARD) != 0;
  // the interface motion differentiated too (VOF, SVOF, curvature and CSF transposed in the reverse
  // sweep); nothing to differentiate when sigma is zero (the interface exerts no force)
  const bool moving = (o->flags & FUGU_ADJ_MOVING_INTERFACE) != 0 && sigma > 1.0e-14;
  // checkpointing: every K-th step's restart state kept, a window of W states recomputed per
  // segment in the reverse sweep (K <= 1: the whole horizon stays resident)
  const int K = (o->checkpoint_every > 1 && o->checkpoint_every < N) ? o->checkpoint_every : 0;
  if (K && reuse) throw Error("adjoint: FUGU_ADJ_REUSE_INTERFACE needs the whole horizon resident (no checkpointing)");
  if ((o->flags & FUGU_ADJ_MOVING_INTERFACE) && reuse)
    throw Error("adjoint: FUGU_ADJ_MOVING_INTERFACE differentiates the interface motion; it cannot reuse a frozen one");
  const int W = K ? K + 5 : N + 1;
  auto slot = [&](int n) { return K ? n % W : n; };
  if (reuse && c->traj_n != N) throw Error("adjoint: no recorded interface of this horizon to reuse");
  if (!mu_in || !rho_in || !u0 || !v0 || (!reuse && !c0)) throw Error("adjoint: null input field");
  adj_alloc(c);
  auto& a = c->aw;
  Work<double>& wd = c->wd;
  const size_t n3 = c->ntrip, nu = g.nu, nv = g.nv;
  while ((int)c->traj_u.size() < W) {
    c->traj_u.push_back(dalloc<double>(c, nu)); c->traj_v.push_back(dalloc<double>(c, nv));
    c->traj_fu.push_back(dalloc<double>(c, nu)); c->traj_fv.push_back(dalloc<double>(c, nv));
  }
  const int ncp = K ? (N + K - 1) / K : 0;
  while ((int)c->cp.size() < ncp) {
    fugu_ctx::Checkpoint q;
    q.u1 = dalloc<double>(c, nu); q.u2 = dalloc<double>(c, nu); q.v1 = dalloc<double>(c, nv);
    q.v2 = dalloc<double>(c, nv); q.p = dalloc<double>(c, g.np); q.c = dalloc<double>(c, g.np);
    c->cp.push_back(q);
  }
  if (moving) {
    while ((int)c->traj_c.size() < W) c->traj_c.push_back(dalloc<double>(c, g.np));
    iw_alloc(c);
  }
  auto CT = [&](int n) { return c->traj_c[slot(n)]; };   // C^n (moving interface)
  auto Us = [&](int n) { return c->traj_u[slot(n)]; };
  auto Vs = [&](int n) { return c->traj_v[slot(n)]; };
  auto FUs = [&](int n) { return c->traj_fu[slot(n)]; };
  auto FVs = [&](int n) { return c->traj_fv[slot(n)]; };
  const double vol = g.vol;
  const int rm = o->restart > 0 ? o->restart : 30;
  const int mi = o->max_iter > 0 ? o->max_iter : 400;
  const double tol = o->tol > 0 ? o->tol : 1.0e-12;
  // properties
  in(c, a.mu, mu_in, g.np); in(c, a.rho, rho_in, g.np);
  LAUNCH(k_cell_props, grid_union(g), dim3(TX, TY), c->s, g, (const double*)a.mu, (const double*)a.rho, c->rho_c,
         c->rho_u, c->rho_v, c->kap, c->zeta, c->eta);
  bind_props_d(c, false, false);
  const bool mixed_adj = (o->flags & FUGU_ADJ_MIXED_ADJOINT) != 0;
  if (mixed || mixed_adj) work_alloc<float>(c, c->wf, true);
  c->use_mg = true;
  c->use_vmg = (o->flags & FUGU_ADJ_PRECOND_VMG) != 0;
  c->pcd_visc = c->pcd_visc_coef;
  c->mg_cycles = o->mg_cycles > 0 ? o->mg_cycles : 1;
  // initial state (BCs and halos as the end of a step leaves them)
  in(c, Us(0), u0, nu); in(c, Vs(0), v0, nv);
  bcs(c, Us(0), Vs(0), 0.0); halo(c, Us(0), 0); halo(c, Vs(0), 1);
  if (!reuse) in(c, c->c, c0, g.np);
  if (moving) HIPCHK(hipMemcpyAsync(CT(0), c->c, g.np * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  HIPCHK(hipMemcpyAsync(c->u, Us(0), nu * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  HIPCHK(hipMemcpyAsync(c->v, Vs(0), nv * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  zero(c, c->p, g.np);
  auto set_uadv = [&](int n) {   // u_adv of step n
    if (n <= 1) {
      HIPCHK(hipMemcpyAsync(c->uadv, Us(0), nu * sizeof(double), hipMemcpyDeviceToDevice, c->s));
      HIPCHK(hipMemcpyAsync(c->vadv, Vs(0), nv * sizeof(double), hipMemcpyDeviceToDevice, c->s));
    } else {
      axpby(c, c->uadv, Us(n - 1), 2.0, Us(n - 2), -1.0, nu);
      axpby(c, c->vadv, Vs(n - 1), 2.0, Vs(n - 2), -1.0, nv);
    }
  };
  auto bdf2_of = [&](int n) { return P.time_order == 2 && n >= 2; };
  auto dte_of = [&](int n) { return bdf2_of(n) ? dt / 1.5 : dt; };
  auto cpy = [&](double* dst, const double* src, size_t n) {
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  };
  int fwd_it = 0, adj_it = 0;
  double fwd_res = 0.0, adj_res = 0.0;
  // one forward step n from the state (c->u, c->v, c->p, c->c) = step n-1; stores u^n, v^n and
  // the unit-sigma force of step n in their slots
  auto fwd_step = [&](int n, bool count) {
    set_uadv(n);
    if (!reuse) {
      vof(c, c->c, c->cnew, Us(n - 1), Vs(n - 1), dt, nullptr, nullptr);
      std::swap(c->c, c->cnew);
      if (moving) cpy(CT(n), c->c, g.np);
      if (sigma > 1.0e-14) {
        const double delta = 1.5 * (g.dx < g.dy ? g.dx : g.dy);
        svof(c, c->c, c->cs, delta, 150, 1.0e-7, nullptr);
        curvature(c, c->cs, c->kcurv);
        LAUNCH(k_csf, grid_union(g), dim3(TX, TY), c->s, g, (const double*)c->c, (const double*)c->kcurv, 1.0,
               FUs(n), FVs(n));
      } else {
        zero(c, FUs(n), nu); zero(c, FVs(n), nv);
      }
    }
    axpby(c, c->fu, FUs(n), sigma, nullptr, 0.0, nu);
    axpby(c, c->fv, FVs(n), sigma, nullptr, 0.0, nv);
    RhsArgs r;
    r.un = Us(n - 1); r.unm1 = (n >= 2) ? Us(n - 2) : Us(0); r.vn = Vs(n - 1); r.vnm1 = (n >= 2) ? Vs(n - 2) : Vs(0);
    r.rho_u = c->rho_u; r.rho_v = c->rho_v; r.fu = c->fu; r.fv = c->fv; r.penu = nullptr; r.penv = nullptr;
    r.zeta = c->zeta; r.eta = c->eta;
    r.Bu = wd.B; r.Bv = wd.B + nu; r.Bp = wd.B + nu + nv;
    r.dt = dt; r.gx = P.gx; r.gy = P.gy; r.u_solid = 0.0; r.v_solid = 0.0; r.u_lid = 0.0;
    r.bdf2 = bdf2_of(n) ? 1 : 0; r.has_lid = 0;
    LAUNCH(k_rhs, grid_union(g), dim3(TX, TY), c->s, g, r);
    int it = 0, outer = 0;
    double res = 0.0;
    if (mixed) {
      bind_props_f(c, false, false);
      solve_fgmres<float>(c, c->wf, c->X, dte_of(n), tol, 0.0, mi, 0, rm, MV_ADV, false, 0.0, &it, &outer, &res);
    } else {
      solve_fgmres<double>(c, wd, c->X, dte_of(n), tol, 0.0, mi, 0, rm, MV_ADV, false, 0.0, &it, &outer, &res);
    }
    if (count) {
      fwd_it += outer;
      fwd_res = std::max(fwd_res, res);
    }
    // u^n = E X^n
    mask_trip(c, c->X);
    bcs(c, c->u, c->v, 0.0); halo(c, c->u, 0); halo(c, c->v, 1); halo(c, c->p, 2);
    cpy(Us(n), c->u, nu);
    cpy(Vs(n), c->v, nv);
  };
  // restart state of step n (the state after step n-1): u^(n-1), u^(n-2), v alike, p^(n-1), C^(n-1)
  auto save_cp = [&](int n) {
    fugu_ctx::Checkpoint& q = c->cp[(n - 1) / K];
    cpy(q.u1, Us(n - 1), nu); cpy(q.u2, Us(n >= 2 ? n - 2 : 0), nu);
    cpy(q.v1, Vs(n - 1), nv); cpy(q.v2, Vs(n >= 2 ? n - 2 : 0), nv);
    cpy(q.p, c->p, g.np); cpy(q.c, c->c, g.np);
  };
  auto load_cp = [&](int n) {
    const fugu_ctx::Checkpoint& q = c->cp[(n - 1) / K];
    cpy(Us(n - 1), q.u1, nu); cpy(Vs(n - 1), q.v1, nv);
    if (n >= 2) { cpy(Us(n - 2), q.u2, nu); cpy(Vs(n - 2), q.v2, nv); }
    cpy(c->u, q.u1, nu); cpy(c->v, q.v1, nv); cpy(c->p, q.p, g.np); cpy(c->c, q.c, g.np);
    if (moving) cpy(CT(n - 1), q.c, g.np);
  };
  HIPCHK(hipEventRecord(c->ev[0], c->s));
  for (int n = 1; n <= N; ++n) {
    if (K && (n - 1) % K == 0) save_cp(n);
    fwd_step(n, true);
  }
  // J = 1/2 coupled_dot((band rho u, band rho v, 0), (u, v, 0))
  LAUNCH(k_band_mask, grid_union(g), dim3(TX, TY), c->s, g, o->y0, a.bmu, a.bmv);
  zero(c, a.y, n3); zero(c, a.t2, n3);
  acc_prod(c, a.y, 1.0, a.bmu, c->rho_u, Us(N), nu);
  acc_prod(c, a.y + nu, 1.0, a.bmv, c->rho_v, Vs(N), nv);
  HIPCHK(hipMemcpyAsync(a.t2, Us(N), nu * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  HIPCHK(hipMemcpyAsync(a.t2 + nu, Vs(N), nv * sizeof(double), hipMemcpyDeviceToDevice, c->s));
  reset_flags(c);
  cdot1<double>(c, a.y, a.t2, PH_STORE);
  read_slots(c);
  const double J = 0.5 * c->h_sc[S_DOT0];
  HIPCHK(hipEventRecord(c->ev[1], c->s));
  c->traj_n = K ? 0 : N;   // a checkpointed horizon is not kept whole: nothing to reuse
  double gs = 0.0;
  if (!fwd_only) {
    zero(c, a.grc, g.np); zero(c, a.gru, nu); zero(c, a.grv, nv); zero(c, a.gk, g.np);
    zero(c, a.gz, g.nw); zero(c, a.ge, g.nw); zero(c, a.gmu, g.np); zero(c, a.grho, g.np);
    // explicit dJ/drho_face (first: u^N may leave the window when segments are recomputed)
    acc_prod(c, a.gru, 0.5 * vol, a.bmu, Us(N), Us(N), nu);
    acc_prod(c, a.grv, 0.5 * vol, a.bmv, Vs(N), Vs(N), nv);
    auto coef_grad = [&](const double* L, int n, double s, int what, double* gua, double* gva) {
      CoefGradArgs q;
      q.Lu = L; q.Lv = L + nu; q.Xu = Us(n); q.Xv = Vs(n); q.uadv = c->uadv; q.vadv = c->vadv;
      q.rho_c = c->rho_c; q.rho_u = c->rho_u; q.rho_v = c->rho_v; q.kap = c->kap; q.zeta = c->zeta; q.eta = c->eta;
      q.grc = a.grc; q.gru = a.gru; q.grv = a.grv; q.gk = a.gk; q.gz = a.gz; q.ge = a.ge; q.gua = gua; q.gva = gva;
      q.dt = dte_of(n); q.s = s; q.adv = 1; q.what = what;
      LAUNCH(k_coef_grad, grid_union(g), dim3(TX, TY), c->s, g, q);
    };
    c->adjoint = true;
    for (int n = N; n >= 1; --n) {
      // checkpointed horizon: on entering a segment [a, b] (below the last), recompute its
      // forward states from the checkpoint at a; states b+1, b+2 are still in the window
      if (K && n % K == 0 && n < N) {
        const int sa = n - K + 1;
        c->adjoint = false;
        load_cp(sa);
        for (int q = sa; q <= n; ++q) fwd_step(q, false);
        c->adjoint = true;
      }
      double* G = a.G;
      zero(c, G, n3);
      if (n == N) {
        acc_prod(c, G, vol, a.bmu, c->rho_u, Us(N), nu);
        acc_prod(c, G + nu, vol, a.bmv, c->rho_v, Vs(N), nv);
      }
      for (int m = n + 1; m <= n + 2 && m <= N; ++m) {
        const double* L = a.lam[m % 3];
        const bool b2 = bdf2_of(m);
        const double cb = (m == n + 1) ? (b2 ? 2.0 : 1.0) : (b2 ? -0.5 : 0.0);
        if (cb != 0.0) {
          acc_prod(c, G, cb / dt, c->rho_u, L, nullptr, nu);
          acc_prod(c, G + nu, cb / dt, c->rho_v, L + nu, nullptr, nv);
        }
        // through u_adv^m = 2 u^{m-1} - u^{m-2} (m >= 2)
        const double cadv = (m == n + 1) ? 2.0 : -1.0;
        set_uadv(m);
        zero(c, a.ga, n3);
        coef_grad(L, m, 1.0, CG_ADV, a.ga, a.ga + nu);
        extend_T(c, a.ga);
        acc_prod(c, G, -cadv, nullptr, a.ga, nullptr, nu);
        acc_prod(c, G + nu, -cadv, nullptr, a.ga + nu, nullptr, nv);
      }
      // moving interface: the velocity u^n advected C^n into C^(n+1) (VOF of step n + 1, transposed
      // at the previous iteration); u^n = E X^n, so E^T of that velocity adjoint joins G
      if (moving && n < N) {
        zero(c, a.ga, n3);
        cpy(a.ga, c->iw.ubar, nu);
        cpy(a.ga + nu, c->iw.vbar, nv);
        extend_T(c, a.ga);
        acc_prod(c, G, 1.0, nullptr, a.ga, nullptr, nu);
        acc_prod(c, G + nu, 1.0, nullptr, a.ga + nu, nullptr, nv);
      }
      mask_trip(c, G);
      // A_n^T lam_n = G
      set_uadv(n);
      HIPCHK(hipMemcpyAsync(wd.B, G, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->s));
      double* Ln = a.lam[n % 3];
      zero(c, Ln, n3);
      int it = 0, outer = 0;
      double res = 0.0;
      if (mixed_adj) {
        bind_props_f(c, false, false);
        solve_fgmres<float>(c, c->wf, Ln, dte_of(n), tol, 0.0, mi, 0, rm, MV_ADV, false, 0.0, &it, &outer, &res);
      } else {
        solve_fgmres<double>(c, wd, Ln, dte_of(n), tol, 0.0, mi, 0, rm, MV_ADV, false, 0.0, &it, &outer, &res);
      }
      adj_it += outer;
      adj_res = std::max(adj_res, res);
      mask_trip(c, Ln);
      // - lam^T dA/dcoef X^n
      coef_grad(Ln, n, -1.0, CG_COEF, nullptr, nullptr);
      // + lam^T dB/dcoef: rho_u (hist / dt + g_x), rho_v alike
      const bool b2 = bdf2_of(n);
      const double* um2 = (n >= 2) ? Us(n - 2) : Us(0);
      const double* vm2 = (n >= 2) ? Vs(n - 2) : Vs(0);
      axpby(c, a.h, Us(n - 1), b2 ? 2.0 : 1.0, b2 ? um2 : nullptr, -0.5, nu);
      axpby(c, a.h + nu, Vs(n - 1), b2 ? 2.0 : 1.0, b2 ? vm2 : nullptr, -0.5, nv);
      acc_prod(c, a.gru, 1.0 / dt, nullptr, Ln, a.h, nu);
      acc_prod(c, a.grv, 1.0 / dt, nullptr, Ln + nu, a.h + nu, nv);
      acc_prod(c, a.gru, P.gx, nullptr, Ln, nullptr, nu);
      acc_prod(c, a.grv, P.gy, nullptr, Ln + nu, nullptr, nv);
      // sigma: lam^T fhat
      zero(c, a.t2, n3);
      HIPCHK(hipMemcpyAsync(a.t2, FUs(n), nu * sizeof(double), hipMemcpyDeviceToDevice, c->s));
      HIPCHK(hipMemcpyAsync(a.t2 + nu, FVs(n), nv * sizeof(double), hipMemcpyDeviceToDevice, c->s));
      reset_flags(c);
      cdot1<double>(c, Ln, a.t2, PH_STORE);
      read_slots(c);
      gs += c->h_sc[S_DOT0] / vol;
      if (moving) {
        // dJ/dfhat^n = sigma lam_n: through CSF (C^n itself and the curvature), the curvature
        // (of the filtered C^n) and the SVOF filter into Cbar^n, which already holds what step n + 1's
        // advection sent back; then the advection of step n (from C^(n-1) with u^(n-1)) transposed:
        // Cbar^(n-1) and the velocity adjoint for the next iteration's G
        auto& w = c->iw;
        if (n == N) zero(c, w.cbar, g.np);
        const double delta = 1.5 * (g.dx < g.dy ? g.dx : g.dy);
        int its = 0;
        svof(c, CT(n), c->cs, delta, 150, 1.0e-7, &its);
        curvature(c, c->cs, c->kcurv);
        axpby(c, w.fub, Ln, sigma, nullptr, 0.0, nu);
        axpby(c, w.fvb, Ln + nu, sigma, nullptr, 0.0, nv);
        csf_T(c, CT(n), c->kcurv, 1.0, w.fub, w.fvb, w.cbar, w.kbar);
        zero(c, w.csb, g.np);
        curv_T(c, c->cs, w.kbar, w.csb);
        svof_T(c, w.csb, delta, its, w.cbar);
        vof_T(c, CT(n - 1), Us(n - 1), Vs(n - 1), dt, w.cbar, w.cbar2, w.ubar, w.vbar);
        std::swap(w.cbar, w.cbar2);
      }
    }
    c->adjoint = false;
    // the property map
    LAUNCH(k_cell_props_T, grid_union(g), dim3(TX, TY), c->s, g, (const double*)a.mu, (const double*)a.grc,
           (const double*)a.gru, (const double*)a.grv, (const double*)a.gk, (const double*)a.gz,
           (const double*)a.ge, a.gmu, a.grho);
    out(c, g_mu, a.gmu, g.np);
    out(c, g_rho, a.grho, g.np);
  }
  HIPCHK(hipEventRecord(c->ev[2], c->s));
  sync(c);
  kt_resolve(c);
  c->use_mg = false; c->use_vmg = false;
  c->pcd_visc = 0.0;
  if (g_sigma) *g_sigma = gs;
  if (st) {
    float f_ms = 0.f, r_ms = 0.f;
    HIPCHK(hipEventElapsedTime(&f_ms, c->ev[0], c->ev[1]));
    HIPCHK(hipEventElapsedTime(&r_ms, c->ev[1], c->ev[2]));
    st->J = J;
    st->forward_seconds = f_ms * 1e-3;
    st->reverse_seconds = r_ms * 1e-3;
    st->forward_iters = fwd_it;
    st->adjoint_iters = adj_it;
    st->max_forward_res = fwd_res;
    st->max_adjoint_res = adj_res;
  }
  int rc = 0;
  if (!(fwd_res == fwd_res) || !(adj_res == adj_res) || !(J == J)) rc |= FUGU_W_NONFINITE;
  if (fwd_res > tol || adj_res > tol) rc |= FUGU_W_MAX_ITER;
  return rc;
}

template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

#define NEED(ctx) \
  if (!(ctx)) { g_err = "null context"; return -2; }

}  // namespace

// =====================================================================================
// Allocation-free coupled dot (fugu_coupled_dot_noalloc): its scratch is this file-scope device
// array -- the reference order's terms, chunk totals, value counts and sum, or the tree order's
// per-workgroup partials -- held by every GPU the library runs on until the process ends.  Host
// code never uses the array's name as an address: each device's address comes from
// hipGetSymbolAddress on the first call on that device (that device current) and is kept per
// device.  Calls that overlap on one GPU share the array and race.
constexpr long NOALLOC_MAX_TERMS = 1L << 20;   // owned u + v + p points (reference order)
constexpr long NOALLOC_NCH = NOALLOC_MAX_TERMS / SQ_CH_MIN;   // chunks of the shortest length
constexpr long NOALLOC_PARTIALS = 2L * RED_MAXB;       // one double-double per workgroup
constexpr long NOALLOC_CTOT = 3 * NOALLOC_NCH;
constexpr long NOALLOC_CNT = NOALLOC_NCH / 2;          // NOALLOC_NCH ints
constexpr long NOALLOC_TOT = 8;
constexpr long NOALLOC_DOUBLES =
    NOALLOC_PARTIALS + NOALLOC_CTOT + NOALLOC_CNT + NOALLOC_TOT + NOALLOC_MAX_TERMS;   // 8.4 MB
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

namespace {
constexpr int NOALLOC_MAX_DEV = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEV] = {};

double* scratch_on_current_device() {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEV) throw Error("fugu_coupled_dot_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error("fugu_coupled_dot_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[dev];
}
}  // namespace

// =====================================================================================
// C ABI
extern "C" {

const char* fugu_last_error(void) { return g_err.c_str(); }

long long fugu_field_size(const fugu_ctx* c, int kind) {
  if (!c) return -1;
  switch (kind) {
    case FUGU_KIND_U: return (long long)c->g.nu;
    case FUGU_KIND_V: return (long long)c->g.nv;
    case FUGU_KIND_P: return (long long)c->g.np;
    case FUGU_KIND_W: return (long long)c->g.nw;
    default: return -1;
  }
}

int fugu_ctx_create(const fugu_grid* grid, const fugu_params* prm, fugu_ctx** out) {
  if (!grid || !prm || !out) { g_err = "null argument"; return -2; }
  if (grid->nh != 2) { g_err = "nh must be 2 (src/main.f90:22)"; return -2; }
  if (grid->nx < 4 || grid->ny < 4) { g_err = "grid too small (nx, ny >= 4)"; return -2; }
  fugu_ctx* c = new fugu_ctx();
  int rc = guarded([&]() {
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev < 1) throw Error("no HIP device");
    c->g = make_grid(grid->nx, grid->ny, grid->nh, grid->lx / grid->nx, grid->ly / grid->ny,
                     grid->periodic_x ? 1 : 0, grid->periodic_y ? 1 : 0);
    c->prm = *prm;
    if (const char* lp = std::getenv("FUGU_LAUNCH_LOG")) c->launch_log = std::fopen(lp, "a");
    if (const char* tc = std::getenv("FUGU_MG_TAIL_CELLS")) c->mg_tail_cells = std::atoi(tc);
    if (const char* fu = std::getenv("FUGU_MG_FUSED")) c->mg_fused = std::atoi(fu) != 0;
    if (const char* ts = std::getenv("FUGU_MG_TAIL_STAGE")) c->mg_tail_stage = std::atoi(ts) != 0;
    if (const char* tq = std::getenv("FUGU_MG_TAIL_QUAD")) c->mg_tail_quad = std::atoi(tq) != 0;
    if (const char* st = std::getenv("FUGU_MG_STRIP")) c->mg_strip = std::atoi(st);
    if (const char* sl = std::getenv("FUGU_MGS_LARGE_LOG2")) c->mgs_large_log2 = std::atoi(sl);
    if (const char* ss = std::getenv("FUGU_MGS_SMALL_LOG2")) c->mgs_small_log2 = std::atoi(ss);
    if (const char* sq = std::getenv("FUGU_SEQ_CHUNK")) c->seq_chunk = std::atoi(sq);
    if (const char* s = std::getenv("FUGU_MG_NU_P")) c->mg_nu_p = std::min(4, std::max(1, std::atoi(s)));
    if (const char* s = std::getenv("FUGU_MV_TILE")) c->mv_tile = std::atoi(s);
    if (const char* s = std::getenv("FUGU_GS_VEC")) c->gs_vec = std::atoi(s) != 0;
    if (const char* s = std::getenv("FUGU_MG_NU_V")) c->mg_nu_v = std::min(4, std::max(1, std::atoi(s)));
    if (const char* fe = std::getenv("FUGU_FGMRES_EST")) c->fg_est = std::atof(fe);
    if (const char* pc = std::getenv("FUGU_PCD_CROSS")) c->pcd_cross = std::atoi(pc) != 0;
    if (const char* vl = std::getenv("FUGU_VMG_LEVELS")) c->vmg_levels = std::atoi(vl);
    if (const char* vt = std::getenv("FUGU_VMG_THETA")) c->vmg_theta = std::atof(vt);
    if (const char* v2 = std::getenv("FUGU_VMG_THETA2")) c->vmg_theta2 = std::atof(v2);
    if (const char* vc = std::getenv("FUGU_VMG_COARSE_NU")) c->vmg_coarse_nu = std::atoi(vc);
    if (const char* pv = std::getenv("FUGU_PCD_VISC_COEF")) c->pcd_visc_coef = std::atof(pv);
    if (const char* zc = std::getenv("FUGU_ZP_CENTRE")) c->zp_centre = std::atoi(zc) != 0;
    if (const char* fg = std::getenv("FUGU_PCD_FUSED_GLUE")) c->pcd_fused_glue = std::atoi(fg) != 0;
    if (const char* vy = std::getenv("FUGU_VMG_CYCLES")) c->vmg_cycles = std::max(0, std::atoi(vy));
    if (const char* vr = std::getenv("FUGU_VMG_CYCLE_RATIO")) c->vmg_cycle_ratio = std::atof(vr);
    if (const char* rf = std::getenv("FUGU_REORTH_FRAC")) c->reorth_frac = std::atof(rf);
    if (const char* gr = std::getenv("FUGU_GRAPHS")) c->graphs = std::atoi(gr) != 0;
    if (const char* kg = std::getenv("FUGU_KRYLOV_GRAPHS")) c->krylov_graphs = std::atoi(kg) != 0;
    if (const char* sb = std::getenv("FUGU_SVOF_BLOCKED")) c->svof_blocked = std::atoi(sb) != 0;
    HIPCHK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    for (auto& e : c->ev) HIPCHK(hipEventCreate(&e));
    for (auto& e : c->fg_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const Grid& g = c->g;
    c->ntrip = g.nu + g.nv + g.np;
    // two-pass reductions: RED_MAXB workgroups x up to 5 double-double values; multi-dot: 6
    // chunks x MD_BX workgroups x (MD_CH + 1) values
    c->partials_n = std::max({(size_t)RED_MAXB * 10, (size_t)6 * MD_BX * (MD_CH + 1),
                              (size_t)RED_MAXB * (MAXPY_MAX + 1)});
    c->partials = dalloc<double>(c, c->partials_n);
    c->sc = dalloc<double>(c, S_COUNT);
    c->fl = (int*)dalloc<double>(c, F_COUNT);
    HIPCHK(hipHostMalloc((void**)&c->h_sc, S_COUNT * sizeof(double), hipHostMallocDefault));
    // two landing slots past the flags for the FGMRES convergence flag (read_flags copies
    // only the first F_COUNT, so an inner solve cannot overwrite them)
    HIPCHK(hipHostMalloc((void**)&c->h_fl, (F_COUNT + 2) * sizeof(int), hipHostMallocDefault));
    HIPCHK(hipHostGetDevicePointer((void**)&c->d_hfl, c->h_fl, 0));
    c->X = dalloc<double>(c, c->ntrip);
    c->u = c->X; c->v = c->X + g.nu; c->p = c->X + g.nu + g.nv;
    c->c = dalloc<double>(c, g.np); c->cnew = dalloc<double>(c, g.np); c->dclamp = dalloc<double>(c, g.np);
    c->u_prev = dalloc<double>(c, g.nu); c->v_prev = dalloc<double>(c, g.nv);
    c->unm1 = dalloc<double>(c, g.nu); c->vnm1 = dalloc<double>(c, g.nv);
    c->rho_c = dalloc<double>(c, g.np); c->rho_u = dalloc<double>(c, g.nu); c->rho_v = dalloc<double>(c, g.nv);
    c->kap = dalloc<double>(c, g.np); c->zeta = dalloc<double>(c, g.nw); c->eta = dalloc<double>(c, g.nw);
    c->kcurv = dalloc<double>(c, g.np); c->fu = dalloc<double>(c, g.nu); c->fv = dalloc<double>(c, g.nv);
    c->phis_u = dalloc<double>(c, g.nu); c->phis_v = dalloc<double>(c, g.nv);
    c->uadv = dalloc<double>(c, g.nu); c->vadv = dalloc<double>(c, g.nv);
    c->csA = dalloc<double>(c, g.np); c->csB = dalloc<double>(c, g.np); c->cs = dalloc<double>(c, g.np);
    c->nxu = dalloc<double>(c, g.nu); c->nyv = dalloc<double>(c, g.nv);
    c->fx = dalloc<double>(c, g.nu); c->fy = dalloc<double>(c, g.nv);
    c->mdu = dalloc<double>(c, g.nu); c->mdv = dalloc<double>(c, g.nv);
    c->penu = dalloc<double>(c, g.nu); c->penv = dalloc<double>(c, g.nv);
    c->Bst = dalloc<double>(c, c->ntrip);
    for (int k = 0; k < 4; ++k) {
      c->su[k] = dalloc<double>(c, g.nu); c->sv[k] = dalloc<double>(c, g.nv); c->sp[k] = dalloc<double>(c, g.np);
    }
    work_alloc<double>(c, c->wd, false);
    bind_props_d(c, false, false);
    sync(c);
    return 0;
  });
  if (rc != 0) {
    fugu_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

int fugu_ctx_destroy(fugu_ctx* c) {
  if (!c) return 0;
  if (c->s) (void)hipStreamSynchronize(c->s);
  fg_graphs_clear(c);
  kry_graph_clear(c);
  for (void* p : c->allocs) (void)hipFree(p);
  for (auto& kv : c->kt)
    for (auto& e : kv.second.ev) (void)hipEventDestroy(e);
  if (c->h_sc) (void)hipHostFree(c->h_sc);
  if (c->h_fl) (void)hipHostFree(c->h_fl);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->fg_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->s) (void)hipStreamDestroy(c->s);
  if (c->launch_log) std::fclose(c->launch_log);
  delete c;
  return 0;
}

int fugu_set_params(fugu_ctx* c, const fugu_params* p) {
  NEED(c);
  if (!p) return -2;
  c->prm = *p;
  return 0;
}

int fugu_set_sum_order(fugu_ctx* c, int order) {
  NEED(c);
  if (order != FUGU_SUM_REFERENCE && order != FUGU_SUM_TREE) {
    g_err = "fugu_set_sum_order: order must be FUGU_SUM_REFERENCE (0) or FUGU_SUM_TREE (1)";
    return -3;
  }
  c->sum_ref = order == FUGU_SUM_REFERENCE;
  return 0;
}

int fugu_set_vof_limiter(fugu_ctx* c, int limiter) {
  NEED(c);
  if (limiter != FUGU_LIMITER_VAN_LEER && limiter != FUGU_LIMITER_SUPERBEE) {
    g_err = "fugu_set_vof_limiter: limiter must be FUGU_LIMITER_VAN_LEER (0) or FUGU_LIMITER_SUPERBEE (1)";
    return -3;
  }
  c->vof_limiter = limiter;
  return 0;
}

int fugu_get_sum_order(const fugu_ctx* c) {
  NEED(c);
  return c->sum_ref ? FUGU_SUM_REFERENCE : FUGU_SUM_TREE;
}

int fugu_exchange_halos(fugu_ctx* c, double* a, int kind) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    if (kind < 0 || kind > 2) throw Error("kind must be FUGU_KIND_U, _V or _P");
    const size_t n = kind == 0 ? g.nu : (kind == 1 ? g.nv : g.np);
    double* d = kind == 0 ? c->su[0] : (kind == 1 ? c->sv[0] : c->sp[0]);
    in(c, d, a, n);
    halo(c, d, kind);
    out(c, a, d, n);
    sync(c);
    return 0;
  });
}

int fugu_op_divergence(fugu_ctx* c, const double* u, const double* v, double* div) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    Work<double>& w = c->wd;
    // divergence is the p-row of the matvec: run the matvec with zero properties' effect
    in(c, w.X, u, g.nu); in(c, w.X + g.nu, v, g.nv);
    HIPCHK(hipMemsetAsync(w.X + g.nu + g.nv, 0, g.np * sizeof(double), c->s));
    MatvecArgs<double> a{};
    a.Xu = w.X; a.Xv = w.X + g.nu; a.Xp = w.X + g.nu + g.nv;
    a.Yu = w.z; a.Yv = w.z + g.nu; a.Yp = w.z + g.nu + g.nv;
    a.uadv = c->uadv; a.vadv = c->vadv; a.rho_c = c->rho_c; a.rho_u = c->rho_u; a.rho_v = c->rho_v;
    a.kap = c->kap; a.zeta = c->zeta; a.eta = c->eta; a.dt = 1.0; a.flags = 0;
    LAUNCH((k_matvec<double>), grid_union(g), dim3(TX, TY), c->s, g, a, (const int*)nullptr);
    out(c, div, w.z + g.nu + g.nv, g.np);
    sync(c);
    return 0;
  });
}

int fugu_apply_velocity_bcs(fugu_ctx* c, double* u, double* v, const double* lid) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    in(c, c->su[0], u, g.nu); in(c, c->sv[0], v, g.nv);
    bcs(c, c->su[0], c->sv[0], lid ? *lid : 0.0);
    out(c, u, c->su[0], g.nu); out(c, v, c->sv[0], g.nv);
    sync(c);
    return 0;
  });
}

static void upload_props(fugu_ctx* c, const double* uadv, const double* vadv, const double* rho_c,
                         const double* rho_u, const double* rho_v, const double* kap, const double* zeta,
                         const double* eta, const double* pu, const double* pv, const double* mu,
                         const double* mv) {
  const Grid& g = c->g;
  in(c, c->uadv, uadv, g.nu); in(c, c->vadv, vadv, g.nv);
  in(c, c->rho_c, rho_c, g.np); in(c, c->rho_u, rho_u, g.nu); in(c, c->rho_v, rho_v, g.nv);
  in(c, c->kap, kap, g.np); in(c, c->zeta, zeta, g.nw); in(c, c->eta, eta, g.nw);
  in(c, c->penu, pu, g.nu); in(c, c->penv, pv, g.nv);
  in(c, c->mdu, mu, g.nu); in(c, c->mdv, mv, g.nv);
  bind_props_d(c, pu && pv, mu && mv);
  if (!pu) c->wd.penu = nullptr;
  if (!pv) c->wd.penv = nullptr;
}

int fugu_matvec_coupled(fugu_ctx* c, double* Xu, double* Xv, double* Xp, double* Yu, double* Yv, double* Yp,
                        const double* uadv, const double* vadv, const double* rho_c, const double* rho_u,
                        const double* rho_v, const double* kap, const double* zeta, const double* eta,
                        double dt, int include_adv, const double* pu, const double* pv, const double* mu,
                        const double* mv) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    Work<double>& w = c->wd;
    upload_props(c, uadv, vadv, rho_c, rho_u, rho_v, kap, zeta, eta, pu, pv, mu, mv);
    in(c, w.X, Xu, g.nu); in(c, w.X + g.nu, Xv, g.nv); in(c, w.X + g.nu + g.nv, Xp, g.np);
    const int flags = (include_adv ? MV_ADV : 0) | ((pu && pv) ? MV_PEN : 0) | ((mu && mv) ? MV_CONS : 0);
    matvec<double>(c, w, w.X, w.z, dt, flags, nullptr);
    out(c, Xu, w.X, g.nu); out(c, Xv, w.X + g.nu, g.nv); out(c, Xp, w.X + g.nu + g.nv, g.np);
    out(c, Yu, w.z, g.nu); out(c, Yv, w.z + g.nu, g.nv); out(c, Yp, w.z + g.nu + g.nv, g.np);
    sync(c);
    return 0;
  });
}

int fugu_coupled_dot(fugu_ctx* c, const double* u1, const double* v1, const double* p1, const double* u2,
                     const double* v2, const double* p2, double* dot_out) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    Work<double>& w = c->wd;
    in(c, w.z, u1, g.nu); in(c, w.z + g.nu, v1, g.nv); in(c, w.z + g.nu + g.nv, p1, g.np);
    in(c, w.q, u2, g.nu); in(c, w.q + g.nu, v2, g.nv); in(c, w.q + g.nu + g.nv, p2, g.np);
    reset_flags(c);
    cdot1<double>(c, w.z, w.q, PH_STORE);
    read_slots(c);
    if (dot_out) *dot_out = c->h_sc[S_DOT0];
    return 0;
  });
}

int fugu_coupled_dot_noalloc(const fugu_grid* grid, const double* u1, const double* v1, const double* p1,
                             const double* u2, const double* v2, const double* p2, double* dot_out,
                             int sum_order, void* stream) {
  if (!grid || !u1 || !v1 || !p1 || !u2 || !v2 || !p2 || !dot_out) { g_err = "null argument"; return -2; }
  if (grid->nh != 2) { g_err = "nh must be 2 (src/main.f90:22)"; return -2; }
  if (grid->nx < 4 || grid->ny < 4) { g_err = "grid too small (nx, ny >= 4)"; return -2; }
  if (sum_order != FUGU_SUM_REFERENCE && sum_order != FUGU_SUM_TREE) {
    g_err = "fugu_coupled_dot_noalloc: sum_order must be FUGU_SUM_REFERENCE (0) or FUGU_SUM_TREE (1)";
    return -3;
  }
  return guarded([&]() {
    const Grid g = make_grid(grid->nx, grid->ny, grid->nh, grid->lx / grid->nx, grid->ly / grid->ny,
                             grid->periodic_x ? 1 : 0, grid->periodic_y ? 1 : 0);
    const long nterms = cdot_terms(g);
    if (sum_order == FUGU_SUM_REFERENCE && nterms > NOALLOC_MAX_TERMS)
      throw Error("fugu_coupled_dot_noalloc: " + std::to_string(nterms) + " owned points exceed the reference-order "
                  "scratch limit of " + std::to_string(NOALLOC_MAX_TERMS) + " (use fugu_coupled_dot)");
    double* base = scratch_on_current_device();
    double* partials = base;
    double* ctot = partials + NOALLOC_PARTIALS;
    int* cnt = reinterpret_cast<int*>(ctot + NOALLOC_CTOT);
    double* tot = ctot + NOALLOC_CTOT + NOALLOC_CNT;
    double* terms = tot + NOALLOC_TOT;
    hipStream_t st = static_cast<hipStream_t>(stream);
    DotSplit d;
    d.a[0] = u1; d.a[1] = v1; d.a[2] = p1;
    d.b[0] = u2; d.b[1] = v2; d.b[2] = p2;
    const dim3 gr = grid_union_red(g);
    // every value the launches read is written by an earlier launch of this call: nothing to clear
    if (sum_order == FUGU_SUM_REFERENCE) {
      const Sink sk{terms, nterms, tot};
      LAUNCH(k_cdot_split, gr, dim3(TX, TY), st, g, d, sk);
      SeqArgs a;
      a.t = terms; a.n = nterms;
      a.ctot = ctot; a.cnt = cnt; a.tot = tot;
      seq_launch_chunks(a, sq_chunk(nterms), 1, Skip{-1, 0}, (const int*)nullptr, st);
      HIPCHK(hipGetLastError());
      LAUNCH(k_fold_dot_out, dim3(1), dim3(NT), st, 0, sk, dot_out);
    } else {
      const Sink sk{partials, 0, nullptr};
      LAUNCH(k_cdot_split, gr, dim3(TX, TY), st, g, d, sk);
      LAUNCH(k_fold_dot_out, dim3(1), dim3(NT), st, (int)nblocks(gr), sk, dot_out);
    }
    return 0;
  });
}

int fugu_solve_poisson_pressure(fugu_ctx* c, double* phi, const double* rhs, const double* rho_u,
                                const double* rho_v, double tol, int max_iter, int* iter_out, double* res_out) {
  NEED(c);
  return guarded([&]() {
    PrecondScope pscope(c);   // the reference solvers
    const Grid& g = c->g;
    Work<double>& w = c->wd;
    if (tol <= 0) tol = 1.0e-5;
    if (max_iter <= 0) max_iter = 150;
    in(c, c->rho_u, rho_u, g.nu); in(c, c->rho_v, rho_v, g.nv);
    bind_props_d(c, false, false);
    double* dphi = c->sp[0];
    double* drhs = c->sp[1];
    in(c, dphi, phi, g.np); in(c, drhs, rhs, g.np);
    reset_flags(c);
    set_slot(c, S_P_TOL, tol);
    pcg_setup(c, w);
    // phi ghosts: halo + Neumann mirror of the initial guess (:123-124)
    halo(c, dphi, 2);
    if (!g.per_x || !g.per_y) LAUNCH(k_mirror_p<double>, grid_union(g), dim3(TX, TY), c->s, g, dphi);
    pcg<double>(c, w, dphi, drhs, false, tol, max_iter);
    if (!g.per_x && !g.per_y) {
      cell_mean<double>(c, (const double*)dphi, (int)S_P_MEAN, 2);
      LAUNCH((k_pcg_submean<double>), grid_cells(g), dim3(TX, TY), c->s, g, dphi, (const double*)c->sc,
             (const int*)c->fl);
    }
    LAUNCH((k_zero_if<double>), dim3(nblk(g.np)), dim3(256), c->s, dphi, g.np, (const int*)c->fl, (int)F_P_DONE, 2);
    out(c, phi, dphi, g.np);
    read_flags(c);
    read_slots(c);
    if (iter_out) *iter_out = c->h_fl[F_P_ITER];
    if (res_out) *res_out = (c->h_fl[F_P_DONE] == 2) ? 0.0 : c->h_sc[S_P_RES] / c->h_sc[S_P_RHSN];
    return 0;
  });
}

int fugu_solve_pressure_mg(fugu_ctx* c, double* phi, const double* rhs, const double* rho_u,
                           const double* rho_v, int n_cycles) {
  NEED(c);
  return guarded([&]() {
    TreeSums tree(c);
    const Grid& g = c->g;
    Work<double>& w = c->wd;
    in(c, c->rho_u, rho_u, g.nu); in(c, c->rho_v, rho_v, g.nv);
    bind_props_d(c, false, false);
    double* drhs = c->sp[1];
    double* dphi = c->sp[0];
    in(c, drhs, rhs, g.np);
    HIPCHK(hipMemsetAsync(dphi, 0, g.np * sizeof(double), c->s));
    mg_setup(c, w);
    c->mg_cycles = n_cycles > 0 ? n_cycles : 1;
    mg_pressure<double>(c, w, dphi, drhs);
    if (!g.per_x && !g.per_y)
      LAUNCH((k_pcg_submean<double>), grid_cells(g), dim3(TX, TY), c->s, g, dphi, (const double*)c->sc,
             (const int*)c->fl);
    out(c, phi, dphi, g.np);
    sync(c);
    return 0;
  });
}

int fugu_solve_helmholtz_u(fugu_ctx* c, double* u_sol, const double* rhs, const double* rho_u, const double* kap,
                           const double* zeta, const double* eta, double dt, int n_sweeps, const double* pen) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    if (n_sweeps <= 0) n_sweeps = 5;
    in(c, c->rho_u, rho_u, g.nu); in(c, c->kap, kap, g.np); in(c, c->zeta, zeta, g.nw); in(c, c->eta, eta, g.nw);
    in(c, c->penu, pen, g.nu);
    double *a = c->su[0], *b = c->su[1], *r = c->su[2];
    in(c, a, u_sol, g.nu); in(c, r, rhs, g.nu);
    for (int k = 0; k < n_sweeps; ++k) {
      halo(c, a, 0);   // exchange at the top of each sweep (:281) -- ghosts persist
      LAUNCH((k_jacobi_u<double>), grid_union(g), dim3(TX, TY), c->s, g, (const double*)a, b, (const double*)r,
             (const double*)c->rho_u, (const double*)c->kap, (const double*)c->zeta, (const double*)c->eta, dt,
             pen ? (const double*)c->penu : nullptr);
      std::swap(a, b);
    }
    out(c, u_sol, a, g.nu);
    sync(c);
    return 0;
  });
}

int fugu_solve_helmholtz_v(fugu_ctx* c, double* v_sol, const double* rhs, const double* rho_v, const double* kap,
                           const double* zeta, const double* eta, double dt, int n_sweeps, const double* pen) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    if (n_sweeps <= 0) n_sweeps = 5;
    in(c, c->rho_v, rho_v, g.nv); in(c, c->kap, kap, g.np); in(c, c->zeta, zeta, g.nw); in(c, c->eta, eta, g.nw);
    in(c, c->penv, pen, g.nv);
    double *a = c->sv[0], *b = c->sv[1], *r = c->sv[2];
    in(c, a, v_sol, g.nv); in(c, r, rhs, g.nv);
    for (int k = 0; k < n_sweeps; ++k) {
      halo(c, a, 1);
      LAUNCH((k_jacobi_v<double>), grid_union(g), dim3(TX, TY), c->s, g, (const double*)a, b, (const double*)r,
             (const double*)c->rho_v, (const double*)c->kap, (const double*)c->zeta, (const double*)c->eta, dt,
             pen ? (const double*)c->penv : nullptr);
      std::swap(a, b);
    }
    out(c, v_sol, a, g.nv);
    sync(c);
    return 0;
  });
}

int fugu_apply_pcd_precond(fugu_ctx* c, const double* zu, const double* zv, const double* zp, double* ru,
                           double* rv, double* rp, const double* uadv, const double* vadv, const double* rho_u,
                           const double* rho_v, const double* kap, const double* zeta, const double* eta,
                           double dt, const double* pu, const double* pv) {
  NEED(c);
  return guarded([&]() {
    PrecondScope pscope(c);   // the reference PCD (Jacobi-PCG pressure, Jacobi velocity blocks)
    const Grid& g = c->g;
    Work<double>& w = c->wd;
    upload_props(c, uadv, vadv, nullptr, rho_u, rho_v, kap, zeta, eta, pu, pv, nullptr, nullptr);
    in(c, w.z, zu, g.nu); in(c, w.z + g.nu, zv, g.nv); in(c, w.z + g.nu + g.nv, zp, g.np);
    reset_flags(c);
    set_slot(c, S_P_TOL, 1.0e-5);
    pcg_setup(c, w);
    pcd<double>(c, w, w.z, w.q, dt);
    out(c, ru, w.q, g.nu); out(c, rv, w.q + g.nu, g.nv); out(c, rp, w.q + g.nu + g.nv, g.np);
    sync(c);
    return 0;
  });
}

static int krylov_entry(fugu_ctx* c, int order, double* Xu, double* Xv, double* Xp, const double* Bu,
                        const double* Bv, const double* Bp, const double* uadv, const double* vadv,
                        const double* rho_c, const double* rho_u, const double* rho_v, const double* kap,
                        const double* zeta, const double* eta, double dt, double tol, int max_iter,
                        int* iter_out, double* res_out, const double* lid, const double* pu, const double* pv,
                        const double* mu, const double* mv, unsigned flags) {
  NEED(c);
  return guarded([&]() {
    PrecondScope pscope(c);   // the reference solvers
    const Grid& g = c->g;
    Work<double>& w = c->wd;
    if (tol <= 0) tol = 1.0e-5;
    if (max_iter <= 0) max_iter = 100;
    upload_props(c, uadv, vadv, rho_c, rho_u, rho_v, kap, zeta, eta, pu, pv, mu, mv);
    in(c, w.X, Xu, g.nu); in(c, w.X + g.nu, Xv, g.nv); in(c, w.X + g.nu + g.nv, Xp, g.np);
    in(c, w.B, Bu, g.nu); in(c, w.B + g.nu, Bv, g.nv); in(c, w.B + g.nu + g.nv, Bp, g.np);
    set_slot(c, S_P_TOL, 1.0e-5);
    const int mvflags = MV_ADV | ((pu && pv) ? MV_PEN : 0) | ((mu && mv) ? MV_CONS : 0);
    int it = 0;
    double res = 0.0;
    if (order == 2)
      bicgstab2<double>(c, w, dt, tol, max_iter, mvflags, lid != nullptr, lid ? *lid : 0.0,
                        (flags & FUGU_B2_CONSISTENT_BCS) != 0, &it, &res);
    else
      bicgstab1<double>(c, w, dt, tol, max_iter, mvflags, lid != nullptr, lid ? *lid : 0.0, &it, &res);
    out(c, Xu, w.X, g.nu); out(c, Xv, w.X + g.nu, g.nv); out(c, Xp, w.X + g.nu + g.nv, g.np);
    sync(c);
    if (iter_out) *iter_out = it;
    if (res_out) *res_out = res;
    if (!(res == res)) return (int)FUGU_W_NONFINITE;
    return it >= max_iter ? (int)FUGU_W_MAX_ITER : 0;
  });
}

int fugu_bicgstab_solve(fugu_ctx* c, double* Xu, double* Xv, double* Xp, const double* Bu, const double* Bv,
                        const double* Bp, const double* uadv, const double* vadv, const double* rho_c,
                        const double* rho_u, const double* rho_v, const double* kap, const double* zeta,
                        const double* eta, double dt, double tol, int max_iter, int* iter_out, double* res_out,
                        const double* lid, const double* pu, const double* pv, const double* mu, const double* mv) {
  return krylov_entry(c, 1, Xu, Xv, Xp, Bu, Bv, Bp, uadv, vadv, rho_c, rho_u, rho_v, kap, zeta, eta, dt, tol,
                      max_iter, iter_out, res_out, lid, pu, pv, mu, mv, 0u);
}

int fugu_bicgstab2_solve(fugu_ctx* c, double* Xu, double* Xv, double* Xp, const double* Bu, const double* Bv,
                         const double* Bp, const double* uadv, const double* vadv, const double* rho_c,
                         const double* rho_u, const double* rho_v, const double* kap, const double* zeta,
                         const double* eta, double dt, double tol, int max_iter, int* iter_out, double* res_out,
                         const double* lid, const double* pu, const double* pv, const double* mu, const double* mv,
                         unsigned flags) {
  return krylov_entry(c, 2, Xu, Xv, Xp, Bu, Bv, Bp, uadv, vadv, rho_c, rho_u, rho_v, kap, zeta, eta, dt, tol,
                      max_iter, iter_out, res_out, lid, pu, pv, mu, mv, flags);
}

int fugu_vof_advection_tvd(fugu_ctx* c, const double* cin, double* cout, const double* u, const double* v,
                           double dt, double* fx, double* fy) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    in(c, c->sp[0], cin, g.np); in(c, c->su[0], u, g.nu); in(c, c->sv[0], v, g.nv);
    vof(c, c->sp[0], c->sp[1], c->su[0], c->sv[0], dt, fx ? c->su[1] : nullptr, fy ? c->sv[1] : nullptr);
    out(c, cout, c->sp[1], g.np);
    if (fx) out(c, fx, c->su[1], g.nu);
    if (fy) out(c, fy, c->sv[1], g.nv);
    sync(c);
    return 0;
  });
}

int fugu_svof_helmholtz_filter(fugu_ctx* c, const double* craw, double* cs, double delta, int max_iter,
                               double tol, int* iter_out) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    if (max_iter <= 0) max_iter = 100;
    if (tol <= 0) tol = 1.0e-6;
    in(c, c->sp[0], craw, g.np);
    svof(c, c->sp[0], c->sp[1], delta, max_iter, tol, iter_out);
    out(c, cs, c->sp[1], g.np);
    sync(c);
    return 0;
  });
}

int fugu_compute_interface_curvature(fugu_ctx* c, const double* cs, double* kc) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    in(c, c->sp[0], cs, g.np);
    curvature(c, c->sp[0], c->sp[1]);
    out(c, kc, c->sp[1], g.np);
    sync(c);
    return 0;
  });
}

int fugu_compute_csf_force(fugu_ctx* c, const double* craw, const double* kc, double* fu, double* fv) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    in(c, c->sp[0], craw, g.np); in(c, c->sp[1], kc, g.np);
    csf(c, c->sp[0], c->sp[1], c->su[0], c->sv[0]);
    out(c, fu, c->su[0], g.nu); out(c, fv, c->sv[0], g.nv);
    sync(c);
    return 0;
  });
}

int fugu_compute_mixture_densities(fugu_ctx* c, const double* cc, double* rho_c, double* rho_u, double* rho_v) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    in(c, c->sp[0], cc, g.np);
    mixture(c, c->sp[0]);
    out(c, rho_c, c->rho_c, g.np); out(c, rho_u, c->rho_u, g.nu); out(c, rho_v, c->rho_v, g.nv);
    sync(c);
    return 0;
  });
}

int fugu_compute_mixture_viscosities(fugu_ctx* c, const double* cc, double* kap, double* zeta, double* eta) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    in(c, c->sp[0], cc, g.np);
    mixture(c, c->sp[0]);
    out(c, kap, c->kap, g.np); out(c, zeta, c->zeta, g.nw); out(c, eta, c->eta, g.nw);
    sync(c);
    return 0;
  });
}

int fugu_compute_adaptive_dt(fugu_ctx* c, const double* u, const double* v, double dt_max, double* dt_out) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    const fugu_params& P = c->prm;
    in(c, c->su[0], u, g.nu); in(c, c->sv[0], v, g.nv);
    const dim3 gr = grid_union_red(g);
    LAUNCH(k_absmax, gr, dim3(TX, TY), c->s, g, (const double*)c->su[0], (const double*)c->sv[0], c->partials);
    LAUNCH(k_fold_absmax, dim3(1), dim3(NT), c->s, (int)nblocks(gr), (const double*)c->partials, c->sc);
    read_slots(c);
    const double um = c->h_sc[S_DOT0], vm = c->h_sc[S_DOT1];
    const double pi = 3.141592653589793238462643383279502884197;
    const double h = g.dx < g.dy ? g.dx : g.dy;
    double dtc = (um / g.dx + vm / g.dy > 1.0e-12) ? P.cfl_target / (um / g.dx + vm / g.dy) : dt_max;
    double dtl = dt_max < dtc ? dt_max : dtc;
    if (P.sigma > 1.0e-14) {
      const double cap = 0.5 * sqrt((P.rho1 + P.rho2) * (h * h * h) / (4.0 * pi * P.sigma));
      dtl = cap < dtl ? cap : dtl;
    }
    const double gm = sqrt(P.gx * P.gx + P.gy * P.gy);
    if (gm > 1.0e-14) {
      const double dg = 0.5 * sqrt(h / gm);
      dtl = dg < dtl ? dg : dtl;
    }
    if (dt_out) *dt_out = dtl;
    return 0;
  });
}

int fugu_state_upload(fugu_ctx* c, const double* u, const double* v, const double* u_prev, const double* v_prev,
                      const double* p, const double* cc, const double* psu, const double* psv) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    in(c, c->u, u, g.nu); in(c, c->v, v, g.nv); in(c, c->u_prev, u_prev, g.nu); in(c, c->v_prev, v_prev, g.nv);
    in(c, c->p, p, g.np); in(c, c->c, cc, g.np);
    c->have_pprev = false;
    c->has_solid = (psu && psv);
    if (c->has_solid) { in(c, c->phis_u, psu, g.nu); in(c, c->phis_v, psv, g.nv); }
    sync(c);
    return 0;
  });
}

int fugu_state_download(fugu_ctx* c, double* u, double* v, double* u_prev, double* v_prev, double* p, double* cc) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    out(c, u, c->u, g.nu); out(c, v, c->v, g.nv); out(c, u_prev, c->u_prev, g.nu); out(c, v_prev, c->v_prev, g.nv);
    out(c, p, c->p, g.np); out(c, cc, c->c, g.np);
    sync(c);
    return 0;
  });
}

int fugu_field_download(fugu_ctx* c, const char* name, double* o) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    struct { const char* n; double* p; size_t sz; } tab[] = {
        {"u", c->u, g.nu}, {"v", c->v, g.nv}, {"p", c->p, g.np}, {"c", c->c, g.np},
        {"u_prev", c->u_prev, g.nu}, {"v_prev", c->v_prev, g.nv}, {"rho_c", c->rho_c, g.np},
        {"rho_u", c->rho_u, g.nu}, {"rho_v", c->rho_v, g.nv}, {"kappa", c->kap, g.np},
        {"zeta", c->zeta, g.nw}, {"eta", c->eta, g.nw}, {"kappa_curv", c->kcurv, g.np},
        {"f_sigma_u", c->fu, g.nu}, {"f_sigma_v", c->fv, g.nv}, {"u_adv", c->uadv, g.nu},
        {"v_adv", c->vadv, g.nv}, {"c_smooth", c->cs, g.np}};
    for (auto& t : tab)
      if (std::strcmp(t.n, name) == 0) {
        out(c, o, t.p, t.sz);
        sync(c);
        return 0;
      }
    throw Error(std::string("unknown field ") + name);
  });
}

int fugu_checkpoint_write(fugu_ctx* c, const char* prefix, int step, double t) {
  NEED(c);
  if (!prefix) { g_err = "fugu_checkpoint_write: null prefix"; return -2; }
  return guarded([&]() {
    const Grid& g = c->g;
    // the six padded arrays, in the reference's write order (src/mod_io.f90:185)
    const double* src[6] = {c->u, c->v, c->u_prev, c->v_prev, c->p, c->c};
    const size_t cnt[6] = {g.nu, g.nv, g.nu, g.nv, g.np, g.np};
    size_t tot = 0;
    for (size_t k : cnt) tot += k;
    std::vector<double> host(tot);
    size_t off = 0;
    for (int k = 0; k < 6; ++k) {
      HIPCHK(hipMemcpyAsync(host.data() + off, src[k], cnt[k] * sizeof(double), hipMemcpyDeviceToHost, c->s));
      off += cnt[k];
    }
    sync(c);
    const std::string name = checkpoint_name(prefix);
    FILE* f = std::fopen(name.c_str(), "wb");
    if (!f) throw Error("fugu_checkpoint_write: cannot open " + name);
    unsigned char hb[24];
    header_bytes(CkptHeader{step, t, g.nx, g.ny, g.nh}, hb);
    const bool ok = std::fwrite(hb, 1, 24, f) == 24 && std::fwrite(host.data(), sizeof(double), tot, f) == tot;
    if (std::fclose(f) != 0 || !ok) throw Error("fugu_checkpoint_write: write failed on " + name);
    return 0;
  });
}

int fugu_checkpoint_read(fugu_ctx* c, const char* prefix, int* step, double* t, int* ok) {
  NEED(c);
  if (!prefix || !ok) { g_err = "fugu_checkpoint_read: null argument"; return -2; }
  *ok = 0;
  return guarded([&]() {
    const Grid& g = c->g;
    const std::string name = checkpoint_name(prefix);
    FILE* f = std::fopen(name.c_str(), "rb");
    if (!f) return 0;   // no checkpoint: ok = 0 (src/mod_io.f90:206-210)
    unsigned char hb[24];
    if (std::fread(hb, 1, 24, f) != 24) { std::fclose(f); return 0; }
    const CkptHeader h = header_from(hb);
    if (step) *step = h.step;
    if (t) *t = h.t;
    if (h.nx != g.nx || h.ny != g.ny || h.nh != g.nh) { std::fclose(f); return 0; }   // :214-218
    double* dst[6] = {c->u, c->v, c->u_prev, c->v_prev, c->p, c->c};
    const size_t cnt[6] = {g.nu, g.nv, g.nu, g.nv, g.np, g.np};
    size_t tot = 0;
    for (size_t k : cnt) tot += k;
    std::vector<double> host(tot);
    const size_t got = std::fread(host.data(), sizeof(double), tot, f);
    std::fclose(f);
    if (got != tot) throw Error("fugu_checkpoint_read: " + name + " is shorter than its header says");
    size_t off = 0;
    for (int k = 0; k < 6; ++k) {
      HIPCHK(hipMemcpyAsync(dst[k], host.data() + off, cnt[k] * sizeof(double), hipMemcpyHostToDevice, c->s));
      off += cnt[k];
    }
    c->have_pprev = false;
    mixture(c, c->c);   // the driver recomputes the mixture after a restart (src/main.f90:185-186)
    sync(c);
    *ok = 1;
    return 0;
  });
}

int fugu_write_vtk(fugu_ctx* c, const char* filename) {
  NEED(c);
  if (!filename) { g_err = "fugu_write_vtk: null filename"; return -2; }
  return guarded([&]() {
    const Grid& g = c->g;
    const int nx = g.nx, ny = g.ny;
    std::vector<double> p(g.np), cc(g.np), rc(g.np), kc(g.np), u(g.nu), v(g.nv);
    out(c, p.data(), c->p, g.np); out(c, cc.data(), c->c, g.np); out(c, rc.data(), c->rho_c, g.np);
    out(c, kc.data(), c->kcurv, g.np); out(c, u.data(), c->u, g.nu); out(c, v.data(), c->v, g.nv);
    sync(c);
    std::string s;
    s.reserve((size_t)nx * ny * 16 * 7 + 1024);
    s += "# vtk DataFile Version 3.0\nFUGU-2D Coupled Two-Phase Solver Output\nASCII\nDATASET STRUCTURED_POINTS\n";
    s += "DIMENSIONS "; fortran_i(s, nx, 8); fortran_i(s, ny, 8); fortran_i(s, 1, 8); s += "\n";
    s += "ORIGIN "; fortran_es(s, 0.5 * g.dx); fortran_es(s, 0.5 * g.dy); fortran_es(s, 0.0); s += "\n";
    s += "SPACING "; fortran_es(s, g.dx); fortran_es(s, g.dy); fortran_es(s, g.dx); s += "\n";
    s += "POINT_DATA "; fortran_i(s, (long long)nx * ny, 10); s += "\n";
    const char* names[4] = {"VOF_C", "Pressure", "Density", "Curvature"};
    const std::vector<double>* fields[4] = {&cc, &p, &rc, &kc};
    for (int k = 0; k < 4; ++k) {
      s += "SCALARS "; s += names[k]; s += " double 1\nLOOKUP_TABLE default\n";
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) { fortran_es(s, (*fields[k])[IP(g, i, j)]); s += "\n"; }
    }
    s += "VECTORS Velocity double\n";
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        // cell-centred velocity, half * (u(i-1) + u(i)) (src/mod_io.f90:48-53)
        const double uc = 0.5 * (u[IU(g, i - 1, j)] + u[IU(g, i, j)]);
        const double vc = 0.5 * (v[IV(g, i, j - 1)] + v[IV(g, i, j)]);
        fortran_es(s, uc); fortran_es(s, vc); fortran_es(s, 0.0); s += "\n";
      }
    FILE* f = std::fopen(filename, "w");
    if (!f) throw Error(std::string("fugu_write_vtk: cannot open ") + filename);
    const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
    if (std::fclose(f) != 0 || !ok) throw Error(std::string("fugu_write_vtk: write failed on ") + filename);
    return 0;
  });
}

int fugu_state_init_properties(fugu_ctx* c) {
  NEED(c);
  return guarded([&]() {
    mixture(c, c->c);
    sync(c);
    return 0;
  });
}

int fugu_kernel_timing(fugu_ctx* c, int enable) {
  NEED(c);
  return guarded([&]() {
    sync(c);
    kt_resolve(c);
    c->kt_on = enable != 0;
    if (enable)
      for (auto& kv : c->kt) { kv.second.total_ms = 0.0; kv.second.count = 0; kv.second.total_bytes = 0.0; }
    return 0;
  });
}

int fugu_kernel_time(fugu_ctx* c, const char* name, double* total_ms, long long* count, double* total_bytes) {
  NEED(c);
  return guarded([&]() {
    sync(c);
    kt_resolve(c);
    auto it = c->kt.find(name);
    if (total_ms) *total_ms = it == c->kt.end() ? 0.0 : it->second.total_ms;
    if (count) *count = it == c->kt.end() ? 0 : it->second.count;
    if (total_bytes) *total_bytes = it == c->kt.end() ? 0.0 : it->second.total_bytes;
    return 0;
  });
}

int fugu_step(fugu_ctx* c, double dt, int step_idx, const fugu_step_opts* o, fugu_step_stats* st) {
  NEED(c);
  return guarded([&]() { return step_impl(c, dt, step_idx, o, st); });
}

int fugu_advance_coupled_step(fugu_ctx* c, double* u, double* v, double* p, double* cc, double* u_prev,
                              double* v_prev, double* rho_c, double* rho_u, double* rho_v, double* kap,
                              double* zeta, double* eta, double* kcurv, double* fu, double* fv, double dt,
                              int step_idx, const fugu_step_opts* o, const double* psu, const double* psv,
                              int* iter_out, double* res_out) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    in(c, c->u, u, g.nu); in(c, c->v, v, g.nv); in(c, c->p, p, g.np); in(c, c->c, cc, g.np);
    in(c, c->u_prev, u_prev, g.nu); in(c, c->v_prev, v_prev, g.nv);
    c->has_solid = (psu && psv);
    if (c->has_solid) { in(c, c->phis_u, psu, g.nu); in(c, c->phis_v, psv, g.nv); }
    fugu_step_stats st{};
    const int rc = step_impl(c, dt, step_idx, o, &st);
    out(c, u, c->u, g.nu); out(c, v, c->v, g.nv); out(c, p, c->p, g.np); out(c, cc, c->c, g.np);
    out(c, u_prev, c->u_prev, g.nu); out(c, v_prev, c->v_prev, g.nv);
    out(c, rho_c, c->rho_c, g.np); out(c, rho_u, c->rho_u, g.nu); out(c, rho_v, c->rho_v, g.nv);
    out(c, kap, c->kap, g.np); out(c, zeta, c->zeta, g.nw); out(c, eta, c->eta, g.nw);
    out(c, kcurv, c->kcurv, g.np); out(c, fu, c->fu, g.nu); out(c, fv, c->fv, g.nv);
    sync(c);
    if (iter_out) *iter_out = st.iters;
    if (res_out) *res_out = st.res;
    return rc;
  });
}

// ---- adjoint (no counterpart in the reference) --------------------------------------
int fugu_vof_advection_transpose(fugu_ctx* c, const double* cin, const double* u, const double* v, double dt,
                                 const double* cbar_out, double* cbar_in, double* ubar, double* vbar) {
  NEED(c);
  if (!cin || !u || !v || !cbar_out) { g_err = "null input"; return -2; }
  return guarded([&]() {
    const Grid& g = c->g;
    in(c, c->sp[0], cin, g.np); in(c, c->su[0], u, g.nu); in(c, c->sv[0], v, g.nv); in(c, c->sp[1], cbar_out, g.np);
    vof_T(c, c->sp[0], c->su[0], c->sv[0], dt, c->sp[1], c->sp[2], c->su[1], c->sv[1]);
    out(c, cbar_in, c->sp[2], g.np); out(c, ubar, c->su[1], g.nu); out(c, vbar, c->sv[1], g.nv);
    sync(c);
    return 0;
  });
}

int fugu_svof_transpose(fugu_ctx* c, const double* craw, double delta, int max_iter, double tol,
                        const double* csbar, double* crawbar, int* iter_out) {
  NEED(c);
  if (!craw || !csbar) { g_err = "null input"; return -2; }
  return guarded([&]() {
    const Grid& g = c->g;
    if (max_iter <= 0) max_iter = 100;
    if (tol <= 0) tol = 1.0e-6;
    in(c, c->sp[0], craw, g.np); in(c, c->sp[2], csbar, g.np);
    int it = 0;
    svof(c, c->sp[0], c->sp[1], delta, max_iter, tol, &it);   // the forward's sweep count
    zero(c, c->sp[3], g.np);
    svof_T(c, c->sp[2], delta, it, c->sp[3]);
    out(c, crawbar, c->sp[3], g.np);
    sync(c);
    if (iter_out) *iter_out = it;
    return 0;
  });
}

int fugu_curvature_transpose(fugu_ctx* c, const double* cs, const double* kbar, double* csbar) {
  NEED(c);
  if (!cs || !kbar) { g_err = "null input"; return -2; }
  return guarded([&]() {
    const Grid& g = c->g;
    in(c, c->sp[0], cs, g.np); in(c, c->sp[1], kbar, g.np);
    zero(c, c->sp[2], g.np);
    curv_T(c, c->sp[0], c->sp[1], c->sp[2]);
    out(c, csbar, c->sp[2], g.np);
    sync(c);
    return 0;
  });
}

int fugu_csf_transpose(fugu_ctx* c, const double* craw, const double* kc, double sigma, const double* fubar,
                       const double* fvbar, double* cbar, double* kbar) {
  NEED(c);
  if (!craw || !kc || !fubar || !fvbar) { g_err = "null input"; return -2; }
  return guarded([&]() {
    const Grid& g = c->g;
    in(c, c->sp[0], craw, g.np); in(c, c->sp[1], kc, g.np); in(c, c->su[0], fubar, g.nu); in(c, c->sv[0], fvbar, g.nv);
    zero(c, c->sp[2], g.np);
    csf_T(c, c->sp[0], c->sp[1], sigma, c->su[0], c->sv[0], c->sp[2], c->sp[3]);
    out(c, cbar, c->sp[2], g.np); out(c, kbar, c->sp[3], g.np);
    sync(c);
    return 0;
  });
}

int fugu_matvec_transpose(fugu_ctx* c, const double* Lu, const double* Lv, const double* Lp, double* Ou, double* Ov,
                          double* Op, const double* uadv, const double* vadv, const double* rho_c,
                          const double* rho_u, const double* rho_v, const double* kap, const double* zeta,
                          const double* eta, double dt, int include_adv) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    Work<double>& w = c->wd;
    upload_props(c, uadv, vadv, rho_c, rho_u, rho_v, kap, zeta, eta, nullptr, nullptr, nullptr, nullptr);
    in(c, w.X, Lu, g.nu); in(c, w.X + g.nu, Lv, g.nv); in(c, w.X + g.nu + g.nv, Lp, g.np);
    matvec_T(c, w.X, w.z, dt, include_adv ? 1 : 0);
    out(c, Ou, w.z, g.nu); out(c, Ov, w.z + g.nu, g.nv); out(c, Op, w.z + g.nu + g.nv, g.np);
    sync(c);
    return 0;
  });
}

int fugu_apply_pcd_transpose(fugu_ctx* c, const double* yu, const double* yv, const double* yp, double* zu,
                             double* zv, double* zp, const double* uadv, const double* vadv, const double* rho_u,
                             const double* rho_v, const double* kap, const double* zeta, const double* eta,
                             double dt, double ptol, int pmax, double visc) {
  NEED(c);
  return guarded([&]() {
    TreeSums tree(c);
    PrecondScope pscope(c);
    const Grid& g = c->g;
    Work<double>& w = c->wd;
    upload_props(c, uadv, vadv, nullptr, rho_u, rho_v, kap, zeta, eta, nullptr, nullptr, nullptr, nullptr);
    adj_alloc(c);
    in(c, w.B, yu, g.nu); in(c, w.B + g.nu, yv, g.nv); in(c, w.B + g.nu + g.nv, yp, g.np);
    if (ptol <= 0.0) {
      c->mg_cycles = pmax > 0 ? pmax : 1;
      mg_setup(c, w);
    }
    pcd_T(c, w.B, w.X, dt, ptol, pmax > 0 ? pmax : 5000, visc);
    out(c, zu, w.X, g.nu); out(c, zv, w.X + g.nu, g.nv); out(c, zp, w.X + g.nu + g.nv, g.np);
    sync(c);
    c->mg_cycles = 1;
    return 0;
  });
}

int fugu_adjoint_solve(fugu_ctx* c, double* Lu, double* Lv, double* Lp, const double* Gu, const double* Gv,
                       const double* Gp, const double* uadv, const double* vadv, const double* rho_c,
                       const double* rho_u, const double* rho_v, const double* kap, const double* zeta,
                       const double* eta, double dt, double tol, int max_iter, int* iter_out, double* res_out) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    Work<double>& w = c->wd;
    upload_props(c, uadv, vadv, rho_c, rho_u, rho_v, kap, zeta, eta, nullptr, nullptr, nullptr, nullptr);
    adj_alloc(c);
    in(c, w.B, Gu, g.nu); in(c, w.B + g.nu, Gv, g.nv); in(c, w.B + g.nu + g.nv, Gp, g.np);
    mask_trip(c, w.B);
    double* L = c->aw.G;
    zero(c, L, c->ntrip);
    in(c, L, Lu, g.nu); in(c, L + g.nu, Lv, g.nv); in(c, L + g.nu + g.nv, Lp, g.np);
    mask_trip(c, L);
    c->use_mg = true; c->pcd_visc = 1.0; c->mg_cycles = 1; c->adjoint = true;
    int it = 0, outer = 0;
    double res = 0.0;
    try {
      solve_fgmres<double>(c, w, L, dt, tol > 0 ? tol : 1.0e-10, 0.0, max_iter > 0 ? max_iter : 400, 0, 30, MV_ADV,
                           false, 0.0, &it, &outer, &res);
    } catch (...) {
      c->adjoint = false; c->use_mg = false; c->use_vmg = false; c->pcd_visc = 0.0;
      throw;
    }
    c->adjoint = false; c->use_mg = false; c->use_vmg = false; c->pcd_visc = 0.0;
    mask_trip(c, L);
    out(c, Lu, L, g.nu); out(c, Lv, L + g.nu, g.nv); out(c, Lp, L + g.nu + g.nv, g.np);
    sync(c);
    if (iter_out) *iter_out = outer;
    if (res_out) *res_out = res;
    return (res <= (tol > 0 ? tol : 1.0e-10)) ? 0 : (int)FUGU_W_MAX_ITER;
  });
}

int fugu_cell_properties(fugu_ctx* c, const double* mu_cell, const double* rho_cell, double* rho_c, double* rho_u,
                         double* rho_v, double* kap, double* zeta, double* eta) {
  NEED(c);
  return guarded([&]() {
    const Grid& g = c->g;
    adj_alloc(c);
    if (!mu_cell || !rho_cell) throw Error("cell_properties: null input");
    in(c, c->aw.mu, mu_cell, g.np); in(c, c->aw.rho, rho_cell, g.np);
    LAUNCH(k_cell_props, grid_union(g), dim3(TX, TY), c->s, g, (const double*)c->aw.mu, (const double*)c->aw.rho,
           c->rho_c, c->rho_u, c->rho_v, c->kap, c->zeta, c->eta);
    out(c, rho_c, c->rho_c, g.np); out(c, rho_u, c->rho_u, g.nu); out(c, rho_v, c->rho_v, g.nv);
    out(c, kap, c->kap, g.np); out(c, zeta, c->zeta, g.nw); out(c, eta, c->eta, g.nw);
    sync(c);
    return 0;
  });
}

int fugu_adjoint_gradient(fugu_ctx* c, const fugu_adjoint_opts* o, const double* mu_cell, const double* rho_cell,
                          double sigma, const double* c0, const double* u0, const double* v0, double* g_mu,
                          double* g_rho, double* g_sigma, fugu_adjoint_stats* st) {
  NEED(c);
  if (!o) { g_err = "adjoint_gradient: null options"; return -2; }
  return guarded([&]() {
    try {
      return adjoint_run(c, o, mu_cell, rho_cell, sigma, c0, u0, v0, g_mu, g_rho, g_sigma, st);
    } catch (...) {
      c->adjoint = false; c->use_mg = false; c->use_vmg = false; c->pcd_visc = 0.0;
      throw;
    }
  });
}

}  // extern "C"
