#!/usr/bin/env python3
"""Coupled solves/s to a 1e-8 true relative residual, FP64 vs mixed FP32/FP64, on one MI355X;
the same at 2048 x 2048; adjoint seconds per gradient; the drop-in (the reference algorithm)
beside the reference's CPU arithmetic.

Headline workload (`value`): the rising-bubble case (domain 1x2, 512x1024 cells, bubble r = 0.2
at (0.5, 0.6) with a seeded 1 % radial perturbation over 8 Fourier modes, rho 1000/1, mu
10/0.1, sigma 2, g = -1, BDF2, dt = 2e-4).  One bench step = one full advance_coupled_step
(VOF, mixture, SVOF/CSF, RHS and one coupled (u, v, p) solve) on HBM-resident state through
the C ABI.  `value` is the FP64 form's solves/s (the reference's precision); the mixed form's
rate and the ratio are reported beside it.  Every timed step must reach the tolerance: a step
that does not is listed in `unconverged_steps` and the run exits with status 3.

Both forms solve every step's system to ||B - A X|| / ||B|| <= 1e-8 (coupled_dot norm, FP64)
with restarted flexible GMRES(m) whose every restart recomputes the true residual and updates
X in FP64 (m by the library's default rule).  The preconditioner is the block PCD with geometric-multigrid pressure and velocity
blocks and the viscous Schur correction.  FP64 form: everything FP64.  Mixed form: the Arnoldi
cycles (basis, directions, matvec operands, preconditioner) in FP32 storage, FP64 products,
sums and Hessenberg.

"m1": the same two forms on 2048 x 2048 (unit square, 256 seeded inclusions r ~ U[0.005,
0.02], carrier rho 1000 / mu 1e-2, inclusions rho 1 / mu 10, sigma 1, g -1, BDF2 dt 5e-5).
"adjoint": seconds per gradient of J (kinetic energy in the band y >= 0.9 after 16 BDF2 steps,
dt 5e-5) with respect to per-cell mu, per-cell rho and sigma on 1024 x 1024 with 64 seeded
inclusions (carrier rho 100 / mu 1e-2, inclusions rho 1 / mu 1e-2 * 10^U[0,3], sigma 0.5).
"dropin": the reference algorithm (BiCGStab + PCD with the 150-iteration PCG, the reference
driver's default) on the 256^2 static droplet (rho 1000/1, mu 1/0.01, sigma 1, BDF2 dt 1e-4),
steps 3 and 4 from the state after step 2 (both converge to the reference stopping test 1e-8):
the GPU in the reference summation order (bit-identical to the oracle, checked), the GPU in the
tree order, the FP64 fast form, and the oracle on one host core -- the `cpu_baseline`.
"adjoint2048": the combined 2048^2 workload in one call -- the mixed forward over 32 BDF2 steps
(dt 5e-5; 256 seeded inclusions r ~ U[0.005, 0.02], carrier rho 1000 / mu 1e-2, inclusions
rho 1 / mu 1e-2 * 10^U[0,3], sigma 1, g -1) with a restart state every 4 steps, then the full
adjoint gradient of the same objective (adjoint solves by the mixed form too: FP32 cycles of
the transposed operator and PCD refined in FP64 to 1e-10; each segment's forward states
recomputed from its checkpoint).
"pcg2048": the drop-in's PCG kernels (k_pcg_iter1/k_pcg_iter2, 112 B per cell and iteration)
on 2048^2, 150 iterations, timed with HIP events -- their roofline.

Prints one JSON line.  --gpus N > 1 exits: the bench measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# rising bubble (512 x 1024): the headline workload
NX, NY, LX, LY = 512, 1024, 1.0, 2.0
RHO1, RHO2, MU1, MU2, SIGMA, GY = 1000.0, 1.0, 10.0, 0.1, 2.0, -1.0
DT = 2.0e-4
TOL = 1.0e-8
SEED = 3


class Workload:
    def __init__(self, name, nx, ny, lx, ly, phys, dt, cfield):
        self.name, self.nx, self.ny, self.lx, self.ly = name, nx, ny, lx, ly
        self.phys, self.dt, self.cfield = phys, dt, cfield   # phys: rho1, rho2, mu1, mu2, sigma, gy


def bubble_c(nx, ny, lx, ly, nh=2, r0=0.2, xc=0.5, yc=0.6, seed=SEED):
    """The rising-bubble volume fraction (coupled_codefort_2d_amd.workloads.rising_bubble_c)."""
    from coupled_codefort_2d_amd import workloads
    return workloads.rising_bubble_c(nx, ny, lx, ly, nh, r0, xc, yc, seed)


def inclusions_c(n, count, rmin, rmax, seed):
    from coupled_codefort_2d_amd import workloads
    return workloads.inclusions(n, count, rmin, rmax, seed, 1.0, 1.0, 1.0, 1.0)[0]


BUBBLE = Workload("rising bubble 512x1024 (1x2 domain), rho 1000/1, mu 10/0.1, sigma 2, g -1, BDF2 dt 2e-4",
                  NX, NY, LX, LY, (RHO1, RHO2, MU1, MU2, SIGMA, GY), DT, lambda: bubble_c(NX, NY, LX, LY))
INCL2048 = Workload("2048x2048 unit square, 256 seeded inclusions r~U[0.005,0.02] (seed 5), carrier rho 1000 "
                    "mu 1e-2, inclusions rho 1 mu 10, sigma 1, g -1, BDF2 dt 5e-5",
                    2048, 2048, 1.0, 1.0, (1000.0, 1.0, 1e-2, 10.0, 1.0, -1.0), 5e-5,
                    lambda: inclusions_c(2048, 256, 0.005, 0.02, 5))


def make_ctx(abi, wl=None):
    wl = wl or BUBBLE
    r1, r2, m1, m2, sg, gy = wl.phys
    p = abi.FuguParams(r1, r2, m1, m2, sg, 0.0, gy, 2, 1.0e12, 0.5)
    ctx = abi.Context(wl.nx, wl.ny, wl.lx, wl.ly, False, False, params=p)
    nh = 2
    u = np.zeros((wl.ny + 2 * nh, wl.nx + 1 + 2 * nh))
    v = np.zeros((wl.ny + 1 + 2 * nh, wl.nx + 2 * nh))
    pp = np.zeros((wl.ny + 2 * nh, wl.nx + 2 * nh))
    ctx.upload(u, v, u.copy(), v.copy(), pp, wl.cfield())
    ctx.init_properties()
    return ctx


def run_form(abi, flags, steps, warmup, max_outer, inner_tol, mg_cycles, inner_max, restart, wl=None,
             instr_steps=None):
    """Warm-up steps, then `steps` timed steps (the rate), then `instr_steps` more steps with
    every named kernel bracketed by HIP events on the library's stream (per-kernel device time
    and algorithmic bytes; kept out of the timed steps, whose rate the events would lower)."""
    wl = wl or BUBBLE
    instr_steps = min(steps, 4) if instr_steps is None else instr_steps   # events per kernel name are pooled
    ctx = make_ctx(abi, wl)
    opts = abi.step_opts(use_bicgstab2=True, tol=TOL, max_iter=max_outer, solver_flags=flags,
                         inner_tol=inner_tol, inner_max_iter=inner_max, mg_cycles=mg_cycles, restart=restart)
    stats = []
    for s in range(1, warmup + 1):
        st, _ = ctx.step(wl.dt, s, opts)
        print(f"# warmup step {s}: outer {st.outer} inner {st.iters} res {st.res:.2e} "
              f"solve {st.solve_seconds:.3f}s", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for s in range(warmup + 1, warmup + steps + 1):
        st, _ = ctx.step(wl.dt, s, opts)
        stats.append(st)
        print(f"# step {s}: outer {st.outer} inner {st.iters} res {st.res:.2e} solve {st.solve_seconds:.3f}s "
              f"step {st.step_seconds:.3f}s", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    ctx.kernel_timing(True)
    for s in range(warmup + steps + 1, warmup + steps + instr_steps + 1):
        st, _ = ctx.step(wl.dt, s, opts)
        print(f"# instrumented step {s}: outer {st.outer} solve {st.solve_seconds:.3f}s", file=sys.stderr, flush=True)
    ktimes = {}
    for name in KERNELS:
        ms, n, nbytes = ctx.kernel_time(name)
        if n:
            ktimes[name] = (ms, n, nbytes)
    ctx.close()
    return dict(wall=wall, stats=stats, ktimes=ktimes)


# kernels timed live with HIP events on the library's stream; the library reports each
# launch's algorithmic bytes (unique fields read + written once)
KERNELS = ("pcg_iter1_f64", "pcg_iter2_f64", "pcg_iter1_f32", "pcg_iter2_f32", "matvec_f64", "matvec_f32",
           "mg_smooth_f64", "mg_smooth_f32", "mg_resid_f64", "mg_resid_f32", "mdot_f64", "mdot_f32", "maxpy_f64",
           "maxpy_f32", "maxpy_update", "maxpy_dot_f64", "maxpy_dot_f32", "mg_pre_f64", "mg_pre_f32", "mg_post_f64",
           "mg_post_f32", "mg_pre3_f64", "mg_pre3_f32", "mg_post3_f64", "mg_post3_f32", "mg_cut_f64", "mg_cut_f32",
           "mgs_pre_f64", "mgs_pre_f32", "mgs_post_f64", "mgs_post_f32", "mgs_pre3_f64", "mgs_pre3_f32",
           "mgs_post3_f64", "mgs_post3_f32", "mgs_cut_f64", "mgs_cut_f32")
DEVICE_NAMES = {"maxpy_f64": "k_maxpy<double, double>", "maxpy_f32": "k_maxpy<float, float>",
                # the matvec from LDS tiles of 8 rows (the library's default)
                "matvec_f64": "k_matvec_tile<double, 8>", "matvec_f32": "k_matvec_tile<float, 8>",
                "maxpy_update": "k_maxpy<double, float>",
                # fused multigrid levels: two red-black sweeps per half (pressure), three (velocity)
                "mg_pre_f64": "k_mg_pre<double, 2>", "mg_pre_f32": "k_mg_pre<float, 2>",
                "mg_post_f64": "k_mg_post<double, 2>", "mg_post_f32": "k_mg_post<float, 2>",
                "mg_pre3_f64": "k_mg_pre<double, 3>", "mg_pre3_f32": "k_mg_pre<float, 3>",
                "mg_post3_f64": "k_mg_post<double, 3>", "mg_post3_f32": "k_mg_post<float, 3>",
                # the cut velocity hierarchy's coarsest level: k_mg_pre with NU sweeps, no restriction
                "mg_cut_f64": "k_mg_pre<double, NU> (cut coarsest)", "mg_cut_f32": "k_mg_pre<float, NU> (cut coarsest)",
                # the register-strip level kernels (the large and the small levels)
                "mgs_pre_f64": "k_mgs_pre<double, 2>", "mgs_pre_f32": "k_mgs_pre<float, 2>",
                "mgs_post_f64": "k_mgs_post<double, 2>", "mgs_post_f32": "k_mgs_post<float, 2>",
                "mgs_pre3_f64": "k_mgs_pre<double, 3>", "mgs_pre3_f32": "k_mgs_pre<float, 3>",
                "mgs_post3_f64": "k_mgs_post<double, 3>", "mgs_post3_f32": "k_mgs_post<float, 3>",
                "mgs_cut_f64": "k_mgs_pre<double, NU> (cut coarsest)", "mgs_cut_f32": "k_mgs_pre<float, NU> (cut coarsest)"}


def device_name(name):
    if name in DEVICE_NAMES:
        return DEVICE_NAMES[name]
    base, prec = name.rsplit("_", 1)
    return f"k_{base}<{'double' if prec == 'f64' else 'float'}>"


def roofline(ktimes, tag="rising_bubble_512x1024"):
    """Roofline object for the timed kernel with the largest total device time."""
    if not ktimes:
        return None
    name, (ms, n, nbytes) = max(ktimes.items(), key=lambda kv: kv[1][0])
    avg_s = ms * 1e-3 / n
    bytes_launch = nbytes / n
    achieved = bytes_launch / avg_s / 1e9
    out = {"bound": "hbm", "kernel": device_name(name), "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
           "bytes_per_launch": round(bytes_launch), "avg_launch_us": round(avg_s * 1e6, 3), "launches": n}
    # measured HBM bytes / algorithmic bytes of this kernel on this workload (two rocprofv3
    # --pmc runs, FETCH_SIZE and WRITE_SIZE, scaled by a plain copy; hbm_traffic.py), one line
    # per kernel: "name kernel launches traffic algorithmic ratio fetch_factor write_factor"
    # the same kernel's average duration in the committed rocprofv3 --kernel-trace --stats summary
    # of this workload: HIP events around a launch also count its dispatch latency (a few us),
    # which the profiler's begin/end timestamps leave out
    trace = os.path.join(ROOT, "profiles", f"kernel_trace_{tag}.md")
    if os.path.exists(trace):
        # the kernel itself, or its instantiations with further template arguments (tile shapes)
        key = device_name(name)
        prefix = key[:-1] + ","
        tot_ms, calls = 0.0, 0
        for ln in open(trace):
            f = [x.strip() for x in ln.split("|")]
            if len(f) < 6 or not f[1].startswith("`"):
                continue
            kname = f[1].strip("`")
            if kname == key or kname.startswith(prefix):
                tot_ms += float(f[3]); calls += int(f[2])
        if calls:
            pavg = tot_ms * 1e3 / calls
            out["profile_avg_launch_us"] = round(pavg, 3)
            out["frac_at_profile_avg"] = round(bytes_launch / (pavg * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
            out["profile_source"] = os.path.relpath(trace, ROOT)
    prof = os.path.join(ROOT, "profiles", f"hbm_traffic_{tag}.txt")
    if os.path.exists(prof):
        for ln in open(prof):
            f = ln.split()
            if f and not ln.startswith("#") and f[0] == name and len(f) >= 6:
                ratio = float(f[-3])
                out["traffic"] = round(ratio * bytes_launch)
                out["traffic_ratio"] = ratio
                out["traffic_source"] = os.path.relpath(prof, ROOT)
    return out


# the drop-in workload: config 2's static droplet, steps 3 and 4 from the state after step 2
DROP_N, DROP_DT, DROP_STEPS = 256, 1.0e-4, (3, 4)
DROP_PHYS = (1000.0, 1.0, 1.0, 0.01, 1.0, 0.0)   # rho1, rho2, mu1, mu2, sigma, gy


def droplet_c(n, nh=2):
    from coupled_codefort_2d_amd import workloads
    X, Y = workloads.cell_centres(n, n, 1.0, 1.0, nh)
    return (np.hypot(X - 0.5, Y - 0.5) <= 0.25).astype(np.float64)


def run_dropin(abi, no_cpu=False):
    """The reference algorithm on the 256^2 static droplet: steps 1-2 on the GPU (tree order),
    then steps 3 and 4 from that state by (a) the GPU in the reference summation order, (b) the
    GPU in the tree order, (c) the FP64 fast form, and (d) the oracle on one host core (the
    cpu_baseline), whose states (a) must equal bit for bit."""
    n, nh = DROP_N, 2
    r1, r2, m1, m2, sg, gy = DROP_PHYS
    prm = abi.FuguParams(r1, r2, m1, m2, sg, 0.0, gy, 2, 1.0e12, 0.5)
    ref_opts = abi.step_opts(use_bicgstab2=False, tol=TOL, max_iter=400)
    shapes = [(n + 2 * nh, n + 1 + 2 * nh), (n + 1 + 2 * nh, n + 2 * nh)] * 2 + [(n + 2 * nh, n + 2 * nh)] * 2
    ctx = abi.Context(n, n, params=prm, sum_order=abi.SUM_TREE)
    z = [np.zeros(sh) for sh in shapes[:5]]
    ctx.upload(*z, droplet_c(n))
    ctx.init_properties()
    for s in (1, 2):
        st, rc = ctx.step(DROP_DT, s, ref_opts)
        print(f"# drop-in: step {s} (GPU tree order) iters {st.iters} res {st.res:.2e}", file=sys.stderr, flush=True)
    S2 = [np.zeros(sh) for sh in shapes]   # u, v, u_prev, v_prev, p, c
    ctx.download(*S2)
    ctx.close()
    base = abi.SOLVE_REFINED | abi.PRECOND_MG | abi.PRECOND_VISC | abi.PRECOND_VMG
    runs = {}
    states_ref = {}
    for name, order, opts in (("gpu_reference_order", abi.SUM_REFERENCE, ref_opts),
                              ("gpu_tree_order", abi.SUM_TREE, ref_opts),
                              ("gpu_fp64_fast_form", abi.SUM_TREE,
                               abi.step_opts(tol=TOL, max_iter=400, solver_flags=base, inner_max_iter=-1))):
        c2 = abi.Context(n, n, params=prm, sum_order=order)
        c2.upload(*S2)
        c2.init_properties()
        per = []
        for s in DROP_STEPS:
            t0 = time.perf_counter()
            st, rc = c2.step(DROP_DT, s, opts)
            per.append(dict(step=s, seconds=round(time.perf_counter() - t0, 4), iters=st.iters, outer=st.outer,
                            res=float(f"{st.res:.3e}"), rc=rc))
            if order == abi.SUM_REFERENCE and opts is ref_opts:
                out = [np.zeros(sh) for sh in shapes]
                c2.download(*out)
                states_ref[s] = out
            print(f"# drop-in {name}: step {s} iters {st.iters}/{st.outer} res {st.res:.2e} "
                  f"{per[-1]['seconds']:.3f}s", file=sys.stderr, flush=True)
        c2.close()
        tot = sum(p["seconds"] for p in per)
        runs[name] = {"solves_per_s": round(len(per) / tot, 4), "steps": per}
    out = {"workload": f"{n}x{n} static droplet R 0.25, rho 1000/1, mu 1/0.01, sigma 1, g 0, BDF2 dt {DROP_DT}; "
                       f"steps {DROP_STEPS[0]}-{DROP_STEPS[-1]} from the state after step 2",
           "solver": "reference BiCGStab + PCD (Jacobi-PCG 1e-5/150, 3 damped-Jacobi sweeps per velocity block), "
                     "reference stopping test 1e-8; fast form: FP64 FGMRES + multigrid PCD to true residual 1e-8",
           **runs}
    if not no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        O.set_sum_mode(0)
        g = O.Grid(n, n)
        ph = O.OgPhys(r1, r2, m1, m2, sg, 0.0, gy, 2, 1.0e12, 0.5)
        st = O.State(g)
        st.u[...], st.v[...], st.u_prev[...], st.v_prev[...], st.p[...], st.c[...] = S2
        per, same = [], True
        for s in DROP_STEPS:
            t0 = time.perf_counter()
            it, res = O.advance_step(st, ph, DROP_DT, s, use_bicgstab2=False, tol=TOL, max_iter=400)
            per.append(dict(step=s, seconds=round(time.perf_counter() - t0, 3), iters=it, res=float(f"{res:.3e}")))
            gpu = states_ref[s]
            same = same and all(np.array_equal(a, b) for a, b in
                                zip(gpu, (st.u, st.v, st.u_prev, st.v_prev, st.p, st.c)))
            print(f"# drop-in oracle (1 core): step {s} iters {it} res {res:.2e} {per[-1]['seconds']:.2f}s; "
                  f"GPU reference-order state bit-identical: {same}", file=sys.stderr, flush=True)
        tot = sum(p["seconds"] for p in per)
        out["cpu_oracle_1core"] = {"solves_per_s": round(len(per) / tot, 5), "steps": per}
        out["gpu_reference_order_bit_identical_to_cpu"] = same
        out["gpu_reference_order_speedup_vs_cpu"] = round(runs["gpu_reference_order"]["solves_per_s"] * tot / len(per), 2)
    return out


def run_physics(abi):
    """Configs 1 and 2 at full length by the FP64 fast form, with time steps/s and the physics the
    reference's validation programs print (coupled_codefort_2d_amd.diagnostics):
    config 1: lid cavity 128^2, rho 1, mu 1e-3 (Re 1000), BDF1 dt 1e-3, 100 steps, solves to 1e-10;
    config 2: static droplet 256^2, rho 1000/1, mu 1/0.01, sigma 1, BDF2 dt 1e-4, 50 steps, 1e-8."""
    from coupled_codefort_2d_amd import diagnostics as D
    from coupled_codefort_2d_amd import workloads
    base = abi.SOLVE_REFINED | abi.PRECOND_MG | abi.PRECOND_VISC | abi.PRECOND_VMG
    out = {}
    # config 1
    n = 128
    ctx = abi.Context(n, n, params=abi.FuguParams(1.0, 1.0, 1e-3, 1e-3, 0.0, 0.0, 0.0, 1, 1.0e12, 0.5))
    u = np.zeros((n + 4, n + 5)); v = np.zeros((n + 5, n + 4)); p = np.zeros((n + 4, n + 4))
    ctx.upload(u, v, u.copy(), v.copy(), p, np.ones((n + 4, n + 4)))
    ctx.init_properties()
    opts = abi.step_opts(tol=1e-10, max_iter=600, solver_flags=base, inner_max_iter=-1, u_lid=1.0)
    res, t0 = [], time.perf_counter()
    for s_ in range(1, 101):
        st, rc = ctx.step(1e-3, s_, opts)
        res.append(st.res)
    wall = time.perf_counter() - t0
    u, v = ctx.field("u", u), ctx.field("v", v)
    ctx.close()
    out["config1"] = {"workload": "lid cavity 128x128, rho 1, mu 1e-3 (Re 1000), BDF1 dt 1e-3, 100 steps, FP64 "
                                  "fast form to true residual 1e-10",
                      "time_steps_per_s": round(100 / wall, 2), "max_res": float(f"{max(res):.3e}"),
                      "centre_line_extrema_t0.1": D.cavity_extrema(u, v, n, n),
                      "div_max": D.div_max(u, v, n, n, 1.0, 1.0)}
    # config 2
    n, R, sigma = 256, 0.25, 1.0
    c0 = droplet_c(n)
    ctx = abi.Context(n, n, params=abi.FuguParams(1000.0, 1.0, 1.0, 0.01, sigma, 0.0, 0.0, 2, 1.0e12, 0.5))
    u = np.zeros((n + 4, n + 5)); v = np.zeros((n + 5, n + 4)); p = np.zeros((n + 4, n + 4))
    ctx.upload(u, v, u.copy(), v.copy(), p, c0)
    ctx.init_properties()
    opts = abi.step_opts(tol=TOL, max_iter=400, solver_flags=base, inner_max_iter=-1)
    res, worst, clamp, t0 = [], 0.0, 0.0, time.perf_counter()
    for s_ in range(1, 51):
        st, rc = ctx.step(DROP_DT, s_, opts)
        res.append(st.res)
        worst = max(worst, abs(st.vol_change - st.vol_clamp) / st.vol_before)
        clamp = max(clamp, abs(st.vol_clamp) / st.vol_before)
    wall = time.perf_counter() - t0
    u, v, p = ctx.field("u", u), ctx.field("v", v), ctx.field("p", p)
    c, kc = ctx.field("c", np.zeros_like(c0)), ctx.field("kappa_curv", np.zeros_like(c0))
    ctx.close()
    out["config2"] = {"workload": "static droplet 256x256, R 0.25, rho 1000/1, mu 1/0.01, sigma 1, BDF2 dt 1e-4, "
                                  "50 steps, FP64 fast form to true residual 1e-8",
                      "time_steps_per_s": round(50 / wall, 2), "max_res": float(f"{max(res):.3e}"),
                      "laplace_jump": D.laplace_jump(p, n, n, 1.0, 1.0, R, sigma),
                      "curvature": D.curvature_mean(kc, n, n, 1.0, 1.0, R),
                      "spurious_velocity_max": D.spurious_velocity(u, v, n, n),
                      "mass_error": D.mass_error(c, c0, n, n), "div_max": D.div_max(u, v, n, n, 1.0, 1.0),
                      "max_step_conservation_error_rel": worst, "max_step_clamp_volume_rel": clamp}
    for k in ("config1", "config2"):
        print(f"# physics {k}: " + json.dumps({a: b for a, b in out[k].items() if a != "workload"}),
              file=sys.stderr, flush=True)
    return out


def run_pcg2048(abi, n=2048, iters=150):
    """The drop-in's PCG iteration kernels on 2048^2 (the 256-inclusion density field, rho 1000/1),
    150 iterations from zero, in both summation orders, timed with HIP events (warm-up call first)."""
    from coupled_codefort_2d_amd import workloads
    c = workloads.inclusions(n, 256, 0.005, 0.02, 5, 1.0, 1.0, 1.0, 1.0)[0]
    prm = abi.FuguParams(1000.0, 1.0, 1e-2, 10.0, 1.0, 0.0, -1.0, 2, 1.0e12, 0.5)
    rng = np.random.default_rng(7)
    rhs = rng.standard_normal((n + 4, n + 4))
    res = {}
    for order, tag in ((abi.SUM_REFERENCE, "reference_order"), (abi.SUM_TREE, "tree_order")):
        ctx = abi.Context(n, n, params=prm, sum_order=order)
        rho_c, rho_u, rho_v = np.zeros((n + 4, n + 4)), np.zeros((n + 4, n + 5)), np.zeros((n + 5, n + 4))
        kap, zeta, eta = np.zeros((n + 4, n + 4)), np.zeros((n + 5, n + 5)), np.zeros((n + 5, n + 5))
        ctx.mixture(c, rho_c, rho_u, rho_v, kap, zeta, eta)
        phi = np.zeros((n + 4, n + 4))
        ctx.poisson_pcg(phi, rhs, rho_u, rho_v, tol=1e-30, max_iter=iters)
        ctx.kernel_timing(True)
        phi[...] = 0.0
        t0 = time.perf_counter()
        it, r = ctx.poisson_pcg(phi, rhs, rho_u, rho_v, tol=1e-30, max_iter=iters)
        wall = time.perf_counter() - t0
        k = {}
        for name in ("pcg_iter1_f64", "pcg_iter2_f64"):
            ms, cnt, nb = ctx.kernel_time(name)
            if cnt:
                avg = ms * 1e-3 / cnt
                k[name] = {"kernel": device_name(name), "launches": cnt, "avg_launch_us": round(avg * 1e6, 3),
                           "bytes_per_launch": round(nb / cnt), "GBps": round(nb / cnt / avg / 1e9, 1),
                           "frac": round(nb / cnt / avg / 1e9 / HBM_PEAK_GBS, 4)}
        ctx.close()
        res[tag] = {"iterations": it, "wall_seconds_incl_host_copies": round(wall, 4), "kernels": k}
        print(f"# pcg {n}^2 {tag}: {it} iterations, {wall:.3f}s; " +
              ", ".join(f"{a} {b['avg_launch_us']}us {b['GBps']}GB/s" for a, b in k.items()), file=sys.stderr, flush=True)
    return res


def pcg_roofline(pcg, tag="pcg_2048x2048"):
    """Roofline object of the drop-in's dominant PCG kernel (reference order): HIP-event figures and
    the committed profile's average for the same kernel."""
    k = pcg["reference_order"]["kernels"]
    if not k:
        return None
    name, v = max(k.items(), key=lambda kv: kv[1]["avg_launch_us"] * kv[1]["launches"])
    out = {"bound": "hbm", "kernel": v["kernel"], "achieved": v["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": v["frac"], "traffic": None, "bytes_per_launch": v["bytes_per_launch"],
           "avg_launch_us": v["avg_launch_us"], "launches": v["launches"]}
    trace = os.path.join(ROOT, "profiles", f"kernel_trace_{tag}.md")
    if os.path.exists(trace):
        for ln in open(trace):
            f = [x.strip() for x in ln.split("|")]
            if len(f) >= 6 and f[1].strip("`") == v["kernel"]:
                pavg = float(f[4])
                out["profile_avg_launch_us"] = pavg
                out["frac_at_profile_avg"] = round(v["bytes_per_launch"] / (pavg * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
                out["profile_source"] = os.path.relpath(trace, ROOT)
    # measured HBM bytes / algorithmic bytes of the same kernel (two rocprofv3 --pmc runs)
    prof = os.path.join(ROOT, "profiles", f"hbm_traffic_{tag}.txt")
    if os.path.exists(prof):
        for ln in open(prof):
            f = ln.split()
            if f and not ln.startswith("#") and f[0] == name and len(f) >= 6:
                ratio = float(f[-3])
                out["traffic"] = round(ratio * v["bytes_per_launch"])
                out["traffic_ratio"] = ratio
                out["traffic_source"] = os.path.relpath(prof, ROOT)
    return out


def timed_rate(ktimes):
    """Algorithmic bytes / device time summed over every timed kernel (GB/s)."""
    ms = sum(v[0] for v in ktimes.values())
    nb = sum(v[2] for v in ktimes.values())
    return round(nb / (ms * 1e-3) / 1e9, 1) if ms > 0 else None


def run_adjoint(abi, n=1024, nsteps=16, dt=5e-5, count=64, seed=4, rmin=0.01, rmax=0.04, rho_c=100.0,
                mu_ratio=1e3, sigma=0.5, checkpoint_every=4, mixed=False, also_resident=False, mg_cycles=1,
                moving=False, restart=30, mixed_adjoint=False):
    """Seconds per gradient on an adjoint workload: the forward horizon with checkpoints (restart
    states every `checkpoint_every` steps), the reverse sweep with each segment's forward states
    recomputed, and the property-path gradients; after a 2-step warm-up call on the same
    context.  also_resident: the same gradient with every step's state kept in HBM (no
    recomputation) is timed too, and reported as seconds_per_gradient when it agrees.  moving: the
    interface motion differentiated too (FUGU_ADJ_MOVING_INTERFACE) instead of held frozen.
    mixed / mixed_adjoint: the forward / adjoint solves by FP32 Arnoldi cycles refined in FP64 to
    the same true residual."""
    from coupled_codefort_2d_amd import workloads
    c, mu, rho, _ = workloads.inclusions(n, count, rmin, rmax, seed, rho_c, 1e-2, 1.0, mu_ratio)
    ctx = abi.Context(n, n, 1.0, 1.0, False, False,
                      params=abi.FuguParams(1.0, 1.0, 1.0, 1.0, sigma, 0.0, -1.0, 2, 1e12, 0.5))
    u0 = np.zeros((n + 4, n + 5))
    v0 = np.zeros((n + 5, n + 4))
    gm, gr = np.zeros_like(mu), np.zeros_like(rho)
    flags = (abi.ADJ_PRECOND_VMG | (abi.ADJ_MIXED_FORWARD if mixed else 0) | (abi.ADJ_MOVING_INTERFACE if moving else 0) |
             (abi.ADJ_MIXED_ADJOINT if mixed_adjoint else 0))
    ctx.adjoint_gradient(abi.FuguAdjointOpts(2, dt, 0.9, 1e-10, 3000, restart, mg_cycles, flags), mu, rho, sigma, c, u0, v0,
                         gm, gr)
    opts = abi.FuguAdjointOpts(nsteps, dt, 0.9, 1e-10, 3000, restart, mg_cycles, flags, checkpoint_every)
    t0 = time.perf_counter()
    J, gs, st, rc = ctx.adjoint_gradient(opts, mu, rho, sigma, c, u0, v0, gm, gr)
    wall = time.perf_counter() - t0
    res = None
    if also_resident:
        gm2, gr2 = np.zeros_like(mu), np.zeros_like(rho)
        t0 = time.perf_counter()
        J2, gs2, st2, rc2 = ctx.adjoint_gradient(abi.FuguAdjointOpts(nsteps, dt, 0.9, 1e-10, 3000, restart, mg_cycles, flags, 0),
                                                 mu, rho, sigma, c, u0, v0, gm2, gr2)
        w2 = time.perf_counter() - t0
        # J bit for bit; the coefficient gradients accumulate with atomics (order not fixed): to rounding
        rel = lambda a, b: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
        same = bool(rc2 == 0 and J2 == J and abs(gs2 - gs) <= 1e-12 * abs(gs) and rel(gm2, gm) <= 1e-12 and
                    rel(gr2, gr) <= 1e-12)
        res = {"seconds_per_gradient": round(w2, 3), "forward_seconds": round(st2.forward_seconds, 3),
               "reverse_seconds": round(st2.reverse_seconds, 3), "agrees_with_checkpointed": same,
               "grad_rel_diff": max(rel(gm2, gm), rel(gr2, gr))}
        print(f"# adjoint resident horizon: {w2:.2f}s, agrees with the checkpointed run (J bitwise, gradients "
              f"to 1e-12): {same}", file=sys.stderr, flush=True)
    ctx.close()
    print(f"# adjoint {n}x{n}, {nsteps} steps{' (moving interface)' if moving else ''}: J {J:.6e} dJ/dsigma {gs:.6e} "
          f"forward {st.forward_seconds:.2f}s "
          f"reverse {st.reverse_seconds:.2f}s iters {st.forward_iters}/{st.adjoint_iters}", file=sys.stderr, flush=True)
    return {"workload": f"{n}x{n} unit square, {count} seeded inclusions r~U[{rmin},{rmax}] (seed {seed}), carrier rho "
                        f"{rho_c:g} mu 1e-2, inclusions rho 1 mu 1e-2*10^U[0,{np.log10(mu_ratio):g}], sigma {sigma}, "
                        f"g -1; J = kinetic energy in y >= 0.9 after {nsteps} BDF2 steps (dt {dt}); gradients w.r.t. "
                        f"per-cell mu, rho and sigma; checkpoints every {checkpoint_every} steps",
            "seconds_per_gradient": round(wall, 3), "forward_seconds": round(st.forward_seconds, 3),
            "reverse_seconds_incl_recompute": round(st.reverse_seconds, 3), "forward_iters": st.forward_iters,
            "adjoint_iters": st.adjoint_iters, "max_forward_res": float(f"{st.max_forward_res:.3e}"),
            "max_adjoint_res": float(f"{st.max_adjoint_res:.3e}"), "rc": rc, "J": J, "dJ_dsigma": gs,
            "interface": "moving (VOF, SVOF, curvature and CSF transposed)" if moving else "frozen",
            "solves": ("mixed (FP32 cycles, FP64 restarts)" if mixed else "FP64") + " forward and " +
                      ("mixed" if mixed_adjoint else "FP64") + f" adjoint FGMRES({restart}) to 1e-10, PCD with multigrid "
                      f"pressure ({mg_cycles} V-cycle(s)) and velocity blocks",
            **({"resident_horizon": res, "checkpointed_seconds_per_gradient": round(wall, 3),
                "seconds_per_gradient": res["seconds_per_gradient"] if res["agrees_with_checkpointed"]
                else round(wall, 3)} if res else {})}


def copy_probe(mib):
    """A plain streaming copy of known size (torch elementwise kernel, FP64): the calibration
    for the HBM byte counters (FETCH_SIZE, WRITE_SIZE) read on the solver's kernels."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(f"# copy probe: {mib} MiB read + {mib} MiB written per launch, 3 launches", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the bubble's first steps are a start-up transient (35, 29, 24 FGMRES steps, then ~21): the
    # headline rate is taken after it
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--max-outer", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,mixed")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run a plain FP64 copy of this size (counter calibration)")
    ap.add_argument("--inner-tol", type=float, default=1e-3)
    ap.add_argument("--mg-cycles", type=int, default=0,
                    help="pressure V-cycles per PCD; 0: the library's default (2 from 2^22 cells, else 1)")
    ap.add_argument("--inner-max", type=int, default=-1, help="inner BiCGStab(2) cap; -1: PCD only")
    ap.add_argument("--restart", type=int, default=0,
                    help="FGMRES restart length; 0: the library's default (10 after a solve of 16-30 steps, "
                         "else 20 up to 2^21 cells, 30 above)")
    ap.add_argument("--sections", default="headline,m1,adjoint,adjoint2048,dropin,pcg2048,physics",
                    help="headline (rising bubble, value), m1 (2048 x 2048 FP64 vs mixed), adjoint (s/grad, "
                         "1024^2, 16 steps), adjoint2048 (2048^2, 256 inclusions, 32 steps, mixed forward and adjoint), "
                         "dropin (reference algorithm on the 256^2 droplet, GPU vs the oracle on one core: "
                         "cpu_baseline), pcg2048 (the drop-in's PCG kernels on 2048^2), physics (configs 1 and 2 "
                         "at full length: time steps/s and the reference's validation outputs)")
    ap.add_argument("--no-m1", action="store_true", help="skip the 2048 x 2048 FP64-vs-mixed section")
    ap.add_argument("--m1-steps", type=int, default=3)
    ap.add_argument("--m1-warmup", type=int, default=1)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint seconds-per-gradient section")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(the solver runs one device-resident domain per process)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
    from coupled_codefort_2d_amd import abi
    abi.load()
    sections = set(args.sections.split(","))
    if args.no_m1:
        sections.discard("m1")
    if args.no_adjoint:
        sections.discard("adjoint")
    forms = {}
    base = abi.SOLVE_REFINED | abi.PRECOND_MG | abi.PRECOND_VISC | abi.PRECOND_VMG
    kw = dict(steps=args.steps, warmup=args.warmup, max_outer=args.max_outer, inner_tol=args.inner_tol,
              mg_cycles=args.mg_cycles, inner_max=args.inner_max, restart=args.restart)
    if "headline" in sections:
        if "fp64" in args.forms:
            forms["fp64"] = run_form(abi, base, **kw)
        if "mixed" in args.forms:
            forms["mixed"] = run_form(abi, base | abi.SOLVE_MIXED, **kw)
    m1 = {}
    if "m1" in sections:
        kw1 = dict(kw, steps=args.m1_steps, warmup=args.m1_warmup, wl=INCL2048)
        m1["fp64"] = run_form(abi, base, **kw1)
        m1["mixed"] = run_form(abi, base | abi.SOLVE_MIXED, **kw1)
    # the 1024^2 horizon (16 states, < 1 GB) is held in HBM; its checkpointed run is timed beside it.
    # Two pressure V-cycles per PCD: measured 3.2 s/grad against 3.7 s with one; four gain < 1 %.
    adj = run_adjoint(abi, also_resident=True, mg_cycles=2) if "adjoint" in sections else None
    if adj is not None:
        mv = run_adjoint(abi, mg_cycles=2, moving=True)
        adj["moving_interface"] = {k: mv[k] for k in ("seconds_per_gradient", "forward_seconds",
                                                      "reverse_seconds_incl_recompute", "adjoint_iters", "rc",
                                                      "dJ_dsigma", "interface")}
    adj2048 = None
    if "adjoint2048" in sections:
        # pressure V-cycles per PCD (the 1e-10 adjoint solves lean on the pressure block): with the
        # register-strip level kernels 40.1 s/grad with 4, 39.3 with 6, 36.7 with 8, 36.9 with 12
        # (FP64 adjoint solves); the mixed adjoint solves: 35.5 -> 28.8 s/grad, the same FGMRES steps
        adj2048 = run_adjoint(abi, n=2048, nsteps=32, count=256, seed=5, rmin=0.005, rmax=0.02, rho_c=1000.0,
                              sigma=1.0, mixed=True, mg_cycles=8, mixed_adjoint=True)
    dropin = run_dropin(abi, no_cpu=args.no_cpu) if "dropin" in sections else None
    pcg = run_pcg2048(abi) if "pcg2048" in sections else None
    physics = run_physics(abi) if "physics" in sections else None

    def summary(f):
        sts = f["stats"]
        solve_s = sum(s.solve_seconds for s in sts)
        return {"solves_per_s": round(len(sts) / f["wall"], 4),
                "solve_only_per_s": round(len(sts) / solve_s, 4) if solve_s > 0 else None,
                "ms_per_step": round(1e3 * f["wall"] / len(sts), 3),
                "inner_iters": [s.iters for s in sts], "outer": [s.outer for s in sts],
                "res": [float(f"{s.res:.3e}") for s in sts]}

    # value: the FP64 form (the reference's precision); the mixed form beside it
    head = "fp64" if "fp64" in forms else ("mixed" if "mixed" in forms else None)
    rep = {k: summary(v) for k, v in forms.items()}
    if head is None:   # headline section not run (profiling another section): no value
        rep[None] = {"solves_per_s": None, "ms_per_step": None}
    all_k = {}
    for f in forms.values():
        for k, v in f["ktimes"].items():   # names carry the precision; sum over forms that share one
            o = all_k.get(k, (0.0, 0, 0.0))
            all_k[k] = (o[0] + v[0], o[1] + v[1], o[2] + v[2])
    line = {
        "metric": "coupled solves/s to 1e-8 residual, FP64 vs mixed; HBM GB/s % peak; adjoint s/grad",
        "value": rep[head]["solves_per_s"],
        "unit": "coupled solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": rep[head]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "rising bubble 512x1024 (1x2 domain), rho 1000/1, mu 10/0.1, sigma 2, g -1, "
                               "BDF2 dt 2e-4, solves to true relative residual 1e-8",
                   "nx": NX, "ny": NY, "form": head},
        "forms": {k: v for k, v in rep.items() if k is not None},
        "form_notes": {"fp64": "FGMRES(m) in FP64 + FP64 PCD (multigrid pressure and velocity blocks), "
                               "true residual 1e-8",
                       "mixed": "FGMRES(m) Arnoldi cycles in FP32 storage (FP64 dots and Hessenberg) + FP32 PCD; "
                                "true residual and update in FP64 at every restart; true residual 1e-8",
                       "restart": (f"m = {args.restart}" if args.restart > 0 else
                                   "the library's default: m = 10 after a solve of 16-30 FGMRES steps on the "
                                   "context, else 20 up to 2^21 cells and 30 above (the bubble's solves take "
                                   "about 20 steps: m = 10 from the second step on; 2048x2048: m = 30)")},
        "roofline": roofline(forms[head]["ktimes"]) if head in forms else None,   # the headline form's kernels
        "kernels": {k: {"avg_us": round(1e3 * v[0] / v[1], 2), "launches": v[1], "total_ms": round(v[0], 2),
                        "GBps": round(v[2] / (v[0] * 1e-3) / 1e9, 1)} for k, v in all_k.items()},
    }
    if "fp64" in rep and "mixed" in rep and rep["fp64"]["solves_per_s"]:
        line["mixed_solves_per_s"] = rep["mixed"]["solves_per_s"]
        line["mixed_speedup_vs_fp64"] = round(rep["mixed"]["solves_per_s"] / rep["fp64"]["solves_per_s"], 3)
    line["sample_note"] = (f"{args.steps} timed steps ({args.warmup + 1}-{args.warmup + args.steps}) of the "
                           "rising bubble's 500 after the start-up transient")
    if m1:
        r1 = {k: summary(v) for k, v in m1.items()}
        line["m1"] = {"workload": INCL2048.name + ", solves to true relative residual 1e-8",
                      "forms": r1,
                      "mixed_speedup_vs_fp64": round(r1["mixed"]["solves_per_s"] / r1["fp64"]["solves_per_s"], 3),
                      "roofline": {k: roofline(v["ktimes"], "inclusions_2048x2048") for k, v in m1.items()},
                      "timed_kernels_GBps": {k: timed_rate(v["ktimes"]) for k, v in m1.items()}}
    if adj:
        line["adjoint"] = adj
    if adj2048:
        line["adjoint_2048"] = adj2048
    if dropin:
        line["dropin"] = dropin
        if "cpu_oracle_1core" in dropin:
            cpu = dropin["cpu_oracle_1core"]
            line["cpu_baseline"] = {
                "value": cpu["solves_per_s"], "unit": "coupled solves/s", "cores": 1, "kind": "port",
                "sample": "the reference algorithm (BiCGStab + PCD, reference stopping test 1e-8) on the 256x256 static "
                          f"droplet, steps {DROP_STEPS[0]}-{DROP_STEPS[-1]} from the GPU's state after step 2, on the "
                          "oracle (C restatement of the reference, gcc -O2, 1 thread, the reference's sequential "
                          f"sums): {[p['iters'] for p in cpu['steps']]} BiCGStab iterations in "
                          f"{sum(p['seconds'] for p in cpu['steps']):.1f} s; the GPU drop-in on the same steps "
                          f"(reference summation order, bit-identical states: "
                          f"{dropin.get('gpu_reference_order_bit_identical_to_cpu')}) runs "
                          f"{dropin['gpu_reference_order']['solves_per_s']} solves/s",
                "workload_differs_from_value": "the rising bubble's reference solver does not converge (the "
                                               "headline uses the fast forms), so the baseline is a workload where "
                                               "the reference path converges"}
    if physics:
        line["physics"] = physics
    if pcg:
        line["pcg2048"] = pcg
        line["dropin_roofline"] = pcg_roofline(pcg)
    # every timed step must have reached the tolerance
    bad = []
    for tag, fs in (("headline", forms), ("m1", m1)):
        for k, f in fs.items():
            bad += [f"{tag}.{k}.step{i}" for i, st in enumerate(f["stats"]) if not (st.res <= TOL)]
    if dropin:
        bad += [f"dropin.{k}.step{p['step']}" for k in ("gpu_reference_order", "gpu_tree_order", "gpu_fp64_fast_form")
                for p in dropin[k]["steps"] if not (p["res"] <= TOL)]
    line["unconverged_steps"] = bad
    print(json.dumps(line), flush=True)
    if bad:
        print(f"bench: {len(bad)} timed step(s) did not reach the tolerance {TOL}: {bad}", file=sys.stderr, flush=True)
        sys.exit(3)


if __name__ == "__main__":
    main()
