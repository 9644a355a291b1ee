#!/usr/bin/env python3
"""Benchmark of the BSSN RHS + RK4 hot path on one MI355X.

Workload: the two-puncture configuration on one 128^3 patch (+3 ghost
zones) over [-16,16]^3, GAUGE 0, KO eps 0.02, dt = h/4, frozen ghost zones.
A "step" is one RK4 step as the reference's Step takes it: 4 right-hand-side
evaluations of every interior point, the predictor's forming the seven
constraint residuals (co = 0), run by the fused RHS + RK4 kernel of
libamssbssn.so, plus the seven constraint L2 norms of every step.

Prints one JSON line.  value = interior points x RHS evaluations / wall
time of the timed steps (Mpoints/s), inputs resident in HBM.  The roofline
object uses the average duration of the fused kernel's launches, measured
with HIP events on the level's stream during the timed region.  The CPU
baseline times the reference Fortran (built from the reference sources,
OpenMP-style z-slab threads, its configured GAUGE 2, residuals in the
predictor) on the host cores, median of several steps; the GPU/CPU ratio is
taken against the GPU at GAUGE 2.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nr-amssncku_amd"))

METRIC = "BSSN RHS Mpoints/s + % FP64/HBM roofline; batched vs plain; adjoint/forward"


def traffic_from_profiles():
    """HBM bytes per launch of the fused LDS kernel on this workload, from the
    committed rocprof PMC summary (request-size-resolved reads + WRITE_SIZE,
    calibrated against a copy kernel of known size), if present."""
    p = os.path.join(ROOT, "profiles", "rk4_lds_c1_128cubed_traffic.json")
    try:
        with open(p) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(cfg, nsteps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cpu_baseline as CB  # the oracle package's harness (CPU baseline leg only)
    t, kind, cores, times = CB.time_reference_steps(cfg, nsteps=nsteps)
    mpts = cfg.interior_pts * 4 / t / 1e6
    what = ("reference Fortran (amdflang -O3), its configured GAUGE 2" if kind == "reference"
            else "C restatement, GAUGE 0")
    sample = (f"median of {nsteps} RK4 steps of the same 128^3 workload after 1 warm-up ({what}; constraint "
              f"residuals in the predictor, frozen ghost zones), z-slab threads with 3 ghost planes; "
              f"{t:.3f} s per step (min {min(times):.3f}, max {max(times):.3f})")
    return dict(value=round(mpts, 3), unit="Mpoints/s", cores=cores, kind=kind, sample=sample,
                gauge=2 if kind == "reference" else 0)


def copy_probe(mib):
    """A plain device-to-device copy of known size: FETCH_SIZE and WRITE_SIZE
    of its dispatches (mib MiB each) give the counters' scale on this GPU."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)  # an elementwise kernel (copy_ would be a DMA memcpy)
    torch.cuda.synchronize()
    assert torch.equal(dst, src)
    print(json.dumps({"copy_probe_mib": mib, "dispatches": 5}), flush=True)
    return 0


def batching(steps, warmup):
    """The 37-patch uneven level (extents multiples of 8 in [24,96], seed 1,
    packed in [-24,24]^3): the same RK4 steps with one launch per patch and
    kind of work (plain), with the batched tile-list launches, and replayed
    from hipGraphs.  Every step forms the constraint residuals in its
    predictor and their norms, as the headline step does."""
    import amssbssn as A
    cfg = A.config2()
    lv = A.Level.from_config(cfg, gauge=0)
    pts = cfg.interior_pts
    choice, table = lv.autotune()
    out = dict(workload=f"{len(cfg.patches)} uneven patches (multiples of 8 in [24,96], seed 1) in [-24,24]^3",
               interior_points=pts, steps=steps, tune=choice)
    for name, mode in (("plain_direct", A.MODE_PLAIN), ("plain", A.MODE_PLAIN), ("batched", A.MODE_BATCHED),
                       ("graph", A.MODE_GRAPH)):
        lv.set_variant(0 if name == "plain_direct" else 1)
        assert lv.step(cfg.dt, warmup, mode, want_ham=A.RES_EVERY_STEP) == 0
        t0 = time.perf_counter()
        assert lv.step(cfg.dt, steps, mode, want_ham=A.RES_EVERY_STEP) == 0
        dt = time.perf_counter() - t0
        out[name + "_mpoints_s"] = round(pts * 4 * steps / dt / 1e6, 2)
        out[name + "_ms_per_step"] = round(dt * 1e3 / steps, 3)
    out["graph_vs_plain"] = round(out["graph_mpoints_s"] / out["plain_mpoints_s"], 3)
    out["batched_vs_plain"] = round(out["batched_mpoints_s"] / out["plain_mpoints_s"], 3)
    out["graph_vs_plain_direct"] = round(out["graph_mpoints_s"] / out["plain_direct_mpoints_s"], 3)
    out["note"] = ("plain = one launch per patch of the LDS kernel (variant 1); plain_direct = one launch per patch "
                   "of the direct one-thread-per-point kernel (variant 0)")
    out["tune_table_ms"] = [[t[0], t[1], t[2], round(t[3], 4)] for t in table]
    lv.close()
    return out


def transfer_leg(reps=20):
    """Inter-level transfer (prolong3 / restrict3, 6-point 5th order, 2:1
    cell-centred): a 76^3 coarse array (h 0.5 over [-19,19]^3) prolonged
    onto the whole of a 134^3 fine array (h 0.25 over [-16.75,16.75]^3), and
    the fine data restricted back onto the 62^3 coarse cells of
    [-15.5,15.5]^3.  Device arrays,
    kernels timed with events on the launch stream.  Algorithmic bytes: 8 B
    written + 1 B read (an eighth of a coarse value) per fine cell; 8 B
    written + 64 B read per restricted coarse cell."""
    import ctypes as C
    import torch
    import amssbssn as A
    L = A.lib()
    dev = torch.device("cuda:0")
    CD, FD = 0.5, 0.25
    extc, extf = (76, 76, 76), (134, 134, 134)
    llbc = [-CD * 38] * 3
    llbf = [-FD * 67] * 3
    uubc = [a + CD * n for a, n in zip(llbc, extc)]
    uubf = [a + FD * n for a, n in zip(llbf, extf)]
    d3 = lambda v: (C.c_double * 3)(*v)  # noqa: E731
    i3 = lambda v: (C.c_int * 3)(*v)  # noqa: E731
    g = torch.Generator(device="cpu").manual_seed(3)
    func = torch.randn(extc[::-1], generator=g, dtype=torch.float64).to(dev)
    funf = torch.zeros(extf[::-1], dtype=torch.float64, device=dev)
    s = torch.cuda.current_stream(dev)
    llbr, uubr = [-15.5] * 3, [15.5] * 3

    def prolong():
        return L.amss_prolong3_dev(d3(llbc), d3(uubc), i3(extc), C.c_void_p(func.data_ptr()), d3(llbf),
                                   d3(uubf), i3(extf), C.c_void_p(funf.data_ptr()), d3(llbf), d3(uubf),
                                   d3((1.0, 1.0, 1.0)), C.c_void_p(s.cuda_stream))

    def restrict():
        return L.amss_restrict3_dev(d3(llbc), d3(uubc), i3(extc), C.c_void_p(func.data_ptr()), d3(llbf),
                                    d3(uubf), i3(extf), C.c_void_p(funf.data_ptr()), d3(llbr), d3(uubr),
                                    d3((1.0, 1.0, 1.0)), C.c_void_p(s.cuda_stream))

    out = dict(workload="76^3 coarse (h 0.5) -> whole 134^3 fine (h 0.25), then restrict onto 62^3 coarse cells")
    nf = extf[0] * extf[1] * extf[2]
    nr = int(round((uubr[0] - llbr[0]) / CD)) ** 3
    for name, fn, pts, bpp in (("prolong", prolong, nf, 9.0), ("restrict", restrict, nr, 72.0)):
        assert fn() == 0
        s.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name] = dict(points=pts, ms=round(ms, 4), mpoints_s=round(pts / ms / 1e3, 1),
                         achieved_gbs=round(pts * bpp / ms / 1e6, 1),
                         frac_hbm=round(pts * bpp / ms / 1e6 / 8000.0, 4), bytes_per_point=bpp)
    return out


def boundary_leg(steps=5, warmup=2):
    """The 128^3 two-puncture step under each outer boundary: frozen ghost
    zones (the headline workload: no boundary kernel), and the reference's
    one-layer copy, no boundary, and level-0 Sommerfeld (the 3-deep shell of
    308,104 points through the boundary kernels with the reference's edge
    stencils).  Wall ms per RK4 step with the predictor's residuals, one
    launch per kind of work (mode 0)."""
    import amssbssn as A
    cfg = A.config1(n=128)
    out = dict(workload="two-puncture 128^3 + 3 ghosts, GAUGE 0, plain mode", steps=steps)
    for name, bc in (("frozen", A.BC_FROZEN_GHOSTS), ("copy", A.BC_COPY_LAYER), ("none", A.BC_NONE),
                     ("sommerfeld", A.BC_SOMMERFELD)):
        lv = A.Level.from_config(cfg, gauge=0, bc=bc)
        assert lv.step(cfg.dt, warmup, want_ham=A.RES_EVERY_STEP) == 0
        t0 = time.perf_counter()
        assert lv.step(cfg.dt, steps, want_ham=A.RES_EVERY_STEP) == 0
        out[name + "_ms_per_step"] = round((time.perf_counter() - t0) * 1e3 / steps, 3)
        lv.close()
    return out


def adjoint_leg(steps, checkpoint=20):
    """The 64^3 differentiation workload (two-puncture data over [-16,16]^3,
    parameters: masses, x positions, 16^3 spline coefficients on A_xx): J and
    dJ/dtheta through the checkpointed step adjoint, the wall time of the
    backward pass over the forward pass, and a central difference of dJ/dm1."""
    import numpy as np
    import amssbssn.gradient as G
    G.value_and_grad(G.config3(64, steps=2, checkpoint=1))  # load kernels, warm caches
    cfg = G.config3(64, steps=steps, checkpoint=checkpoint)
    J, g, tm = G.value_and_grad(cfg)
    theta = cfg.theta0()
    h = 1e-6
    tp, tmn = theta.copy(), theta.copy()
    tp[0] += h
    tmn[0] -= h
    fd = (G.value(cfg, tp) - G.value(cfg, tmn)) / (2 * h)
    return dict(workload=f"64^3 patch + 3 ghosts over [-16,16]^3, {steps} RK4 steps, checkpoint every {checkpoint}; "
                         "J = shell integral of alpha chi (r = 6, width 1); theta = (m1, m2, x1, x2, 16^3 A_xx spline)",
                J=J, dJ_dtheta_head=[float(x) for x in g[:4]], spline_grad_norm=float(np.linalg.norm(g[4:])),
                forward_s=round(tm["forward_s"], 4), backward_s=round(tm["backward_s"], 4),
                adjoint_over_forward=round(tm["backward_s"] / tm["forward_s"], 3),
                fd_dJ_dm1=fd, fd_rel_err_m1=abs(fd - g[0]) / abs(g[0]))


def batched_adjoint_leg(steps):
    """The 16-patch differentiation workload (uneven patches, multiples of 8 in
    [48,128], seed 2; masses 0.55 / 0.45 at x = +-4): the batched forward
    (hipGraph mode) with the checkpoint spacing planned from the free HBM, the
    level adjoint (exchange reversed) backwards; dJ/d(masses, positions)."""
    import amssbssn.gradient as G
    cfg = G.config4(steps=steps)
    J, g, info = G.value_and_grad(cfg, mode=2)
    return dict(workload=f"16 uneven patches (multiples of 8 in [48,128], seed 2) in [-24,24]^3, {steps} RK4 steps; "
                         "theta = (m1, m2, x1, x2)",
                interior_points=cfg.interior_pts, J=J, dJ_dtheta=[float(x) for x in g[:4]],
                checkpoint=info["checkpoint"], forward_s=round(info["forward_s"], 4),
                backward_s=round(info["backward_s"], 4),
                adjoint_over_forward=round(info["backward_s"] / info["forward_s"], 3),
                peak_device_GB=round(info["peak_device_bytes"] / 1e9, 2) if info["peak_device_bytes"] else None)


def headline_variants(cfg, steps, mode, bc, warmup=2):
    """The headline step at GAUGE 2 (the reference build's gauge, for the
    GPU/CPU ratio) with residuals, and at GAUGE 0 without residuals."""
    import amssbssn as A
    out = {}
    for name, gauge, res in (("gauge2", 2, A.RES_EVERY_STEP), ("no_residuals", 0, A.RES_NONE)):
        lv = A.Level.from_config(cfg, gauge=gauge, bc=bc)
        assert lv.step(cfg.dt, warmup, mode, want_ham=res) == 0
        t0 = time.perf_counter()
        assert lv.step(cfg.dt, steps, mode, want_ham=res) == 0
        wall = time.perf_counter() - t0
        out[name + "_mpoints_s"] = round(cfg.interior_pts * 4 * steps / wall / 1e6, 2)
        out[name + "_ms_per_step"] = round(wall * 1e3 / steps, 3)
        lv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128, help="interior points per axis")
    ap.add_argument("--mode", choices=["plain", "batched", "graph"], default="plain")
    ap.add_argument("--bc", choices=["frozen", "copy", "none", "sommerfeld"], default="frozen",
                    help="outer boundary of the 128^3 step: frozen ghost zones (the benchmark workload), or the "
                         "reference's one-layer copy, none, or level-0 Sommerfeld (boundary kernel on the shell)")
    ap.add_argument("--variant", type=int, choices=[0, 1], default=None,
                    help="interior kernel: 0 direct stencil reads, 1 LDS-staged tiles (default: the level's)")
    ap.add_argument("--batch-steps", type=int, default=50,
                    help="timed steps of each mode on the 37-patch level (0 skips the batching comparison)")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="steps of the 64^3 gradient run (0 skips the adjoint/forward measurement)")
    ap.add_argument("--batched-adjoint-steps", type=int, default=20,
                    help="steps of the 16-patch batched gradient run (0 skips it; the full workload is 400)")
    ap.add_argument("--transfer-reps", type=int, default=20,
                    help="timed repetitions of the inter-level transfer leg (0: skip)")
    ap.add_argument("--boundary-steps", type=int, default=5,
                    help="timed steps of the boundary-mode leg (0: skip)")
    ap.add_argument("--cpu-steps", type=int, default=5, help="timed CPU-baseline steps (the median is reported)")
    ap.add_argument("--no-residuals", action="store_true",
                    help="time steps without the predictor's constraint residuals (not the reference's step)")
    ap.add_argument("--no-extra-legs", dest="extra_legs", action="store_false",
                    help="skip the GAUGE 2 and no-residual variants of the headline step")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device copies of MIB MiB (calibrates rocprof's FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design only)", file=sys.stderr)
        return 2

    import numpy as np  # noqa: F401
    import amssbssn as A
    from amssbssn import perf

    cfg = A.config1(n=args.n)
    mode = {"plain": A.MODE_PLAIN, "batched": A.MODE_BATCHED, "graph": A.MODE_GRAPH}[args.mode]
    bc = {"frozen": A.BC_FROZEN_GHOSTS, "copy": A.BC_COPY_LAYER, "none": A.BC_NONE,
          "sommerfeld": A.BC_SOMMERFELD}[args.bc]
    lv = A.Level.from_config(cfg, gauge=0, bc=bc)
    if args.variant is not None:
        lv.set_variant(args.variant)
    variant = lv.variant()
    pts = cfg.interior_pts

    res_mode = A.RES_NONE if args.no_residuals else A.RES_EVERY_STEP
    if args.warmup:
        rc = lv.step(cfg.dt, args.warmup, mode, want_ham=res_mode)
        assert rc == 0, f"warm-up step returned {rc}"
    lv.set_timing(True)
    t0 = time.perf_counter()
    rc = lv.step(cfg.dt, args.steps, mode, want_ham=res_mode)   # synchronises at the end
    wall = time.perf_counter() - t0
    assert rc == 0, f"timed steps returned {rc}"
    ms, nl = lv.kernel_ms()
    lv.set_timing(False)
    series = lv.constraint_series() if res_mode else None

    value = pts * 4 * args.steps / wall / 1e6
    ktime = sum(ms) / 1e3
    launches = sum(nl)
    timing = "HIP events around each fused-kernel launch on the level's stream"
    if launches == 0:  # graph replays record no per-kernel events: whole-step wall time instead
        ktime, launches = wall, 4 * args.steps
        timing = "wall time of the graph-replayed steps / (4 sub-steps), fills and exchange included"
        nl = [args.steps] * 4
    stage_bytes = perf.BYTES_PER_PT_STAGE_RES if res_mode else perf.BYTES_PER_PT_STAGE
    stage_flops = perf.FLOPS_PER_PT_STAGE_RES if res_mode else (perf.FLOPS_PER_PT_RHS,) * 4
    flops = sum(f * pts * n for f, n in zip(stage_flops, nl))
    nbytes = sum(b * pts * n for b, n in zip(stage_bytes, nl))
    # plain launches on frozen ghost zones run the ghost fill inside the
    # fused kernel's launch in sub-steps 0 and 1
    fill_pts = (args.n + 6) ** 3 - args.n ** 3
    fused_fill = variant == 1 and args.mode == "plain" and args.bc == "frozen"
    fill_bytes = sum(perf.BYTES_PER_PT_GHOST_FILL * fill_pts * f * n
                     for f, n in zip(perf.GHOST_FILL_STAGES, nl)) if fused_fill else 0
    nbytes += fill_bytes
    ach_gbs = nbytes / ktime / 1e9
    ach_tf = flops / ktime / 1e12
    ai = sum(stage_flops) / sum(stage_bytes)
    ridge = perf.PEAK_FP64_TFLOPS * 1e12 / (perf.PEAK_HBM_GBS * 1e9)
    traffic = traffic_from_profiles() if (variant == 1 and args.n == 128 and args.bc == "frozen"
                                          and res_mode and args.mode == "plain") else None
    roof = dict(bound="hbm" if ai < ridge else "fp64", achieved=round(ach_gbs, 1), peak=perf.PEAK_HBM_GBS,
                unit="GB/s", frac=round(ach_gbs / perf.PEAK_HBM_GBS, 4), traffic=traffic,
                kernel=("rk4_lds_kernel" if variant == 1 else "rk4_interior_kernel") +
                " (fused RHS + RK4 sub-step, stages 0-3)",
                avg_launch_ms=round(ktime * 1e3 / launches, 4), launches=launches, timing=timing,
                algorithmic_bytes_per_launch=int(nbytes / launches),
                ghost_fill_bytes_per_launch=int(fill_bytes / launches),
                algorithmic_flops_per_launch=int(flops / launches),
                fp64=dict(achieved=round(ach_tf, 3), peak=perf.PEAK_FP64_TFLOPS, unit="TFLOP/s",
                          frac=round(ach_tf / perf.PEAK_FP64_TFLOPS, 4)),
                intensity_flop_per_byte=round(ai, 2), ridge_flop_per_byte=round(ridge, 2))
    if roof["bound"] == "fp64":
        roof.update(achieved=round(ach_tf, 3), peak=perf.PEAK_FP64_TFLOPS, unit="TFLOP/s",
                    frac=round(ach_tf / perf.PEAK_FP64_TFLOPS, 4))
    out = dict(metric=METRIC, value=round(value, 2), unit="Mpoints/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=round(wall * 1e3 / args.steps, 3), higher_is_better=True,
               scaling="weak", vs_baseline=None, dtype="f64",
               data="synthetic: analytic two-puncture (Brill-Lindquist) initial data, Philox A_ij noise",
               config=dict(workload=f"two-puncture BSSN RK4 step, one {args.n}^3 patch + 3 ghosts over [-16,16]^3",
                           interior_points=pts, rhs_evals_per_step=4, gauge=0, ko_eps=cfg.eps, dt=cfg.dt,
                           mode=args.mode, boundary=args.bc, variant=variant,
                           residuals=("constraint residuals in every predictor (co = 0) and their 7 L2 norms "
                                      "every step" if res_mode else "none (every predictor co = 1)"),
                           stage_kernel_ms=[round(m / max(n, 1), 4) for m, n in zip(ms, nl)]),
               roofline=roof)
    if series is not None and len(series):
        out["constraint_norms"] = dict(names=["Ham", "Px", "Py", "Pz", "Gx", "Gy", "Gz"],
                                       first_timed_step=[float(x) for x in series[0]],
                                       last_timed_step=[float(x) for x in series[-1]])
    lv.close()
    if args.extra_legs:
        out["headline_variants"] = headline_variants(cfg, args.steps, mode, bc)
    if args.batch_steps > 0:
        out["batching"] = batching(args.batch_steps, 2)
    if args.transfer_reps > 0:
        out["transfer"] = transfer_leg(args.transfer_reps)
    if args.boundary_steps > 0:
        out["boundary_modes"] = boundary_leg(args.boundary_steps)
    if args.adjoint_steps > 0:
        out["adjoint"] = adjoint_leg(args.adjoint_steps)
    if args.batched_adjoint_steps > 0:
        out["batched_adjoint"] = batched_adjoint_leg(args.batched_adjoint_steps)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_steps)
        g2 = out.get("headline_variants", {}).get("gauge2_mpoints_s")
        if out["cpu_baseline"]["gauge"] == 2 and g2:
            out["cpu_baseline"]["gpu_speedup"] = round(g2 / out["cpu_baseline"]["value"], 1)
            out["cpu_baseline"]["gpu_speedup_note"] = "GPU at GAUGE 2 with residuals / reference Fortran at GAUGE 2"
        else:
            out["cpu_baseline"]["gpu_speedup"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
