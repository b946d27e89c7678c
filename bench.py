#!/usr/bin/env python3
"""Frequency points per second of the per-point MoM path (spatial sums, spectral
tables, fill, LU, RHS/solve, extraction, cascade) on one MI355X.

Workload (default): PEC annular ring (a = 3.0 mm, b = 3.6 mm) in a 9 mm cell on a
0.787 mm eps_r = 2.2 / tan_delta = 0.0009 slab, 2016 faces / 2856 unknowns,
theta = 30 deg, phi in {0, 45} deg, 51 frequencies 5..25 GHz: 102 points.
One step = the whole sweep (all 102 points) with the sheet resident on the
device; meshing, setup_rwg and mode selection are outside the timed region.

value = points/s of the mixed path (FP32 LU and smooth spatial sums, FP64
refinement, singular terms and tables); the plain FP64 path is timed beside it
and max|dS| between them is reported, as is max|dS| of both against the oracle
on the CPU-baseline sample point.  detail.two_layer_patch holds the
differentiation workload: forward points/s in FP64 and mixed, and adjoint
gradient vectors per second; detail.meander_polarizer a sample of the cascaded
4-sheet meanderline polarizer at full mesh size (about 12 k unknowns per sheet);
detail.cross_batch a sample of the batched Jerusalem / loaded cross problems.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _sub in ("pssfss.jl_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, _sub))

# HBM bytes per launch pair of the mixed spatial-sum kernels on this workload
# (k_jk_mfma<1,2> and the half-width k_jk_mfma<1,1>), from the FETCH_SIZE /
# WRITE_SIZE passes in profiles/ring_on_slab_traffic_pmc.md (counters calibrated
# with the 256 MiB copy of --calibrate copy: FETCH x2.00, WRITE x1.00).
JK_MFMA_TRAFFIC_BYTES = ((2860.8 + 558.3) * 1024 * 2.0
                         + (40685465.2 + 2554272.0) * 1024 * 1.0)
PEAK = {"fp64": 78.6, "mixed": 157.3}    # TFLOP/s, MI355X dense FP64 (vector = matrix) / FP32 matrix


def max_dS(a, b):
    import numpy as np
    m = 0.0
    for x, y in zip(a, b):
        for blk in ("s11", "s12", "s21", "s22"):
            m = max(m, float(np.abs(getattr(x["gsm"], blk) - getattr(y["gsm"], blk)).max()))
    return m


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ring", choices=["ring", "patch"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--precisions", default="fp64,mixed")
    ap.add_argument("--no-two-layer", action="store_true",
                    help="skip the two-layer patch forward / adjoint measurements")
    ap.add_argument("--no-meander", action="store_true",
                    help="skip the 4-sheet meanderline polarizer sample")
    ap.add_argument("--no-cross", action="store_true",
                    help="skip the batched cross-problem sample")
    ap.add_argument("--calibrate", choices=["copy", "fma"],
                    help="run only a counter-calibration kernel of known work (under rocprofv3): "
                         "copy = 256 MiB device copy x 3, fma = FP64 and FP32 FMA loops of "
                         "2 * 4096 * 256 * 4096 flops each")
    args = ap.parse_args()
    if args.calibrate:
        import ctypes
        cl = ctypes.CDLL(os.path.join(ROOT, "pssfss.jl_amd", "libpssfss_calib.so"))
        cl.pssfss_calib_copy.argtypes = [ctypes.c_int64, ctypes.c_int]
        rc = (cl.pssfss_calib_copy(256 << 20, 3) if args.calibrate == "copy"
              else cl.pssfss_calib_fma(4096, 4096))
        print(f"[bench] calibration {args.calibrate}: status {rc}", flush=True)
        sys.exit(0 if rc == 0 else 1)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(frequency points would shard across GPUs independently, one process "
              "per GPU; this build runs on one)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pssfss_amd as pk
    from pssfss_amd import hip

    if args.workload == "ring":
        strata, freqs, steer = pk.workloads.ring_on_slab()
        wname = ("ring_on_slab: PEC annulus a=3.0 b=3.6 mm, 9 mm cell, 0.787 mm eps_r=2.2 "
                 "tan_d=0.0009 slab, theta=30, phi={0,45}, 51 freqs 5-25 GHz")
    else:
        strata, freqs, steer = pk.workloads.patch_free_space()
        wname = "patch_free_space: 5 mm PEC patch, 10 mm cell, 21 freqs 2-30 GHz, normal incidence"
    npts, nface, nbf = pk.workloads.sheet_points(strata, freqs, steer)
    ctx = hip.get_context()
    dev = ctx.device_info()
    print(f"[bench] device {dev}; workload {args.workload}: {npts} points, {nface} faces, "
          f"{nbf} unknowns", flush=True)
    setup = pk.analyze.prepare(strata, freqs, steer)
    precs = [p for p in args.precisions.split(",") if p]
    solvers = {}
    shared = None
    for p in precs:
        s = hip.SheetSolver(ctx, setup, p)
        if shared is None:
            shared = s.dsheets
        s.dsheets = shared
        solvers[p] = s

    timing, results, stages = {}, {}, {}
    for p in precs:
        for w in range(args.warmup):
            t0 = time.perf_counter()
            solvers[p].run()
            print(f"[bench] warmup {p} step {w}: {time.perf_counter() - t0:.3f} s", flush=True)
        ctx.stats_read(reset=True)
        ctx.stats_enable(True)
        times = []
        for k in range(args.steps):
            t0 = time.perf_counter()
            res, info = solvers[p].run()
            times.append(time.perf_counter() - t0)
            print(f"[bench] {p} step {k}: {times[-1]:.3f} s", flush=True)
        ctx.stats_enable(False)
        stages[p] = ctx.stats_read(reset=True)
        results[p] = res
        timing[p] = times
        iters = [i.refine_iters for i in info["infos"]]
        fb = sum(i.fallback_fp64 for i in info["infos"])
        if p == "mixed":
            print(f"[bench] mixed refinement sweeps min/max {min(iters)}/{max(iters)}, "
                  f"FP64 fallbacks {fb}", flush=True)

    head = "mixed" if "mixed" in timing else precs[0]
    step_s = sum(timing[head]) / len(timing[head])
    value = npts / step_s
    extra = {}
    for p in precs:
        extra[f"points_per_s_{p}"] = npts / (sum(timing[p]) / len(timing[p]))
    if "mixed" in results and "fp64" in results:
        extra["max_dS_mixed_vs_fp64"] = max_dS(results["mixed"], results["fp64"])
        extra["speedup_mixed_vs_fp64"] = extra["points_per_s_mixed"] / extra["points_per_s_fp64"]

    # Roofline of the dominant kernel: the spatial sums (the stage that takes the
    # most time), algorithmic flops = 8 per (class, quadrature point, phasing,
    # lattice point) that the reference's ring rule sums, counted on the device.
    st = stages[head]
    jk = st["spatial_sums"]
    jk_kernel = ("k_jk_mfma<1,2> + k_jk_mfma<1,1> (spatial sums, FP32 MFMA; the second "
                 "launch is the last phasing block at half width)" if head == "mixed"
                 else "k_jk_mfma64<2> + k_jk_mfma64<1> (spatial sums, FP64 MFMA)")
    achieved = jk["flops"] / (jk["ms"] * 1e-3) / 1e12 if jk["ms"] > 0 else 0.0
    roof = {"bound": "mfma", "kernel": jk_kernel, "achieved": round(achieved, 3),
            "peak": PEAK[head], "unit": "TFLOP/s", "frac": round(achieved / PEAK[head], 4),
            "avg_launch_ms": round(jk["ms"] / max(1, jk["launches"]), 3),
            "flops_per_launch": jk["flops"] / max(1, jk["launches"]),
            "traffic": (JK_MFMA_TRAFFIC_BYTES if head == "mixed" and args.workload == "ring"
                        else None)}
    g = st["lu_gemm"]
    g_ach = g["flops"] / (g["ms"] * 1e-3) / 1e12 if g["ms"] > 0 else 0.0
    extra["roofline_lu_gemm"] = {"kernel": "k_gemm_mfma (LU trailing update)",
                                 "achieved": round(g_ach, 3), "peak": PEAK[head],
                                 "unit": "TFLOP/s", "frac": round(g_ach / PEAK[head], 4),
                                 "avg_launch_ms": round(g["ms"] / max(1, g["launches"]), 4)}
    tot_ms = sum(v["ms"] for k, v in st.items() if k != "lu_total_flops")
    stage_ms = {k: round(v["ms"] / args.steps, 2) for k, v in st.items() if k != "lu_total_flops"}
    lu_flops = st["lu_total_flops"]["flops"] / args.steps
    extra["stage_ms_per_step"] = stage_ms
    extra["dominant_stage"] = max(stage_ms, key=stage_ms.get)
    extra["lu_tflops_per_s_whole_step"] = round(lu_flops / step_s / 1e12, 3)

    if not args.no_two_layer:
        extra["two_layer_patch"] = two_layer_patch(pk, hip, ctx)
    if not args.no_meander:
        extra["meander_polarizer"] = meander_polarizer(pk, hip, ctx)
    if not args.no_cross:
        extra["cross_batch"] = cross_batch(pk, hip, ctx)

    cpu = None
    if not args.no_cpu:
        cpu, errs = cpu_baseline(pk, strata, freqs, steer, results)
        extra.update(errs)

    line = {
        "metric": "freq points/s (fill+LU+GSM) mixed vs FP64 at |dS|<=1e-6",
        "value": round(value, 3), "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(step_s * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": wname, "points": npts, "faces": nface, "unknowns": nbf,
                   "precision": head},
        "roofline": roof, "cpu_baseline": cpu, "detail": extra,
    }
    print(json.dumps(line), flush=True)


def two_layer_patch(pk, hip, ctx):
    """Corner-parameterised 8 mm patch (39 x 39 bilinear grid: 3042 faces, 4485
    unknowns, every face pair its own class) in a 12 mm cell on eps_r 3.0 (1.0 mm)
    + eps_r 1.1 foam (3.0 mm); 12 frequencies 8..12 GHz at normal incidence.
    Forward points/s of the FP64 and mixed paths, and gradient vectors per second of
    |S21|^2 co-polar TE and TM with respect to the 8 corner coordinates (one
    gradient vector = 8 derivatives of one objective at one point)."""
    import numpy as np
    from pssfss_amd import adjoint as A
    corners = [[2.0, 2.0], [10.0, 2.0], [10.0, 10.0], [2.0, 10.0]]
    sh, vt = A.bilinear_patch(corners, 39, 39, 12.0, 12.0)
    L = pk.Layer.make
    strata = [L(), sh, L(width=1.0, units="mm", epsr=3.0), L(width=3.0, units="mm", epsr=1.1),
              L()]
    freqs = np.linspace(8.0, 12.0, 12)
    steer = {"theta": [0.0], "phi": [0.0]}
    setup = pk.analyze.prepare(strata, freqs, steer)
    out = {"faces": sh.nface, "unknowns": int(setup.rwgdat[0].nbf), "points": len(freqs)}
    res = {}
    for p in ("fp64", "mixed"):
        s = hip.SheetSolver(ctx, setup, p)
        s.run()                                   # warm-up (J/K integrals, allocations)
        t0 = time.perf_counter()
        res[p], _ = s.run()
        dt = time.perf_counter() - t0
        out[f"points_per_s_{p}"] = round(len(freqs) / dt, 3)
        s.close()
        print(f"[bench] two-layer patch {p}: {dt:.3f} s for {len(freqs)} points", flush=True)
    out["speedup_mixed_vs_fp64"] = round(out["points_per_s_mixed"] / out["points_per_s_fp64"], 3)
    out["max_dS_mixed_vs_fp64"] = max_dS(res["mixed"], res["fp64"])
    obj = [(2, 1, 1, 1), (2, 1, 2, 2)]
    grads = {}
    for p in ("fp64", "mixed"):
        t0 = time.perf_counter()
        grads[p], _ = A.gradients(strata, freqs, steer, vt, obj, ctx=ctx, precision=p)
        dt = time.perf_counter() - t0
        g = grads[p]
        out[f"adjoint_grads_per_s_{p}"] = round(g.shape[0] * g.shape[1] / dt, 3)
        out[f"adjoint_s_{p}"] = round(dt, 3)
        print(f"[bench] two-layer patch adjoint {p}: {dt:.3f} s for {g.shape[0]} points x "
              f"{g.shape[1]} objectives x {g.shape[2]} parameters", flush=True)
    d = np.abs(grads["mixed"] - grads["fp64"]) / np.abs(grads["fp64"]).max(axis=-1, keepdims=True)
    out["adjoint_max_rel_mixed_vs_fp64"] = float(d.max())
    return out


def _sweep(hip, ctx, setup, label):
    """FP64 and mixed runs of a prepared sweep (one untimed warm-up each); points
    are those analysed (the sweep skips cut-off points as the reference does)."""
    out, res = {}, {}
    for p in ("fp64", "mixed"):
        s = hip.SheetSolver(ctx, setup, p)
        s.run()
        t0 = time.perf_counter()
        res[p], info = s.run()
        dt = time.perf_counter() - t0
        s.close()
        npts = len(res[p])
        out["points"] = npts
        out[f"points_per_s_{p}"] = round(npts / dt, 4)
        if p == "mixed":
            it = [i.refine_iters for i in info["infos"]]
            out["mixed_refinement_sweeps_min_max"] = [min(it), max(it)]
            out["mixed_fp64_fallbacks"] = int(sum(i.fallback_fp64 for i in info["infos"]))
        print(f"[bench] {label} {p}: {dt:.3f} s for {npts} points", flush=True)
    out["speedup_mixed_vs_fp64"] = round(out["points_per_s_mixed"] / out["points_per_s_fp64"], 3)
    out["max_dS_mixed_vs_fp64"] = max_dS(res["mixed"], res["fp64"])
    return out


def meander_polarizer(pk, hip, ctx):
    """Four seeded meanderline sheets of about 8000 faces (11.3-12.2 k unknowns
    each) on eps_r 1.07 foam with eps_r 2.8 skins, 30 modes per foam layer
    (dbmin 60), 2 frequencies (8, 14 GHz) at theta 0 and 40 deg: 4 points, 16
    sheet systems, 32-60 right-hand sides per system.  Face-pair classes come from
    the device classifier."""
    strata, freqs, steer = pk.workloads.meander_polarizer(ntri=8000, nfreq=2)
    t0 = time.perf_counter()
    setup = pk.analyze.prepare(strata, freqs, steer, dbmin=60.0,
                               classify=lambda sh: hip.face_pair_classes(ctx, sh))
    prep = time.perf_counter() - t0
    out = {"unknowns": [int(r.nbf) for r in setup.rwgdat],
           "face_pair_classes": [int(r.nufp) for r in setup.rwgdat],
           "modes_per_layer": [len(l.P) for l in setup.layers],
           "prepare_s": round(prep, 3)}
    out.update(_sweep(hip, ctx, setup, "meander polarizer"))
    return out


def cross_batch(pk, hip, ctx):
    """The first Jerusalem cross and the first loaded cross of the seeded 16-problem
    batch (full meshes), each at 8 of its 64 frequencies (every 8th from 1 GHz) and
    the 4 scan angles: 64 points of two uneven problem sizes, run one problem after
    the other."""
    probs = pk.workloads.cross_batch()[:2]
    out = {"problems": [], "points": 0}
    tot = {"fp64": 0.0, "mixed": 0.0}
    for strata, freqs, steer, par in probs:
        f = list(freqs[::8])
        setup = pk.analyze.prepare(strata, f, steer)
        r = _sweep(hip, ctx, setup, par["kind"])
        npts = r["points"]
        r.update({"kind": par["kind"], "period_mm": round(par["P"], 3),
                  "faces": int(strata[1].nface), "unknowns": int(setup.rwgdat[0].nbf)})
        out["problems"].append(r)
        out["points"] += npts
        for p in tot:
            tot[p] += npts / r[f"points_per_s_{p}"]
    for p in tot:
        out[f"points_per_s_{p}"] = round(out["points"] / tot[p], 4)
    out["speedup_mixed_vs_fp64"] = round(tot["fp64"] / tot["mixed"], 3)
    return out


def cpu_baseline(pk, strata, freqs, steer, gpu_results):
    """Oracle (FP64 CPU restatement) on the first point of the sweep: points/s on
    the host's cores, and max|dS| of each GPU path against it at that point."""
    import oracle
    f0 = [float(freqs[0])]
    st0 = {k: [list(v)[0]] for k, v in steer.items()}
    t0 = time.perf_counter()
    ref = oracle.analyze(pk, strata, f0, st0)
    dt = time.perf_counter() - t0
    cores = oracle.lib().orc_num_threads()
    errs = {}
    for p, res in gpu_results.items():
        errs[f"max_dS_{p}_vs_oracle_point0"] = max_dS(res[:1], ref)
    keys = list(steer.keys())
    sample = (f"1 point ({f0[0]} GHz, {keys[0]}={st0[keys[0]][0]}, {keys[1]}={st0[keys[1]][0]}) "
              f"of the same workload, full mesh, {dt:.1f} s")
    return ({"value": round(1.0 / dt, 5), "unit": "points/s", "cores": cores, "kind": "port",
             "sample": sample}, errs)


if __name__ == "__main__":
    main()
