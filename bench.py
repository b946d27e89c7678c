#!/usr/bin/env python3
"""Frequency points per second of the per-point MoM path (spatial sums, spectral
tables, fill, LU, RHS/solve, extraction, cascade) on one MI355X.

Headline workload (default, --workload two_layer): a corner-parameterised 8 mm PEC
patch (39 x 39 bilinear grid: 3042 faces, 4485 unknowns, every face pair its own
class) in a 12 mm cell on eps_r 3.0 (1.0 mm) + eps_r 1.1 foam (3.0 mm), 12
frequencies 8..12 GHz at normal incidence: 12 points.  One step = the whole sweep
from a freshly uploaded sheet, so the frequency-independent spatial sums (one
phasing key at normal incidence) run once per step, as once per analyze call in
the reference; the sheet upload (mesh and maps) is outside the timed region.

value = points/s of the mixed path (FP32 LU and smooth spatial sums, FP64
refinement, singular terms and tables), all 12 points batched.  Beside it: the
plain FP64 path batched the same way, the plain FP64 path one point per library
call (the reference's calling pattern), max|dS| between the paths, per-stage
times, the LU's whole-step flop rate F = (8/3) N^3 + 8 N^2 nrhs per point, and
the adjoint gradient rate of |S21|^2 co-polar TE and TM with respect to the 8
corner coordinates.  detail.ring_on_slab times the oblique ring sweep (spatial
sums on every point); detail.meander_polarizer and detail.cross_batch sample the
cascaded and batched workloads.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _sub in ("pssfss.jl_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, _sub))

# HBM bytes per launch pair of the mixed spatial-sum kernels on the ring workload
# (k_jk_mfma<1,2> and the half-width k_jk_mfma<1,1>), from the FETCH_SIZE /
# WRITE_SIZE passes in profiles/ring_on_slab_traffic_pmc.md (counters calibrated
# with the 256 MiB copy of --calibrate copy: FETCH x2.00, WRITE x1.00).
JK_MFMA_TRAFFIC_BYTES = ((2860.8 + 558.3) * 1024 * 2.0
                         + (40685465.2 + 2554272.0) * 1024 * 1.0)
# HBM bytes per launch of the mixed LU trailing update k_gemm_mfma<cf,2,false,1,4> on the
# two-layer patch (mean over its 34 launches per sweep), from the FETCH_SIZE / WRITE_SIZE
# passes in profiles/two_layer_patch_traffic_pmc.md, calibrated the same way.
TRAIL_GEMM_TRAFFIC_BYTES = 1.024e6 * 1024 * 2.0 + 6.930e5 * 1024 * 1.0
PEAK = {"fp64": 78.6, "mixed": 157.3}    # TFLOP/s, MI355X dense FP64 (vector = matrix) / FP32 matrix


def max_dS(a, b):
    import numpy as np
    m = 0.0
    for x, y in zip(a, b):
        for blk in ("s11", "s12", "s21", "s22"):
            m = max(m, float(np.abs(getattr(x["gsm"], blk) - getattr(y["gsm"], blk)).max()))
    return m


def two_layer_strata(pk):
    from pssfss_amd import adjoint as A
    import numpy as np
    corners = [[2.0, 2.0], [10.0, 2.0], [10.0, 10.0], [2.0, 10.0]]
    sh, vt = A.bilinear_patch(corners, 39, 39, 12.0, 12.0)
    L = pk.Layer.make
    strata = [L(), sh, L(width=1.0, units="mm", epsr=3.0), L(width=3.0, units="mm", epsr=1.1),
              L()]
    return strata, np.linspace(8.0, 12.0, 12), {"theta": [0.0], "phi": [0.0]}, vt


def timed_sweeps(hip, ctx, setup, prec, steps, warmup, label, max_batch=0):
    """`warmup` untimed then `steps` timed sweeps, each from freshly uploaded sheets
    (uploaded before the clock starts).  Returns (per-step seconds, last results,
    last infos, stage stats summed over the timed steps)."""
    for w in range(warmup):
        with hip.SheetSolver(ctx, setup, prec, max_batch) as s:
            s.upload()
            t0 = time.perf_counter()
            s.run()
            print(f"[bench] {label} {prec} warmup {w}: {time.perf_counter() - t0:.3f} s",
                  flush=True)
    ctx.stats_read(reset=True)
    times, res, info = [], None, None
    for k in range(steps):
        with hip.SheetSolver(ctx, setup, prec, max_batch) as s:
            s.upload()
            ctx.stats_enable(True)
            t0 = time.perf_counter()
            res, info = s.run()
            times.append(time.perf_counter() - t0)
            ctx.stats_enable(False)
        print(f"[bench] {label} {prec} step {k}: {times[-1]:.3f} s", flush=True)
    return times, res, info, ctx.stats_read(reset=True)


def stage_summary(st, steps):
    return {k: round(v["ms"] / steps, 3) for k, v in st.items() if k != "lu_total_flops"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="two_layer", choices=["two_layer", "ring", "patch"])
    ap.add_argument("--only-detail", choices=["cross_batch", "meander_polarizer"],
                    help="run one detail sample alone and print its JSON")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-detail", action="store_true",
                    help="headline workload only (no per-point FP64 sweep, adjoint, ring, "
                         "meander or cross samples)")
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
    if args.only_detail:
        ctx = hip.get_context()
        fn = {"cross_batch": cross_batch, "meander_polarizer": meander_polarizer}[args.only_detail]
        print(json.dumps({args.only_detail: fn(pk, hip, ctx)}), flush=True)
        return

    vt = None
    if args.workload == "two_layer":
        strata, freqs, steer, vt = two_layer_strata(pk)
        wname = ("two_layer_patch: 8 mm PEC patch (39x39 bilinear grid), 12 mm cell, eps_r 3.0 "
                 "1.0 mm + eps_r 1.1 3.0 mm, 12 freqs 8-12 GHz, normal incidence")
    elif args.workload == "ring":
        strata, freqs, steer = pk.workloads.ring_on_slab()
        wname = ("ring_on_slab: PEC annulus a=3.0 b=3.6 mm, 9 mm cell, 0.787 mm eps_r=2.2 "
                 "tan_d=0.0009 slab, theta=30, phi={0,45}, 51 freqs 5-25 GHz")
    else:
        strata, freqs, steer = pk.workloads.patch_free_space()
        wname = "patch_free_space: 5 mm PEC patch, 10 mm cell, 21 freqs 2-30 GHz, normal incidence"
    ctx = hip.get_context()
    dev = ctx.device_info()
    setup = pk.analyze.prepare(strata, freqs, steer)
    npts = len(list(setup.points()))
    nface, nbf = setup.sheets[0].nface, int(setup.rwgdat[0].nbf)
    nrhs = len(setup.layers[0].P) + len(setup.layers[-1].P)
    print(f"[bench] device {dev}; workload {args.workload}: {npts} points, {nface} faces, "
          f"{nbf} unknowns, {nrhs} right-hand sides", flush=True)

    timing, results, stages, extra = {}, {}, {}, {}
    for p in ("fp64", "mixed"):
        times, res, info, st = timed_sweeps(hip, ctx, setup, p, args.steps, args.warmup,
                                            args.workload)
        timing[p], results[p], stages[p] = times, res, st
        extra[f"points_per_s_{p}"] = round(npts / (sum(times) / len(times)), 3)
        extra[f"stage_ms_per_step_{p}"] = stage_summary(st, args.steps)
        if p == "mixed":
            it = [i.refine_iters for i in info["infos"]]
            extra["mixed_refinement_sweeps_min_max"] = [min(it), max(it)]
            extra["mixed_fp64_fallbacks"] = int(sum(i.fallback_fp64 for i in info["infos"]))
    step_s = sum(timing["mixed"]) / len(timing["mixed"])
    value = npts / step_s
    extra["max_dS_mixed_vs_fp64"] = max_dS(results["mixed"], results["fp64"])
    extra["speedup_mixed_vs_fp64_batched"] = round(value / extra["points_per_s_fp64"], 3)
    if not args.no_detail:
        # the reference's calling pattern: one point per library call, FP64
        t, res1, _, st1 = timed_sweeps(hip, ctx, setup, "fp64", 1, 1, args.workload + " per-point",
                                       max_batch=1)
        extra["points_per_s_fp64_per_point"] = round(npts / t[0], 3)
        extra["stage_ms_per_step_fp64_per_point"] = stage_summary(st1, 1)
        extra["speedup_mixed_vs_fp64_per_point"] = round(value / extra["points_per_s_fp64_per_point"], 3)
        extra["max_dS_fp64_per_point_vs_batched"] = max_dS(res1, results["fp64"])

    # Roofline of the dominant kernel of the mixed step: the LU's trailing-update
    # GEMM (k_gemm_mfma, FP32 matrix cores), algorithmic flops 8 m n k per complex
    # update, time from HIP events on the library's stream around every launch.
    st = stages["mixed"]
    g = st["lu_gemm"]
    g_ach = g["flops"] / (g["ms"] * 1e-3) / 1e12 if g["ms"] > 0 else 0.0
    roof = {"bound": "mfma", "kernel": "k_gemm_mfma<cf,2,false,1,4> (LU trailing updates, FP32 "
            "matrix cores, three real products per complex k)",
            "achieved": round(g_ach, 3), "peak": PEAK["mixed"], "unit": "TFLOP/s",
            "frac": round(g_ach / PEAK["mixed"], 4),
            "avg_launch_ms": round(g["ms"] / max(1, g["launches"]), 4),
            "flops_per_launch": g["flops"] / max(1, g["launches"]),
            "launches_per_step": g["launches"] / args.steps,
            "traffic": TRAIL_GEMM_TRAFFIC_BYTES if args.workload == "two_layer" else None}
    # the LU's whole-step rate of the metric's definition: F per point x points/s
    F = 8.0 / 3.0 * nbf ** 3 + 8.0 * nbf ** 2 * nrhs
    extra["F_flops_per_point"] = F
    extra["roofline_step_F"] = {
        p: {"achieved": round(F * extra[f"points_per_s_{p}"] / 1e12, 3), "peak": PEAK[p],
            "unit": "TFLOP/s",
            "frac": round(F * extra[f"points_per_s_{p}"] / 1e12 / PEAK[p], 4)}
        for p in ("fp64", "mixed")}
    stm = extra["stage_ms_per_step_mixed"]
    extra["dominant_stage_mixed"] = max(stm, key=stm.get)

    if not args.no_detail:
        if vt is not None:
            extra["adjoint"] = adjoint_rates(pk, ctx, strata, freqs, steer, vt)
        if args.workload != "ring":
            extra["ring_on_slab"] = ring_detail(pk, hip, ctx)
        extra["meander_polarizer"] = meander_polarizer(pk, hip, ctx)
        extra["meander_polarizer_gradient"] = meander_gradient(pk, ctx)
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
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": wname, "points": npts, "faces": nface, "unknowns": nbf,
                   "rhs": nrhs, "precision": "mixed"},
        "roofline": roof, "cpu_baseline": cpu, "detail": extra,
    }
    print(json.dumps(line), flush=True)


def adjoint_rates(pk, ctx, strata, freqs, steer, vt):
    """Gradient vectors per second of |S21|^2 co-polar TE and TM with respect to the 8
    corner coordinates (one gradient vector = 8 derivatives of one objective at one
    point), FP64 and mixed forward/adjoint solves."""
    import numpy as np
    from pssfss_amd import adjoint as A
    obj = [(2, 1, 1, 1), (2, 1, 2, 2)]
    out, grads = {}, {}
    for p in ("fp64", "mixed"):
        A.gradients(strata, freqs[:1], steer, vt, obj, ctx=ctx, precision=p)   # warm-up
        t0 = time.perf_counter()
        grads[p], _ = A.gradients(strata, freqs, steer, vt, obj, ctx=ctx, precision=p)
        dt = time.perf_counter() - t0
        g = grads[p]
        out[f"grads_per_s_{p}"] = round(g.shape[0] * g.shape[1] / dt, 3)
        print(f"[bench] adjoint {p}: {dt:.3f} s for {g.shape[0]} points x {g.shape[1]} "
              f"objectives x {g.shape[2]} parameters", flush=True)
    d = np.abs(grads["mixed"] - grads["fp64"]) / np.abs(grads["fp64"]).max(axis=-1, keepdims=True)
    out["max_rel_mixed_vs_fp64"] = float(d.max())
    return out


def ring_detail(pk, hip, ctx, steps=2):
    """The oblique ring sweep (2016 faces, 2856 unknowns, 102 points, a new phasing
    key on every point): points/s and stages of both precisions, and the spatial
    sums' rate (8 flops per lattice point of the reference's ring rule)."""
    strata, freqs, steer = pk.workloads.ring_on_slab()
    setup = pk.analyze.prepare(strata, freqs, steer)
    npts = len(list(setup.points()))
    out, res = {"points": npts, "unknowns": int(setup.rwgdat[0].nbf)}, {}
    for p in ("fp64", "mixed"):
        t, res[p], _, st = timed_sweeps(hip, ctx, setup, p, steps, 1, "ring")
        out[f"points_per_s_{p}"] = round(npts / (sum(t) / len(t)), 3)
        out[f"stage_ms_per_step_{p}"] = stage_summary(st, steps)
        jk = st["spatial_sums"]
        ach = jk["flops"] / (jk["ms"] * 1e-3) / 1e12 if jk["ms"] > 0 else 0.0
        out[f"spatial_sums_{p}"] = {"achieved": round(ach, 3), "peak": PEAK[p],
                                    "unit": "TFLOP/s", "frac": round(ach / PEAK[p], 4)}
    out["spatial_sums_mixed"]["traffic"] = JK_MFMA_TRAFFIC_BYTES
    out["speedup_mixed_vs_fp64"] = round(out["points_per_s_mixed"] / out["points_per_s_fp64"], 3)
    out["max_dS_mixed_vs_fp64"] = max_dS(res["mixed"], res["fp64"])
    return out


def _sweep(hip, ctx, setup, label):
    """FP64 and mixed runs of a prepared sweep (one untimed warm-up each); points
    are those analysed (the sweep skips cut-off points as the reference does)."""
    out, res = {}, {}
    for p in ("fp64", "mixed"):
        t, res[p], info, _ = timed_sweeps(hip, ctx, setup, p, 1, 1, label)
        npts = len(res[p])
        out["points"] = npts
        out[f"points_per_s_{p}"] = round(npts / t[0], 4)
        out[f"cascade_ms_{p}"] = round(1e3 * info["cascade_s"], 3)
        if p == "mixed":
            it = [i.refine_iters for i in info["infos"]]
            out["mixed_refinement_sweeps_min_max"] = [min(it), max(it)]
            out["mixed_fp64_fallbacks"] = int(sum(i.fallback_fp64 for i in info["infos"]))
    out["speedup_mixed_vs_fp64"] = round(out["points_per_s_mixed"] / out["points_per_s_fp64"], 3)
    out["max_dS_mixed_vs_fp64"] = max_dS(res["mixed"], res["fp64"])
    return out


def meander_polarizer(pk, hip, ctx):
    """Four seeded meanderline sheets of about 8000 faces (11.3-12.2 k unknowns
    each) on eps_r 1.07 foam with eps_r 2.8 skins, 34 modes per foam layer
    (dbmin 68: the reference's selection keeps symmetric groups, 30 or 34, and 34 is
    the first count at or above 32), 2 frequencies (8, 14 GHz) at theta 0 and 40
    deg: 4 points, 16 sheet systems, 36-68 right-hand sides per system, the
    9-Gblock cascade batched on the device.  Face-pair classes come from the device
    classifier."""
    strata, freqs, steer = pk.workloads.meander_polarizer(ntri=8000, nfreq=2)
    t0 = time.perf_counter()
    setup = pk.analyze.prepare(strata, freqs, steer, dbmin=68.0,
                               classify=lambda sh: hip.face_pair_classes(ctx, sh))
    prep = time.perf_counter() - t0
    out = {"unknowns": [int(r.nbf) for r in setup.rwgdat],
           "face_pair_classes": [int(r.nufp) for r in setup.rwgdat],
           "modes_per_layer": [len(l.P) for l in setup.layers],
           "prepare_s": round(prep, 3)}
    out.update(_sweep(hip, ctx, setup, "meander polarizer"))
    return out


def meander_gradient(pk, ctx):
    """Gradient of the mean axial ratio AR21(H) over two points (11 GHz, theta 0
    and 40 deg) of the full-size 4-sheet meander polarizer (about 8000 faces per
    sheet, every face pair its own class) with respect to the 4 meander amplitudes:
    device forward solves of all sheets, reverse pass through the cascade on the
    host, one adjoint call per sheet (FP64)."""
    from pssfss_amd import adjoint as A
    from pssfss_amd import outputs as O
    hs = pk.workloads.meander_amplitudes()
    vts = []
    for h in hs:
        _, vt = A.meander_amplitude_tangent(dict(a=6.0, b=10.0, h=float(h), w1=0.3, w2=0.3,
                                                 ntri=8000, units="mm", fufp=False))
        vts.append(vt)
    strata, _, _ = pk.workloads.meander_polarizer(ntri=8000, nfreq=2, fufp=False, hs=hs)
    freqs, steer = [11.0], {"phi": [0.0], "theta": [0.0, 40.0]}
    t0 = time.perf_counter()
    grad, val = A.gradients_stack(strata, freqs, steer, [(i, vts[i]) for i in range(4)],
                                  O.ardb(2, 1, O.H), ctx=ctx, dbmin=30.0)
    dt = time.perf_counter() - t0
    print(f"[bench] meander gradient: {dt:.3f} s, mean AR {val:.3f} dB, grad {grad}", flush=True)
    return {"points": 2, "parameters": 4, "seconds": round(dt, 3),
            "grads_per_s": round(2 / dt, 4), "mean_ar21_h_db": val,
            "d_mean_ar_d_h_db_per_mm": [float(g) for g in grad]}


def cross_batch(pk, hip, ctx):
    """The 16 seeded Jerusalem / loaded cross problems (full meshes, 800-4000
    requested triangles, uneven unknown counts), each at 16 of its 64 frequencies
    (every 4th from 1 GHz) and the 4 scan angles: 1024 points.  Timed four ways per
    precision after one untimed pass over two problems (kernels loaded): one problem
    after another with the library's default chunking; one after another with the
    tuner's points per chunk for each size class (tuned before the timed region);
    four problems at a time (four contexts on one GPU, tuned chunks); problems of
    similar size merged into shared batches (size classes within 1.1x, systems
    padded to each batch's largest unknown count)."""
    from pssfss_amd.batch import analyze_batch, ChunkTuner
    probs = pk.workloads.cross_batch()
    problems = [(s, list(f[::4]), st) for s, f, st, _ in probs]
    t0 = time.perf_counter()
    setups = [pk.analyze.prepare(s, f, st, classify=lambda sh: hip.face_pair_classes(ctx, sh))
              for s, f, st in problems]
    out = {"problems": len(problems), "unknowns": [int(su.rwgdat[0].nbf) for su in setups],
           "prepare_s": round(time.perf_counter() - t0, 3)}
    tuner = ChunkTuner()
    analyze_batch(problems[:2], "fp64", workers=1, setups=setups[:2])
    one = {}
    for prec in ("fp64", "mixed"):
        r1, s1 = analyze_batch(problems, prec, workers=1, setups=setups)
        one[prec] = r1
        rt, st = analyze_batch(problems, prec, workers=1, tuner=tuner, setups=setups)
        r4, s4 = analyze_batch(problems, prec, workers=4, tuner=tuner, setups=setups)
        rm, sm = analyze_batch(problems, prec, setups=setups, merge=True, class_ratio=1.1)
        out["points"] = s1["points"]
        out[f"points_per_s_{prec}_one_at_a_time"] = round(s1["points_per_s"], 3)
        out[f"points_per_s_{prec}_one_at_a_time_tuned"] = round(st["points_per_s"], 3)
        out[f"points_per_s_{prec}_four_at_a_time"] = round(s4["points_per_s"], 3)
        out[f"max_dS_{prec}_tuned_vs_default"] = max(max_dS(a, b) for a, b in zip(r1, rt))
        out[f"max_dS_{prec}_four_vs_one"] = max(max_dS(a, b) for a, b in zip(r1, r4))
        out[f"points_per_s_{prec}_merged"] = round(sm["points_per_s"], 3)
        out[f"max_dS_{prec}_merged_vs_one"] = max(max_dS(a, b) for a, b in zip(r1, rm))
        out["merged_size_classes"] = sm["classes"]
        print(f"[bench] cross batch {prec}: one at a time {s1['points_per_s']:.1f} (tuned "
              f"{st['points_per_s']:.1f}), four at a time {s4['points_per_s']:.1f}, merged "
              f"{sm['points_per_s']:.1f} points/s", flush=True)
    out["max_dS_mixed_vs_fp64"] = max(max_dS(a, b) for a, b in zip(one["mixed"], one["fp64"]))
    out["tuned_chunks"] = tuner.table
    # few points per problem (one frequency, 10 GHz, the 4 angles): one problem's
    # batch is too small to fill the GPU, the merged batches are not
    few = [(s, [10.0], st) for s, _, st in problems]
    fsets = [pk.analyze.prepare(s, f, st, classify=lambda sh: hip.face_pair_classes(ctx, sh))
             for s, f, st in few]
    for prec in ("fp64", "mixed"):
        analyze_batch(few[:2], prec, workers=1, setups=fsets[:2])
        ra, sa = analyze_batch(few, prec, workers=1, setups=fsets)
        rb, sb = analyze_batch(few, prec, setups=fsets, merge=True, class_ratio=1.25)
        out[f"few_points_{prec}"] = {
            "points": sa["points"], "one_at_a_time_points_per_s": round(sa["points_per_s"], 3),
            "merged_points_per_s": round(sb["points_per_s"], 3),
            "merged_size_classes": sb["classes"],
            "max_dS_merged_vs_one": max(max_dS(a, b) for a, b in zip(ra, rb))}
        print(f"[bench] cross batch, 4 points per problem, {prec}: one at a time "
              f"{sa['points_per_s']:.1f}, merged {sb['points_per_s']:.1f} points/s", flush=True)
    return out


def cpu_baseline(pk, strata, freqs, steer, gpu_results):
    """Oracle (FP64 CPU restatement) on the first point of the sweep, which carries
    the sweep's spatial sums as the reference's first point does: points/s on the
    host's cores, and max|dS| of each GPU path against it at that point."""
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
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    sample = (f"1 point ({f0[0]} GHz, {keys[0]}={st0[keys[0]][0]}, {keys[1]}={st0[keys[1]][0]}) "
              f"of the same workload, full mesh, spatial sums included, {dt:.1f} s")
    return ({"value": round(1.0 / dt, 5), "unit": "points/s", "cores": cores, "kind": "port",
             "cpu": cpu_model, "sample": sample}, errs)


if __name__ == "__main__":
    main()
