"""Benchmark: the 1.0M-element hex8 Taylor bar, FP64, on one MI355X.

One step = one increment of ExplicitSolverCH (time-step control, element
stress/force pass over every element, nodal predictor + Fint gather +
corrector + grid update) on the O-grid cylinder of 2,000 hexahedra per
section x 500 layers (1,000,000 elements, 1,022,541 nodes), copper, linear
hardening A = 400 MPa, B = 100 MPa, 227 m/s impact, central difference
(rho_b = 1), Courant time step every 10 increments.  Inputs live in HBM
before the timed region; the timed region is K steps issued back to back on
the library's stream, bracketed by device synchronisation.

Prints one JSON line: value = element-steps/s,
roofline of the element kernel from HIP events recorded around every element
launch of the timed region, and a CPU baseline: the oracle (C restatement of
the reference, OpenMP element loop) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic HBM bytes per element-step (FP64 hex8), the minimum the algorithm moves:
#   integration-point history S(6)+p+evp = 8 words x 8 B x 2 (read+write) x 8 points = 1024 B
#   connectivity 8 x 4 B = 32 B
#   nodal traffic 276 B per node x (nodes / elements)
HIST_B, CONN_B, NODE_B, ELEM_NODE_READ_B = 1024.0, 32.0, 276.0, 48.0
# the mixed form's element pass: FP32 point history, connectivity, FP64 node
# reads and FP32 force slots, 798 B per element
MIXED_ELEM_B = 798.0
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4  # wave64 VALU instructions per ns: 256 CUs x 4 SIMDs, 4 cycles each at 2.4 GHz


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--layers", type=int, default=500)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-only", action="store_true", help="only the batched-bars leg (FP64 and mixed)")
    ap.add_argument("--batch-engine", default="auto", choices=["auto", "persistent", "stepped", "stepped-wide", "stepped-plain"])
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-shard", action="store_true")
    ap.add_argument("--no-pulse", action="store_true")
    ap.add_argument("--no-copy-peak", action="store_true")
    ap.add_argument("--no-full-state", action="store_true")
    ap.add_argument("--precision-report", action="store_true",
                    help="only the O-grid Johnson-Cook bar (4,000 elements per section) in both forms")
    ap.add_argument("--precision-layers", type=int, default=1000)
    ap.add_argument("--precision-tend", type=float, default=60e-6)
    ap.add_argument("--precision-forms", default="fp64,mixed", help="forms of the precision report, comma-separated")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.exit("bench.py: this build runs one domain on one GPU; --gpus %d is not supported "
                 "(a multi-GPU slab decomposition is not built)" % args.gpus)

    import numpy as np
    import redynela_amd as R
    from redynela_amd import _lib as L

    if args.precision_report:
        print(json.dumps({"precision_report": precision_report(args.precision_layers, args.precision_tend,
                                                               args.precision_forms.split(","))}), flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batched": {form: batched_bars(form, engine=args.batch_engine)
                                      for form in ("fp64", "mixed")}}), flush=True)
        return
    mesh = R.mesh.ogrid_taylor_bar(args.layers)
    ne, nn = mesh["conn"].shape[0], mesh["x"].shape[0]
    mat = R.Material(115e9, 0.3, 8930.0, R.IsoHardElastoplastic(400e6, 100e6, 1.0))
    sol = R.ExplicitSolverCH()
    sol.setTimeStepMethod("Courant")
    sol.setDissipation(1.0)
    dom = R.Domain(mesh, mat, sol)
    dom.set_option(L.OPT_STORE_FINT, 0)
    # dnlg_synchronize waits for the library's stream, the only one the steps run on
    assert dom.begin(1.0)
    dom.step_async(args.warmup)
    dom.synchronize()
    dom.enable_kernel_timing(True)
    t0 = time.perf_counter()
    dom.step_async(args.steps)
    dom.synchronize()
    t1 = time.perf_counter()
    n_el, ms_el = dom.kernel_stats(0)
    n_nd, ms_nd = dom.kernel_stats(1)
    dom.enable_kernel_timing(False)
    elapsed = t1 - t0
    tinfo = dom.time()
    fp64_state = (dom.coords, dom.evp) if not args.no_fp32 else None
    value = ne * args.steps / elapsed
    avg_el = ms_el / max(n_el, 1) * 1e-3
    avg_nd = ms_nd / max(n_nd, 1) * 1e-3
    ratio = nn / ne
    elem_bytes = (HIST_B + CONN_B + ELEM_NODE_READ_B * ratio) * ne
    step_bytes = (HIST_B + CONN_B + NODE_B * ratio) * ne
    achieved = elem_bytes / avg_el / 1e9

    dom.close()
    fp32 = None
    if not args.no_fp32:
        fp32 = mixed_form(mesh, args, value, fp64_state)
    full_state = None
    if not args.no_full_state:
        full_state = full_state_leg(mesh, args)
    batched = None
    if not args.no_batch:
        batched = {form: batched_bars(form, engine=args.batch_engine) for form in ("fp64", "mixed")}
    shard = None
    if not args.no_shard:
        shard = shard_leg(args)
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_leg(mesh)
    pulse = None
    if not args.no_pulse:
        pulse = pulse_leg(args, elapsed / args.steps * 1e3)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mesh, args)
    copy = None if args.no_copy_peak else copy_peak()
    # per-launch counters of the element kernel from the committed profile of
    # the same window (warm-up and timed increments); none for another window
    window = "%d+%d" % (args.warmup, args.steps)
    traffic = profiled("traffic_bytes_per_launch", window)
    valu_insts = profiled("valu_insts_per_launch", window)

    out = {
        "metric": "element-steps/s, 1M hex8 FP64 vs FP32 (err vs FP64 oracle); batched bars/s",
        "value": value,
        "unit": "element-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated O-grid mesh, uniform 227 m/s initial velocity)",
        "config": {"workload": "hex8 O-grid Taylor bar, %d elements (2000/section x %d layers), %d nodes, FP64, "
                               "copper A=400MPa B=100MPa n=1, 227 m/s, central difference" % (ne, args.layers, nn),
                   "elements": ne, "nodes": nn, "form": "fp64 bit-exact vs oracle",
                   "sim_time_end_s": tinfo["t"], "dt_s": tinfo["dt"], "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_elem_hex8 (element stress/force pass)",
                     "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": traffic, "avg_ms": avg_el * 1e3, "launches": n_el,
                     "window": "increments %d..%d after the preamble (warm-up %d, timed %d)"
                               % (args.warmup + 1, args.warmup + args.steps, args.warmup, args.steps),
                     "traffic_source": None if traffic is None else
                     "profiles/k_elem_hex8_fp64_1M.md, FETCH_SIZE x 2.0 + WRITE_SIZE averaged over the same window",
                     "algorithmic_bytes_per_launch": elem_bytes, "copy_peak": copy},
        # the bound that applies to the bit-exact FP64 kernel: VALU issue, one
        # wave64 instruction per 4 cycles on each of 1,024 SIMDs at 2.4 GHz
        "valu_roofline": None if valu_insts is None else {
            "insts_per_launch": valu_insts, "achieved_Ginst_s": valu_insts / avg_el / 1e9,
            "peak_Ginst_s": VALU_PEAK_GINST, "frac": valu_insts / avg_el / 1e9 / VALU_PEAK_GINST},
        "step_roofline": {"algorithmic_bytes_per_step": step_bytes, "achieved_GBs": step_bytes / (elapsed / args.steps) / 1e9,
                          "nodal_kernel_avg_ms": avg_nd * 1e3},
        "fp32_form": fp32,
        "full_state": full_state,
        "batched": batched,
        "adjoint": adjoint,
        "shard": shard,
        "pulse": pulse,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def shape(mesh, x, evp):
    """final bar length (max z), mushroom radius (max r on the z = 0 layer) and peak plastic strain"""
    import numpy as np
    base = mesh["x"][:, 2] == 0.0
    return {"length": float(x[:, 2].max()), "radius": float(np.hypot(x[base, 0], x[base, 1]).max()),
            "max_evp": float(evp.max())}


def mixed_form(mesh, args, fp64_value, fp64_state=None):
    """The mixed-precision form (FP32 element pass, FP64 nodal state) on the
    same workload, and its error against the FP64 run at the end of the same
    increments: bar length, mushroom radius, peak plastic strain and the
    displacement field (max |du| / max |u|)."""
    import numpy as np
    import redynela_amd as R
    mat = R.Material(115e9, 0.3, 8930.0, R.IsoHardElastoplastic(400e6, 100e6, 1.0))
    sol = R.ExplicitSolverCH()
    sol.setTimeStepMethod("Courant")
    sol.setDissipation(1.0)
    sol.setPrecision("mixed")
    dom = R.Domain(mesh, mat, sol)
    from redynela_amd import _lib as L
    dom.set_option(L.OPT_STORE_FINT, 0)
    assert dom.begin(1.0)
    dom.step_async(args.warmup)
    dom.synchronize()
    dom.enable_kernel_timing(True)  # HIP events around the element launches, as in the FP64 window
    t0 = time.perf_counter()
    dom.step_async(args.steps)
    dom.synchronize()
    el = time.perf_counter() - t0
    n_el, ms_el = dom.kernel_stats(0)
    dom.enable_kernel_timing(False)
    ne = mesh["conn"].shape[0]
    v = ne * args.steps / el
    avg_el = ms_el / max(n_el, 1) * 1e-3
    achieved = MIXED_ELEM_B * ne / avg_el / 1e9
    out = {"value": v, "unit": "element-steps/s", "ms_per_step": el / args.steps * 1e3,
           "speedup_vs_fp64": v / fp64_value,
           "element_kernel_ms": avg_el * 1e3,
           "roofline": {"bound": "hbm", "kernel": "k_elem_hex8_f32 (FP32 element pass)", "achieved": achieved,
                        "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                        "algorithmic_bytes_per_element": MIXED_ELEM_B,
                        "counters": "profiles/k_elem_hex8_f32_1M.md"},
           "note": "same mesh, steps and warm-up; FP32 kinematics / stress / B^T sigma, FP64 nodal state and time"}
    if fp64_state is not None:
        x64, e64 = fp64_state
        x32, e32 = dom.coords, dom.evp
        a, b = shape(mesh, x64, e64), shape(mesh, x32, e32)
        u64, u32 = x64 - mesh["x"], x32 - mesh["x"]
        out["err_vs_fp64"] = {k: abs(b[k] - a[k]) / abs(a[k]) for k in a}
        out["err_vs_fp64"]["displacement"] = float(np.abs(u32 - u64).max() / np.abs(u64).max())
        out["err_vs_fp64"]["at"] = "end of the timed increments (%d after the preamble)" % (args.warmup + args.steps)
    dom.close()
    return out


def full_state_leg(mesh, args):
    """The FP64 form with the diagnostic point tensors the reference always
    updates (Eps, EpsPlas, dEps, FE/Element.C:362-391) kept: same mesh and
    window as the headline, which leaves them out (they never feed forces)."""
    import redynela_amd as R
    from redynela_amd import _lib as L
    mat = R.Material(115e9, 0.3, 8930.0, R.IsoHardElastoplastic(400e6, 100e6, 1.0))
    sol = R.ExplicitSolverCH()
    sol.setTimeStepMethod("Courant")
    sol.setDissipation(1.0)
    dom = R.Domain(mesh, mat, sol, track_strain=True)
    dom.set_option(L.OPT_STORE_FINT, 0)
    assert dom.begin(1.0)
    dom.step_async(args.warmup)
    dom.synchronize()
    dom.enable_kernel_timing(True)
    t0 = time.perf_counter()
    dom.step_async(args.steps)
    dom.synchronize()
    el = time.perf_counter() - t0
    n_el, ms_el = dom.kernel_stats(0)
    dom.close()
    ne = mesh["conn"].shape[0]
    return {"value": ne * args.steps / el, "unit": "element-steps/s", "ms_per_step": el / args.steps * 1e3,
            "element_kernel_ms": ms_el / max(n_el, 1),
            "note": "FP64 with Eps, EpsPlas and dEps per point also read and written (+1,728 B per element-step)"}


def batched_bars(form, n=512, t_end=40e-6, engine="auto"):
    """The batched workload: 512 seeded Taylor bars (nr ~ U{5..20}, nz = 5 nr,
    125..2,000 quads, drawn speed and hardening) advanced to 40 us; bars/s
    from HIP events around the whole solve, per-bar device wall time (first
    to last instruction that advanced the bar)."""
    import numpy as np
    import redynela_amd as R
    meshes, mats = R.batch.taylor_batch(n)
    sol = R.ExplicitSolverCH()
    sol.setTimeStepMethod("Courant")
    sol.setDissipation(1.0)
    sol.setPrecision(form)
    tuning = None
    if engine == "auto":
        # the autotuner probes the engines on this batch's first solve (48
        # increments each, from a copy of the state); a batch of the same
        # bars is then timed with the engine it chose, as a caller that
        # solves batches of this shape over and over would run it
        warm = R.Batch(meshes, mats, sol)
        warm.solve(t_end)
        choice, probes = warm.tuning()
        tuning = {"chosen": choice, "probe_ms": probes, "probe_increments": 48}
        engine = choice
    else:
        warm = R.Batch(meshes[:8], mats[:8], sol)  # loads the code object
        warm.set_engine(engine)
        warm.solve(t_end)
    warm.close()
    b = R.Batch(meshes, mats, sol)
    b.set_engine(engine)
    t0 = time.perf_counter()
    wall_us, ms = b.solve(t_end)
    host_s = time.perf_counter() - t0
    steps = np.array([b.time(i)["increment"] for i in range(n)])
    ne = np.array([m["conn"].shape[0] for m in meshes])
    b.close()
    nr, v0, sy0, hm = R.batch.taylor_params(n)
    slow = [{"bar": int(i), "nr": int(nr[i]), "v0": float(v0[i]), "sigma_y0": float(sy0[i]), "H": float(hm[i]),
             "increments": int(steps[i]), "wall_us": float(wall_us[i])} for i in np.argsort(-steps)[:5]]
    # algorithmic HBM bytes per axisymmetric element-step: point history 4 points x 8 words x 8 B x 2, connectivity
    # 16 B and nodal traffic 188 B (FP64); 444 B in the mixed form (FP32 point history)
    per_step = 716.0 if form == "fp64" else 444.0
    es = float((ne * steps).sum())
    roof = {"bound": "hbm", "algorithmic_bytes_per_element_step": per_step,
            "achieved": es * per_step / (ms * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": es * per_step / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "over": "the whole batched solve (element, scalar and nodal launches), device time",
            "counters": "profiles/batch_taylor512_counters.md"}
    return {"bars": n, "engine": engine, "autotune": tuning, "bars_per_s": n / (ms * 1e-3), "solve_ms": ms,
            "roofline": roof, "host_solve_s": host_s,
            "p50_us_per_bar": float(np.percentile(wall_us, 50)), "p99_us_per_bar": float(np.percentile(wall_us, 99)),
            "max_us_per_bar": float(wall_us.max()), "element_steps": int((ne * steps).sum()),
            "element_steps_per_s": float((ne * steps).sum() / (ms * 1e-3)),
            "increments_min_max": [int(steps.min()), int(steps.max())],
            "increments_p50_p99": [float(np.percentile(steps, 50)), float(np.percentile(steps, 99))],
            "elements_total": int(ne.sum()), "most_increments": slow}


def shard_leg(args, layers=64):
    """One GPU's share of the 1M bar split in axial slabs over 8 GPUs: 63
    layers plus a one-layer halo (128,000 elements), same material and speed,
    FP64; the element and nodal kernels at that shape."""
    import redynela_amd as R
    from redynela_amd import _lib as L
    mesh = R.mesh.ogrid_taylor_bar(layers)
    ne = mesh["conn"].shape[0]
    mat = R.Material(115e9, 0.3, 8930.0, R.IsoHardElastoplastic(400e6, 100e6, 1.0))
    sol = R.ExplicitSolverCH()
    sol.setTimeStepMethod("Courant")
    sol.setDissipation(1.0)
    d = R.Domain(mesh, mat, sol)
    d.set_option(L.OPT_STORE_FINT, 0)
    d.begin(1.0)
    d.step_async(args.warmup)
    d.synchronize()
    d.enable_kernel_timing(True)
    t0 = time.perf_counter()
    d.step_async(args.steps)
    d.synchronize()
    el = time.perf_counter() - t0
    n_el, ms_el = d.kernel_stats(0)
    n_nd, ms_nd = d.kernel_stats(1)
    d.close()
    return {"workload": "hex8 O-grid 2,000 el/section x %d layers (1/8 of the 1M bar + 1 halo layer), FP64" % layers,
            "elements": ne, "element_steps_per_s": ne * args.steps / el, "ms_per_step": el / args.steps * 1e3,
            "element_kernel_ms": ms_el / max(n_el, 1), "nodal_kernel_ms": ms_nd / max(n_nd, 1)}


def pulse_leg(args, courant_ms_per_step, steps=200):
    """The same 1M hex8 bar with the Pulse time step (the reference's default
    for ExplicitSolverCH, FE/ExplicitSolverCH.C:83): the cold first call of
    the preamble, then `steps` increments with a warm call every 10."""
    import redynela_amd as R
    from redynela_amd import _lib as L
    mesh = R.mesh.ogrid_taylor_bar(args.layers)
    ne = mesh["conn"].shape[0]
    mat = R.Material(115e9, 0.3, 8930.0, R.IsoHardElastoplastic(400e6, 100e6, 1.0))
    sol = R.ExplicitSolverCH()  # Pulse
    sol.setDissipation(1.0)
    d = R.Domain(mesh, mat, sol)
    d.set_option(L.OPT_STORE_FINT, 0)
    t0 = time.perf_counter()
    d.begin(1.0)
    cold_s = time.perf_counter() - t0
    cold = d.pulse_info()
    d.step(10)
    calls0 = d.pulse_info()["calls"]
    t0 = time.perf_counter()
    d.step_async(steps)
    d.synchronize()
    el = time.perf_counter() - t0
    info = d.pulse_info()
    tm = d.time()
    d.close()
    calls = info["calls"] - calls0
    ms = el / steps * 1e3
    return {"workload": "hex8 O-grid, %d elements, FP64, Pulse time step" % ne,
            "elements": ne, "element_steps_per_s": ne * steps / el, "ms_per_step": ms,
            "preamble_s": cold_s, "cold_iterations": cold["iterations"],
            "warm_calls": calls, "last_call_iterations": info["iterations"],
            "ms_per_warm_call": (ms - courant_ms_per_step) * steps / max(calls, 1),
            "dt_s": tm["dt"], "stiffness_bytes": ne * 24 * 24 * 8}


def adjoint_leg(hex_mesh):
    """Gradient of the final-shape misfit w.r.t. (A, B, n) on the axisymmetric
    20 x 100 bar, 4,000 increments, checkpoints every 100; time and memory of
    the reverse sweep against one forward run.  The same on the 1M hex8 bar
    (100 increments, checkpoints every 10), timing only: the gradient is
    checked against the oracle on a small hex8 bar in the tests."""
    import numpy as np
    from redynela_amd import adjoint as AD
    mesh = AD.bar_20x100()
    p = AD.HardeningAdjoint(mesh)
    xt = p.forward(AD.TARGET_THETA)[:, :2].copy()
    t0 = time.perf_counter()
    x = p.forward(AD.perturbed_start())[:, :2]
    g = p.backward(2 * (x - xt))
    wall = time.perf_counter() - t0
    st = p.stats
    p.close()
    # the same gradient with the state of every increment kept: no
    # recomputation, memory for time
    q = AD.HardeningAdjoint(mesh, ckpt_every=1)
    xq = q.forward(AD.perturbed_start())[:, :2]
    gq = q.backward(2 * (xq - xt))
    sq = q.stats
    q.close()
    # the Johnson-Cook temperature factor with adiabatic heating (no rate term)
    r = AD.HardeningAdjoint(mesh, thermal=AD.THERMAL)
    xr_t = r.forward(AD.TARGET_THETA)[:, :2].copy()
    xr = r.forward(AD.perturbed_start())[:, :2]
    gr = r.backward(2 * (xr - xr_t))
    sr = r.stats
    r.close()
    hx = AD.HardeningAdjoint(hex_mesh, nsteps=100, ckpt_every=10)
    xh_t = hx.forward(AD.TARGET_THETA).copy()
    xh = hx.forward(AD.perturbed_start())
    gh = hx.backward(2 * (xh - xh_t))
    sh = hx.stats
    hx.close()
    return {"workload": "axi 20x100 quads, 200 m/s, 4000 increments, checkpoints every 100, FP64",
            "grad": [float(v) for v in g], "forward_ms": st["forward_ms"], "backward_ms": st["backward_ms"],
            "recompute_ms": st["recompute_ms"], "backward_over_forward": st["backward_ms"] / st["forward_ms"],
            "device_MB": st["device_bytes"] / 1e6, "checkpoint_MB": st["checkpoint_bytes"] / 1e6,
            "checkpoints": st["n_checkpoints"], "wall_s": wall,
            "every_increment": {"forward_ms": sq["forward_ms"], "backward_ms": sq["backward_ms"],
                                "backward_over_forward": sq["backward_ms"] / sq["forward_ms"],
                                "device_MB": sq["device_bytes"] / 1e6, "same_gradient": bool(np.array_equal(g, gq))},
            "thermal": {"law": "Johnson-Cook temperature factor and adiabatic heating, m 1.09, T0 293 K, "
                               "Tm 1356 K, cp 383 J/kg/K, Taylor-Quinney 0.9, no rate term",
                        "grad": [float(v) for v in gr], "forward_ms": sr["forward_ms"],
                        "backward_ms": sr["backward_ms"], "backward_over_forward": sr["backward_ms"] / sr["forward_ms"],
                        "device_MB": sr["device_bytes"] / 1e6},
            "hex8": {"workload": "hex8 O-grid, %d elements, 227 m/s, power law (A, B, n), 100 increments, "
                                 "checkpoints every 10, FP64" % hex_mesh["conn"].shape[0],
                     "grad": [float(v) for v in gh], "forward_ms": sh["forward_ms"],
                     "backward_ms": sh["backward_ms"], "recompute_ms": sh["recompute_ms"],
                     "backward_over_forward": sh["backward_ms"] / sh["forward_ms"],
                     "device_MB": sh["device_bytes"] / 1e6, "checkpoint_MB": sh["checkpoint_bytes"] / 1e6}}


def hex_lengths(x, conn, chunk=262144):
    """the hex8 Courant length V / A_max of every element (the kernels' and
    the oracle's build-defined length), vectorised in chunks"""
    import numpy as np
    sg = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]],
                  dtype=float)
    g = sg / np.sqrt(3.0)
    dn0 = np.zeros((8, 8, 3))
    for p in range(8):
        for k in range(8):
            f = (1 + sg[k] * g[p]) / 2.0
            for d in range(3):
                dn0[p, k, d] = sg[k, d] / 2.0 * np.prod([f[e] for e in range(3) if e != d])
    faces = np.array([[0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7], [7, 4, 5, 6], [2, 1, 0, 3]])
    out = np.empty(conn.shape[0])
    for c0 in range(0, conn.shape[0], chunk):
        xe = x[conn[c0:c0 + chunk]]  # [n, 8, 3]
        J = np.einsum("pkd,nki->npid", dn0, xe)
        V = np.linalg.det(J).sum(axis=1)
        d1 = xe[:, faces[:, 2]] - xe[:, faces[:, 0]]
        d2 = xe[:, faces[:, 3]] - xe[:, faces[:, 1]]
        A = 0.5 * np.linalg.norm(np.cross(d1, d2), axis=2).max(axis=1)
        out[c0:c0 + chunk] = V / A
    return out


def courant_element(mesh, d, form):
    """where the Courant time step is set: the element of least V / A_max,
    its centroid, axial and radial extent against the initial ones, and the
    largest plastic strain and temperature of its points"""
    import numpy as np
    from redynela_amd import _lib as L
    x, x0, conn = d.coords, mesh["x"], mesh["conn"]
    ne = conn.shape[0]
    lens = hex_lengths(x, conn)
    e = int(np.argmin(lens))
    xe, xe0 = x[conn[e]], x0[conn[e]]
    c = xe.mean(axis=0)
    ext = lambda p: (float(np.ptp(p[:, 2])), float(np.ptp(np.hypot(p[:, 0], p[:, 1]))))
    (hz, hr), (hz0, hr0) = ext(xe), ext(xe0)
    # point fields: [element][point] in the FP64 form, [point][element] in the mixed one
    pt = (lambda f: f.reshape(ne, 8)[e]) if form == "fp64" else (lambda f: f.reshape(8, ne)[:, e])
    out = {"element": e, "length_m": float(lens[e]), "initial_length_m": float(hex_lengths(x0, conn[e:e + 1])[0]),
           "centroid_r_m": float(np.hypot(c[0], c[1])), "centroid_z_m": float(c[2]),
           "axial_extent_m": hz, "initial_axial_extent_m": hz0, "radial_extent_m": hr,
           "initial_radial_extent_m": hr0, "evp_max_of_points": float(pt(d.field(L.F_EVP)).max())}
    T = d.field(L.F_TEMP)
    if T.size:
        out["temperature_max_of_points_K"] = float(pt(T).max())
    return out


def precision_report(layers=1000, t_end=60e-6, forms=("fp64", "mixed")):
    """The O-grid bar of 4,000 hexahedra per section (40x40 core + 4 rings of
    40x15; 1,000 layers = 4.0M elements), 250 m/s, Johnson-Cook with rate and
    adiabatic heating, to t_end in the FP64 and the mixed form: final length,
    mushroom radius and peak plastic strain, their relative differences, and
    the element that sets the Courant time step at the end."""
    import numpy as np
    import redynela_amd as R
    mesh = R.mesh.ogrid_bar(40, 15, layers, 3.2e-3, 32.4e-3, 0.4, 250.0)
    base = mesh["x"][:, 2] == 0.0
    out = {"workload": "hex8 O-grid 4,000 el/section x %d layers (%d elements), 250 m/s, Johnson-Cook "
                       "A=90MPa B=292MPa n=0.31 C=0.025 m=1.09 adiabatic, to %.0f us"
                       % (layers, mesh["conn"].shape[0], t_end * 1e6)}
    for form in forms:
        h = R.IsoHardJohnsonCook(90e6, 292e6, 0.31, C=0.025, m=1.09, depsp0=1.0, Tm=1356.0, T0=293.0)
        h.enableRateAndThermal(383.0, 0.9)
        sol = R.ExplicitSolverCH()
        sol.setTimeStepMethod("Courant")
        sol.setDissipation(1.0)
        sol.setPrecision(form)
        d = R.Domain(mesh, R.Material(115e9, 0.3, 8930.0, h), sol)
        d.set_option(2, 0)
        t0 = last = time.perf_counter()
        d.begin(t_end)
        dt0 = d.time()["dt"]
        while d.step(256):
            if time.perf_counter() - last > 15.0:
                last = time.perf_counter()
                tt = d.time()
                print("  %s: increment %d, t = %.3e s, dt = %.3e s" % (form, tt["increment"], tt["t"], tt["dt"]),
                      file=sys.stderr, flush=True)
        el = time.perf_counter() - t0
        tm = d.time()
        x = d.coords
        out[form] = {"increments": tm["increment"], "t_end": tm["t"], "solve_s": el, "length": float(x[:, 2].max()),
                     "radius": float(np.hypot(x[base, 0], x[base, 1]).max()), "max_evp": float(d.evp.max()),
                     "first_dt": dt0, "last_computed_dt": tm["computed_dt"],
                     "courant_element": courant_element(mesh, d, form)}
        d.close()
    if "fp64" in out and "mixed" in out:
        a, b = out["fp64"], out["mixed"]
        out["rel_err"] = {k: abs(b[k] - a[k]) / abs(a[k]) for k in ("length", "radius", "max_evp")}
        out["speedup_mixed_vs_fp64"] = a["solve_s"] / b["solve_s"]
    return out


def profiled(key, window):
    """A per-launch figure of the element kernel from the committed rocprofv3
    summary, for the bench window `warmup+steps`: traffic_bytes_per_launch
    (FETCH_SIZE / WRITE_SIZE, calibrated against a copy kernel) or
    valu_insts_per_launch (SQ_INSTS_VALU); None when that window has not been
    profiled."""
    p = os.path.join(ROOT, "profiles", "k_elem_hex8_fp64_1M.md")
    tag = "%s@%s:" % (key, window)
    try:
        for line in open(p):
            if line.startswith(tag):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_model():
    """the host CPU's model name (lscpu, else /proc/cpuinfo) and its logical CPU count"""
    import subprocess
    name = None
    try:
        r = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10)
        for line in r.stdout.splitlines():
            if line.startswith("Model name:"):
                name = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    if not name:
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    return name or "unknown", os.cpu_count() or 0


def copy_peak():
    """A plain device copy of 1 GiB (tests/capi/hbm_copy: three launch shapes
    of a streaming copy kernel, read + write bytes counted, best of 10 each),
    timed with HIP events in a child process: the HBM bandwidth a streaming
    copy reaches on this box."""
    import subprocess
    prog = os.path.join(ROOT, "tests", "capi", "hbm_copy")
    if not os.path.exists(prog):
        return {"error": "tests/capi/hbm_copy is not built"}
    try:
        r = subprocess.run([prog, "1024", "10"], capture_output=True, text=True, timeout=120)
    except subprocess.TimeoutExpired:
        return {"error": "hbm_copy timed out"}
    if r.returncode != 0:
        return {"error": "hbm_copy exited with %d: %s" % (r.returncode, r.stderr.strip()[-300:])}
    c = json.loads(r.stdout.strip().splitlines()[-1])
    return {"GB_s": c["best_GB_s"], "mean_GB_s": c["mean_GB_s"], "bytes_per_copy": c["bytes_read"] + c["bytes_written"],
            "kernel": c["kernel"], "blocks_per_cu": c["blocks_per_cu"],
            "how": "tests/capi/hbm_copy: the fastest of three launch shapes of a 16-byte-per-lane copy kernel "
                   "(grid-stride, or four loads in flight per lane), 1 GiB read + 1 GiB written"}


def cpu_baseline(mesh, args):
    """The CPU baseline: the oracle (a C restatement of the reference's path)
    on the host's cores, same mesh and material, a bounded sample of
    increments for each assembly strategy of Domain::computeInternalForces
    restated from the reference (FE/Domain.C:130-738: case 2, dynamic
    element list; 3, static centroid-sorted jobs with a critical-section sum;
    6, static jobs with gatherFrom and a fixed-order reduction; 7, the same
    with forward loops) and the oracle's own order (element-parallel work,
    sequential element-order assembly); the fastest of the reference's
    strategies is the baseline."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        import oracle_lib as O
    except Exception as e:  # pragma: no cover
        return {"value": None, "unit": "element-steps/s", "cores": 0, "kind": "port", "sample": "unavailable: %s" % e}
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    O.lib().dnlo_set_threads(threads)
    ne = mesh["conn"].shape[0]
    m = O.Model(mesh, dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0), dict(rho_b=1.0), mathmode=1)
    m.begin(1.0)
    m.step()
    table = {}
    total = 0.0
    for name, strategy in (("oracle element order", 0), ("case 2", 2), ("case 3", 3), ("case 6", 6), ("case 7", 7)):
        m.set_strategy(strategy, threads)
        m.step()  # warm-up: partition, per-thread vectors
        m.forces_seconds(reset=True)
        t0 = time.perf_counter()
        for _ in range(args.cpu_steps):
            m.step()
        dt = time.perf_counter() - t0
        fs = m.forces_seconds()
        total += dt
        table[name] = {"element_steps_per_s": ne * args.cpu_steps / dt, "ms_per_step": dt / args.cpu_steps * 1e3,
                       "internal_forces_ms_per_step": fs / args.cpu_steps * 1e3}
    # the baseline is the reference's fastest strategy; the oracle's own order stays in the table
    best = max((k for k in table if k.startswith("case")), key=lambda k: table[k]["element_steps_per_s"])
    model, ncpu = cpu_model()
    return {"value": table[best]["element_steps_per_s"], "unit": "element-steps/s", "cores": threads, "kind": "port",
            "strategy": best, "strategies": table, "cpu_model": model, "logical_cpus": ncpu,
            "sample": "%d increments per strategy of the same %d-element mesh (after 1 warm-up increment each), "
                      "whole explicit step (predictor, element pass and assembly, corrector, time step; "
                      "internal_forces_ms_per_step is the Domain::computeInternalForces part, the region the "
                      "reference times), oracle in det math, OpenMP %d threads; %.1f s timed in all"
                      % (args.cpu_steps, ne, threads, total)}


if __name__ == "__main__":
    main()
