"""Accuracy and adjoint reports on MI355X, checked against the CPU oracle.

  python tests/reports.py accuracy yee64:fp32_loop,dielectric256:fp32_loop,slab512
      probe and DFT error of every storage form against the FP64 oracle
  python tests/reports.py adjoint --workload adjoint128 --fd-cells 16
  python tests/reports.py adjoint --workload adjoint768 --storage f16 --fd-cells 4 --checkpoint-every 64
      adjoint wall time over a forward run, and gradient checks

The oracle is the checker here (test infrastructure), never the thing
measured.  Results go to stdout as one JSON line, progress to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from fdtd_amd import api, scenes  # noqa: E402


def _progress(msg):
    """A timestamped progress line on stderr (long reports run for minutes)."""
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def accuracy_report(names, storages=("f32", "f16", "bf16")):
    """Probe and DFT error of each storage form against the FP64 oracle (the
    reference's equations restated in FP64, run on the host) on the named
    workloads at their full sizes and step counts.

    Probe error: relative L2 of each probe's series, over the probes that
    carry signal (row norm >= 1e-3 of the largest); DFT error: |X - X64|
    over the probe's spectral peak max_f |X64|.  With fp32_loop the FP32
    restatement is run too and the FP32 engine is compared with it
    bit for bit (the parity bar of the FP32 form)."""
    from oracle.oracle import Oracle
    out = {}
    for name in names:
        name, _, flag = name.partition(":")
        sc = scenes.WORKLOADS[name]()
        t0 = time.perf_counter()
        o = Oracle(sc, 64)
        for s in range(0, sc.steps, 250):
            o.steps(min(250, sc.steps - s))
            _progress(f"{name}: FP64 oracle at step {min(s + 250, sc.steps)} of {sc.steps}")
        ref, X64 = o.probes(), (o.dft() if sc.freqs else None)
        o.close()
        rec = {"grid": [sc.nx, sc.ny, sc.nz], "steps": sc.steps, "probes": len(sc.probes),
               "fp64_oracle_s": round(time.perf_counter() - t0, 1)}
        ref32 = None
        if flag == "fp32_loop":
            o = Oracle(sc, 32)
            for s in range(0, sc.steps, 250):
                o.steps(min(250, sc.steps - s))
                _progress(f"{name}: FP32 host loop at step {min(s + 250, sc.steps)} of {sc.steps}")
            ref32 = o.probes()
            o.close()
        rn = np.linalg.norm(ref, axis=1)
        sig = rn >= 1e-3 * rn.max()
        for st in storages:
            sim = api.Simulation.from_scene(sc, st)
            sim.run(sc.steps)
            got = sim.probes()[:, :sc.steps].astype(np.float64)
            X = sim.dft() if sc.freqs else None
            sim.close()
            err = np.linalg.norm(got - ref, axis=1)[sig] / rn[sig]
            r = {"probe_rel_l2_max": float(f"{err.max():.3e}"), "probe_rel_l2_median": float(f"{np.median(err):.3e}"),
                 "probes_compared": int(sig.sum())}
            if X is not None:
                r["dft_rel_err_max"] = float(f"{(np.abs(X - X64) / np.abs(X64).max(axis=1, keepdims=True)).max():.3e}")
            if st == "f32" and ref32 is not None:
                r["max_abs_diff_vs_fp32_host_loop"] = float(np.abs(got - ref32).max())
            rec[st] = r
        out[name] = rec
        print(json.dumps({name: rec}), file=sys.stderr, flush=True)
    return out


def adjoint_report(args):
    """Adjoint gradient of J = sum |DFT_Ez(probe, f)|^2 w.r.t. the design
    cells' eps_r.

    Timing: adjoint wall time over one forward run of the same steps, for
    the checkpointed path (forward checkpoints every K steps, each segment
    recomputed) and for the stored-history path (the design region's E
    history kept for the whole run, no recompute).

    Checks (seeded design cells): central differences with h = 1e-3 of the
    FP64 oracle's forward runs (host cores) and the FP64 oracle's adjoint,
    for the small workload; for the large one, central differences of FP32
    engine forward runs (per cell at the --fd-h steps, and along whole-region
    directions with --fd-dirs) and the FP32 engine adjoint as the reference
    of a 16-bit-storage gradient."""
    wl = args.workload if args.workload.startswith("adjoint") else "adjoint128"
    sc = scenes.WORKLOADS[wl](**({"steps": args.adjoint_steps} if args.adjoint_steps > 0 else {}))
    terms = [(p, f, 1.0) for p in range(len(sc.probes)) for f in range(len(sc.freqs))]
    _progress(f"{wl}: building the {args.storage} context")
    sim = api.Simulation.from_scene(sc, args.storage)
    _progress(f"{wl}: forward run of {sc.steps} steps")
    sim.run(sc.steps)
    fwd_ms = sim.info()["last_run_ms"]
    out = {"workload": wl, "grid": [sc.nx, sc.ny, sc.nz], "steps": sc.steps, "storage": args.storage,
           "design_cells": int(np.prod(sc.design_eps.shape)), "forward_ms": round(fwd_ms, 3)}
    grads = {}
    for K in (args.checkpoint_every, 0):
        _progress(f"{wl}: adjoint, checkpoint interval {K} (0: stored design history)")
        sim.reset()
        t0 = time.perf_counter()
        J, g = sim.adjoint_gradient(terms, sc.steps, K)
        inf = sim.info()
        grads[K] = g
        out["checkpoint_every_%d" % K if K else "stored_history"] = {
            "adjoint_total_ms": round(inf["adjoint_total_ms"], 3), "segments": inf["adjoint_segments"],
            "peak_device_gb": round(inf["adjoint_device_bytes"] / 1e9, 2),
            "ratio_vs_forward": round(inf["adjoint_total_ms"] / fwd_ms, 3), "wall_s": round(time.perf_counter() - t0, 3)}
    out["J"] = J
    g = grads[args.checkpoint_every]
    out["checkpointed_equals_stored"] = bool(np.array_equal(grads[args.checkpoint_every], grads[0]))
    sim.close()
    rng = np.random.default_rng(sc.seed)
    cells = rng.choice(g.size, min(args.fd_cells, g.size), replace=False)
    eps = sc.design_eps.astype(np.float32)
    gmax = float(np.abs(g).max())
    if wl == "adjoint128" and not args.no_cpu:
        from oracle.oracle import Oracle
        t0 = time.perf_counter()
        _progress(f"{wl}: FP64 oracle adjoint on the host")
        Jo, go = Oracle(sc, 64).adjoint(terms)
        out["fp64_cpu_adjoint"] = {"grad_rel_l2": float(np.linalg.norm(g - go) / np.linalg.norm(go)),
                                   "J_rel": abs(J - Jo) / Jo, "cpu_s": round(time.perf_counter() - t0, 1)}
        h = 1e-3
        errs, errs64 = [], []
        t0 = time.perf_counter()
        for q in cells:
            _progress(f"{wl}: FP64 oracle central difference, cell {q}")
            vals, ev = [], []
            for sgn in (1, -1):
                e2 = eps.copy()
                e2[q] = np.float32(eps[q] + sgn * h)
                ev.append(float(e2[q]))
                s2 = scenes.WORKLOADS[wl](steps=sc.steps)
                s2.design_eps = e2
                o = Oracle(s2, 64)
                o.steps(s2.steps)
                X = o.dft()
                o.close()
                vals.append(sum(w * abs(X[p, f]) ** 2 for p, f, w in terms))
            fd = (vals[0] - vals[1]) / (ev[0] - ev[1])
            errs.append(abs(g[q] - fd) / gmax)
            errs64.append(abs(go[q] - fd) / gmax)
        out["fd_fp64_oracle"] = {"cells": len(cells), "h": h,
                                 "gpu_err_over_max_grad": {"median": float(np.median(errs)), "max": float(np.max(errs))},
                                 "fp64_adjoint_err_over_max_grad": {"max": float(np.max(errs64))},
                                 "cpu_s": round(time.perf_counter() - t0, 1)}
    else:
        _progress(f"{wl}: building the FP32 reference context")
        ref = api.Simulation.from_scene(sc, "f32")
        if args.storage != "f32":
            _progress(f"{wl}: FP32 adjoint (stored design history)")
            _, g32 = ref.adjoint_gradient(terms, sc.steps, 0)
            out["grad_rel_l2_vs_fp32_adjoint"] = float(np.linalg.norm(g - g32) / np.linalg.norm(g32))
        else:
            g32 = g
        out["fd_fp32_engine"] = []
        for h in args.fd_h:
            errs = []
            for q in cells:
                _progress(f"{wl}: FP32 engine central difference, h {h}, cell {q}")
                vals, ev = [], []
                for sgn in (1, -1):
                    e2 = eps.copy()
                    e2[q] = np.float32(eps[q] + sgn * h)
                    ev.append(float(e2[q]))
                    ref.set_design_eps(e2)
                    ref.run(sc.steps)
                    X = ref.dft()
                    vals.append(sum(w * abs(X[p, f]) ** 2 for p, f, w in terms))
                fd = (vals[0] - vals[1]) / (ev[0] - ev[1])
                errs.append(abs(g32[q] - fd) / float(np.abs(g32).max()))
            out["fd_fp32_engine"].append(
                {"cells": len(cells), "h": h,
                 "fp32_adjoint_err_over_max_grad": {"median": float(np.median(errs)) if errs else None,
                                                    "max": float(np.max(errs)) if errs else None,
                                                    "each": [float(f"{e:.3e}") for e in errs]}})
        if args.fd_dirs > 0:
            # Directional central differences: J(eps + h v) - J(eps - h v)
            # against g . (e+ - e-) over every design cell at once (v = sign
            # of the gradient, then seeded random signs).  The change in J is
            # a sum over the whole design region, far above the FP32
            # forwards' rounding of J, so h = 1e-3 is usable here.
            h = args.fd_dir_h
            rngd = np.random.default_rng(sc.seed + 1)
            dirs = [("sign(g)", np.sign(g32).astype(np.float32))]
            for d in range(1, args.fd_dirs):
                dirs.append((f"random {d}", rngd.choice(np.array([-1.0, 1.0], dtype=np.float32), g32.size)))
            res = []
            for name, v in dirs:
                _progress(f"{wl}: FP32 engine directional difference, h {h}, direction {name}")
                vals, es = [], []
                for sgn in (1, -1):
                    e2 = (eps + np.float32(sgn * h) * v).astype(np.float32)
                    es.append(e2.astype(np.float64))
                    ref.set_design_eps(e2)
                    ref.run(sc.steps)
                    X = ref.dft()
                    vals.append(sum(w * abs(X[p, f]) ** 2 for p, f, w in terms))
                fd = vals[0] - vals[1]
                pred = float(np.dot(g32, es[0] - es[1]))
                pred16 = float(np.dot(g, es[0] - es[1]))
                res.append({"direction": name, "dJ_fd": fd, "dJ_fp32_adjoint": pred, "dJ_adjoint": pred16,
                            "rel_err_fp32_adjoint": abs(fd - pred) / abs(pred),
                            "rel_err_adjoint": abs(fd - pred16) / abs(pred16)})
            out["fd_directional_fp32_engine"] = {"h": h, "directions": res}
        ref.close()
    out["grad_max_abs"] = gmax
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("report", choices=["accuracy", "adjoint"])
    ap.add_argument("workloads", nargs="?", default="yee64:fp32_loop,dielectric256:fp32_loop,slab512",
                    help="accuracy: comma-separated workloads (name[:fp32_loop])")
    ap.add_argument("--workload", default="adjoint128", choices=["adjoint128", "adjoint768"])
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"])
    ap.add_argument("--checkpoint-every", type=int, default=64)
    ap.add_argument("--fd-cells", type=int, default=16)
    ap.add_argument("--fd-h", type=lambda v: [float(x) for x in v.split(",")], default=[5e-2],
                    help="adjoint768: comma-separated eps_r steps of the FP32-engine central differences")
    ap.add_argument("--adjoint-steps", type=int, default=0, help="steps of the run (0: the workload's own)")
    ap.add_argument("--fd-dirs", type=int, default=0,
                    help="adjoint, large workload: directional central differences (sign(g), then random signs)")
    ap.add_argument("--fd-dir-h", type=float, default=1e-3, help="step of the directional differences")
    ap.add_argument("--no-cpu", action="store_true", help="adjoint: skip the host FP64 checks")
    args = ap.parse_args()
    if args.report == "accuracy":
        print(json.dumps({"accuracy_vs_fp64_oracle": accuracy_report(args.workloads.split(","))}), flush=True)
    else:
        print(json.dumps(adjoint_report(args)), flush=True)


if __name__ == "__main__":
    main()
