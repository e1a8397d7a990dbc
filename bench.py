"""Benchmark: Mcell-updates/s of the Yee/CPML time step on one MI355X.

Workload (default): slab512 -- 512x512x256 cells, 0.5 mm, CPML 16 on every
face, 256 seeded dielectric boxes, 8 lumped z-sources, 128 Ez probes with an
FP64 DFT at 16 frequencies -- one GPU's share of the reference's z-slab run.
A step is one full time step (E half-step then H half-step) over the whole
grid; inputs are resident in HBM before timing starts.

Prints one JSON line with the FP32 form as `value`, the FP16-storage form
beside it, the roofline of the dominant kernel (algorithmic bytes per launch
over its HIP-event launch time) and the CPU oracle timed on a bounded sample
on the host as a reported baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fdtd_amd import api, scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-updates/s and % HBM peak; FP16-store vs FP32 at stated probe error"


def b_alg(sc, info, storage_bytes):
    """Algorithmic HBM bytes per cell for the E and the H sweep.

    Each sweep reads and writes its own three components and reads the other
    three (9 values); a classed sweep reads one class index per cell; each
    CPML psi (2 per axis per sweep) is read and written over its slab."""
    s = storage_bytes
    f = [(sc.cpml.thick[0] + sc.cpml.thick[1]) / sc.nx, (sc.cpml.thick[2] + sc.cpml.thick[3]) / sc.ny,
         (sc.cpml.thick[4] + sc.cpml.thick[5]) / sc.nz]
    psi = 4.0 * s * sum(f)
    cls = {0: 0.0}
    ncls = info["n_classes"]
    cls_bytes = 0.0 if ncls <= 1 else (2.0 if ncls <= 65536 else 4.0)
    e = 9 * s + cls_bytes + psi
    h = 9 * s + psi   # H coefficients are uniform in these workloads (mu_r = 1, sigma_m = 0)
    del cls
    return e, h


def run_form(sc, storage, steps, warmup, profile_steps):
    sim = api.Simulation.from_scene(sc, storage, n_steps=warmup + steps + profile_steps)
    sim.run(warmup)
    t0 = time.perf_counter()
    sim.run(steps)
    wall = time.perf_counter() - t0
    inf = sim.info()
    ev_ms = inf["last_run_ms"]
    if profile_steps > 0:
        sim.profile(profile_steps)
    prof = sim.info()
    probes = sim.probes()[:, warmup:warmup + steps]
    sim.close()
    return dict(wall_s=wall, event_ms=ev_ms, e_ms=prof["e_kernel_ms"], h_ms=prof["h_kernel_ms"],
                info=inf, probes=probes)


def cpu_baseline(sc, seconds_target=15.0):
    """The CPU oracle (FP32 restatement of the reference's equations, OpenMP
    over z-planes) on the same workload, timed over its first few steps."""
    from oracle.oracle import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = Oracle(sc, 32, nsteps=64)
    t0 = time.perf_counter()
    o.steps(1)
    one = time.perf_counter() - t0
    n = int(max(1, min(63, seconds_target / max(one, 1e-3))))
    t0 = time.perf_counter()
    o.steps(n)
    dt = time.perf_counter() - t0
    o.close()
    return {"value": sc.cells * n / dt / 1e6, "unit": "Mcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle/fdtd_oracle.cpp FP32, OpenMP {threads} threads, steps 2-{n + 1} of "
                      f"the full {sc.nx}x{sc.ny}x{sc.nz} {sc.name} grid ({dt:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="slab512", choices=sorted(scenes.WORKLOADS))
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only time the plain HBM copy kernel (measurement calibration)")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--copy-iters", type=int, default=20)
    args = ap.parse_args()
    if args.copy_probe:
        import ctypes
        lib = ctypes.CDLL(os.path.join(ROOT, "fdtd_amd", "libhbmcopy.so"))
        lib.hbm_copy_gbs.restype = ctypes.c_double
        lib.hbm_copy_gbs.argtypes = [ctypes.c_uint64, ctypes.c_int]
        gbs = lib.hbm_copy_gbs(args.copy_bytes, args.copy_iters)
        print(json.dumps({"copy_gbs": round(gbs, 1), "bytes_each_way": args.copy_bytes,
                          "iters": args.copy_iters, "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}))
        return
    if args.gpus != 1:
        print(f"bench.py runs one GPU: this build has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    sc = scenes.WORKLOADS[args.workload]()
    cells = sc.cells
    f32 = run_form(sc, "f32", args.steps, args.warmup, args.profile_steps)
    ms_per_step = f32["event_ms"] / args.steps
    value = cells * args.steps / (f32["event_ms"] * 1e-3) / 1e6
    be, bh = b_alg(sc, f32["info"], 4)
    dom_e = f32["e_ms"] >= f32["h_ms"]
    kms = f32["e_ms"] if dom_e else f32["h_ms"]
    bpc = be if dom_e else bh
    achieved = bpc * cells / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    step_gbs = (be + bh) * cells / (ms_per_step * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": sc.name, "grid": [sc.nx, sc.ny, sc.nz], "cpml": sc.cpml.thick[0],
                   "boxes": len(sc.boxes), "sources": len(sc.sources), "probes": len(sc.probes),
                   "dft_freqs": len(sc.freqs), "seed": sc.seed,
                   "coefficient_classes": f32["info"]["n_classes"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "e_kernel" if dom_e else "h_kernel",
                     "bytes_per_cell": round(bpc, 3), "kernel_ms": round(kms, 5)},
        "kernels": {"e_ms": round(f32["e_ms"], 5), "h_ms": round(f32["h_ms"], 5),
                    "b_alg_e": round(be, 3), "b_alg_h": round(bh, 3),
                    "e_frac": round(be * cells / (max(f32["e_ms"], 1e-9) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    "h_frac": round(bh * cells / (max(f32["h_ms"], 1e-9) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    "step_frac": round(step_gbs / HBM_PEAK_GBS, 4)},
        "wall_s": round(f32["wall_s"], 3),
    }
    if not args.no_fp16:
        f16 = run_form(sc, "f16", args.steps, args.warmup, args.profile_steps)
        v16 = cells * args.steps / (f16["event_ms"] * 1e-3) / 1e6
        e16, h16 = b_alg(sc, f16["info"], 2)
        p32 = f32["probes"].astype(np.float64)
        p16 = f16["probes"].astype(np.float64)
        live = np.linalg.norm(p32, axis=1) > 0
        rel = np.linalg.norm(p16 - p32, axis=1)[live] / np.linalg.norm(p32, axis=1)[live]
        out["fp16"] = {"value": round(v16, 1), "ms_per_step": round(f16["event_ms"] / args.steps, 5),
                       "speedup_vs_fp32": round(v16 / value, 3),
                       "probe_rel_l2_vs_fp32_max": float(f"{rel.max():.3e}"),
                       "probe_rel_l2_vs_fp32_median": float(f"{np.median(rel):.3e}"),
                       "e_ms": round(f16["e_ms"], 5), "h_ms": round(f16["h_ms"], 5),
                       "e_frac": round(e16 * cells / (max(f16["e_ms"], 1e-9) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "h_frac": round(h16 * cells / (max(f16["h_ms"], 1e-9) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
