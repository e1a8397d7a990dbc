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

# HBM bytes per time step measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (separate runs) on MI355X, scaled by what the same counters read
# for a copy kernel of known size (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0);
# summaries and commands in profiles/split_sweeps_slab512_hbm_traffic.md.
# Keyed by (workload, storage, kernel path); re-measure after kernel changes.
MEASURED_TRAFFIC = {
    # E and H sweeps, each the sum of its interior, x-edge and general launches
    ("slab512", "f32", "split"): ((600573.0 + 216707.7 + 215343.1) * 2.0 + (483868.2 + 188264.3 + 231356.3) +
                                  (618460.7 + 196852.4 + 217374.5) * 2.0 + (483840.0 + 188206.0 + 239390.0)) * 1024,
    ("slab512", "f16", "split"): ((214628.7 + 221534.2 + 111830.4) * 2.0 + (161298.1 + 174776.8 + 115636.4) +
                                  (193349.0 + 190926.0 + 100927.2) * 2.0 + (161280.0 + 174803.8 + 119693.1)) * 1024,
}
METRIC = "Mcell-updates/s and % HBM peak; FP16-store vs FP32 at stated probe error"


def b_alg(sc, info, storage_bytes):
    """Algorithmic HBM bytes per cell-update of one time step.

    One-kernel step: every field component is read once and written once (12
    values); a classed E update reads one class index per cell; each CPML psi
    (4 per axis) is read and written over its slab (fraction 2L/N of the
    axis).  With the two half-step kernels each sweep reads and writes its
    own three components and reads the other three (18 values per step)."""
    s = storage_bytes
    f = [(sc.cpml.thick[0] + sc.cpml.thick[1]) / sc.nx, (sc.cpml.thick[2] + sc.cpml.thick[3]) / sc.ny,
         (sc.cpml.thick[4] + sc.cpml.thick[5]) / sc.nz]
    psi = 8.0 * s * sum(f)
    ncls = info["n_classes"]
    cls_bytes = 0.0 if ncls <= 1 else (2.0 if ncls <= 65536 else 4.0)
    return 18 * s + cls_bytes + psi


def _traffic(workload, storage, info):
    """Measured HBM bytes per step for this workload and kernel path, or None."""
    t = MEASURED_TRAFFIC.get((workload, storage, "fused" if info["fused"] else "split"))
    return None if t is None else round(t)


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
    dft = sim.dft()
    sim.close()
    return dict(wall_s=wall, event_ms=ev_ms, e_ms=prof["e_kernel_ms"], h_ms=prof["h_kernel_ms"],
                step_ms=prof["step_kernel_ms"], info=inf, probes=probes, dft=dft)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sc, seconds_target=15.0, single_seconds=8.0):
    """The CPU oracle (FP32 restatement of the reference's equations, OpenMP
    over z-planes) on the same workload, timed over its first steps: on the
    box's share of host cores (16 threads), then on one thread over a few
    further steps.  Returns (cpu_baseline, host)."""
    from oracle.oracle import Oracle, load
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = Oracle(sc, 32, nsteps=256)
    load().oracle_set_threads(threads)
    t0 = time.perf_counter()
    o.steps(1)
    one = time.perf_counter() - t0
    n = int(max(1, min(200, seconds_target / max(one, 1e-3))))
    t0 = time.perf_counter()
    o.steps(n)
    dt = time.perf_counter() - t0
    base = {"value": sc.cells * n / dt / 1e6, "unit": "Mcell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle/fdtd_oracle.cpp FP32, OpenMP {threads} threads, steps 2-{n + 1} of "
                      f"the full {sc.nx}x{sc.ny}x{sc.nz} {sc.name} grid ({dt:.1f} s)"}
    # one thread: the per-step rate over the next few steps
    load().oracle_set_threads(1)
    n1 = int(max(1, min(255 - 1 - n, single_seconds / max(one * threads, 1e-3))))
    t0 = time.perf_counter()
    o.steps(n1)
    dt1 = time.perf_counter() - t0
    load().oracle_set_threads(threads)
    o.close()
    host = {"cpu_model": _cpu_model(), "nproc": os.cpu_count(), "omp_threads": threads,
            "single_thread_value": round(sc.cells * n1 / dt1 / 1e6, 2), "single_thread_unit": "Mcell-updates/s",
            "single_thread_sample": f"1 thread, steps {n + 2}-{n + n1 + 1} of the same grid ({dt1:.1f} s)"}
    return base, host


def copy_ceiling_gbs(nbytes=1 << 30, iters=5):
    """GB/s of the plain HBM copy kernel (non-temporal 16 B per lane, read and
    written bytes counted), the practical ceiling the sweeps are set against."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "fdtd_amd", "libhbmcopy.so"))
    lib.hbm_copy_variant_gbs.restype = ctypes.c_double
    lib.hbm_copy_variant_gbs.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
    return round(lib.hbm_copy_variant_gbs(nbytes, iters, 5), 1)   # variant 5: u1_nt


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
    ap.add_argument("--copy-variants", action="store_true")
    args = ap.parse_args()
    if args.copy_probe or args.copy_variants:
        import ctypes
        lib = ctypes.CDLL(os.path.join(ROOT, "fdtd_amd", "libhbmcopy.so"))
        lib.hbm_copy_gbs.restype = ctypes.c_double
        lib.hbm_copy_gbs.argtypes = [ctypes.c_uint64, ctypes.c_int]
        gbs = lib.hbm_copy_gbs(args.copy_bytes, args.copy_iters)
        out = {"copy_gbs": round(gbs, 1), "bytes_each_way": args.copy_bytes,
               "iters": args.copy_iters, "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}
        if args.copy_variants:
            lib.hbm_copy_variant_gbs.restype = ctypes.c_double
            lib.hbm_copy_variant_gbs.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
            names = ["grid_stride", "u1", "u2", "u4", "u8", "u1_nt", "u2_nt", "u4_nt", "u8_nt"]
            out["variants_gbs"] = {nm: round(lib.hbm_copy_variant_gbs(args.copy_bytes, args.copy_iters, v), 1)
                                   for v, nm in enumerate(names)}
        print(json.dumps(out))
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
    bpc = b_alg(sc, f32["info"], 4)
    kms = f32["step_ms"]
    achieved = bpc * cells / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    step_gbs = bpc * cells / (ms_per_step * 1e-3) / 1e9
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
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": _traffic(sc.name, "f32", f32["info"]),
                     "kernel": "step_kernel" if f32["info"]["fused"] else "e_kernel+h_kernel",
                     "bytes_per_cell": round(bpc, 3), "kernel_ms": round(kms, 5),
                     "copy_ceiling": copy_ceiling_gbs()},
        "timed_loop_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "wall_s": round(f32["wall_s"], 3),
    }
    if not args.no_fp16:
        f16 = run_form(sc, "f16", args.steps, args.warmup, args.profile_steps)
        v16 = cells * args.steps / (f16["event_ms"] * 1e-3) / 1e6
        b16 = b_alg(sc, f16["info"], 2)
        p32 = f32["probes"].astype(np.float64)
        p16 = f16["probes"].astype(np.float64)
        # probes carrying signal (a probe the pulse has not reached is noise
        # over noise); the FP32 run is the reference: it is within 1e-5 of the
        # FP64 oracle on every scene the tests check
        n32 = np.linalg.norm(p32, axis=1)
        live = (n32 > 0) & (n32 >= 1e-3 * np.abs(p32).max(initial=0.0) * np.sqrt(max(p32.shape[1], 1)))
        rel = np.linalg.norm(p16 - p32, axis=1)[live] / n32[live]
        d32, d16 = f32["dft"], f16["dft"]
        dft_err = float(np.abs(d16 - d32).max() / np.abs(d32).max()) if d32.size and np.abs(d32).max() > 0 else None
        out["fp16"] = {"value": round(v16, 1), "ms_per_step": round(f16["event_ms"] / args.steps, 5),
                       "speedup_vs_fp32": round(v16 / value, 3),
                       "probe_rel_l2_vs_fp32_max": float(f"{rel.max():.3e}") if rel.size else None,
                       "probe_rel_l2_vs_fp32_median": float(f"{np.median(rel):.3e}") if rel.size else None,
                       "dft_rel_err_vs_fp32_max": float(f"{dft_err:.3e}") if dft_err is not None else None,
                       "probes_compared": int(live.sum()),
                       "kernel_ms": round(f16["step_ms"], 5), "bytes_per_cell": round(b16, 3),
                       "traffic": _traffic(sc.name, "f16", f16["info"]),
                       "frac": (round(b16 * cells / (f16["step_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                if f16["step_ms"] > 0 else None)}
    if not args.no_cpu:
        out["cpu_baseline"], out["cpu_host"] = cpu_baseline(sc)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
