"""Benchmark: Mcell-updates/s of the Yee/CPML time step on one MI355X.

Workload (default): slab512 -- 512x512x256 cells, 0.5 mm, CPML 16 on every
face, 256 seeded dielectric boxes, 8 lumped z-sources, 128 Ez probes with an
FP64 DFT at 16 frequencies -- one GPU's share of the reference's z-slab run.
A step is one full time step (E half-step then H half-step) over the whole
grid; inputs are resident in HBM before timing starts.

Prints one JSON line with the FP32 form as `value` and, beside it:
  * the throughput of the FP16-storage form and of the block-scaled int16
    form (int16 mantissas with one exponent byte per 8 x-cells, FP16 psi);
  * accuracy: every storage form run for the workload's full 3000 steps
    (whatever --steps is), its probe series and DFTs against the committed
    FP64 fixture tests/golden/slab512_fp64.npz;
  * the roofline of the sweep kernels (algorithmic bytes per step over their
    HIP-event launch time) with the HBM traffic of the running library measured live by
    rocprofv3 (FETCH_SIZE / WRITE_SIZE, calibrated on a copy kernel);
  * the adjoint128 workload: adjoint/forward wall-time ratios and gradient
    error against the committed FP64 fixture tests/golden/adjoint128.npz;
  * the CPU oracle timed on a bounded sample on the host as a reported
    baseline.
"""
import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fdtd_amd import api, scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0

GOLDEN = os.path.join(ROOT, "tests", "golden")
METRIC = "Mcell-updates/s and % HBM peak; FP16-store vs FP32 at stated probe error"


def b_alg(sc, info, storage_bytes, psi_bytes=None):
    """Algorithmic HBM bytes per cell-update of one time step.

    Each sweep reads and writes its own three field components and reads the
    other three (18 values per step); a classed E update reads one class
    index per cell; each CPML psi (4 per axis) is read and written over its
    slab (fraction 2L/N of the axis).  storage_bytes per field value
    (block-scaled int16: 2 + 4/24, a 4-byte word holding the three
    components' exponents per 8 cells), psi_bytes per psi value (default the
    same)."""
    s = storage_bytes
    sp = s if psi_bytes is None else psi_bytes
    f = [(sc.cpml.thick[0] + sc.cpml.thick[1]) / sc.nx, (sc.cpml.thick[2] + sc.cpml.thick[3]) / sc.ny,
         (sc.cpml.thick[4] + sc.cpml.thick[5]) / sc.nz]
    psi = 8.0 * sp * sum(f)
    ncls = info["n_classes"]
    cls_bytes = 0.0 if ncls <= 1 else (2.0 if ncls <= 65536 else 4.0)
    return 18 * s + cls_bytes + psi


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


def accuracy_vs_fp64(sc, storages=("f32", "f16", "bf16", "bq16")):
    """Every storage form run for the workload's own step count (slab512:
    3000), its probe series and DFTs against the committed FP64 oracle
    fixture of the same workload.  Probe error: relative L2 of each probe's
    series, over the probes carrying signal (row norm >= 1e-3 of the
    largest); DFT error: |X - X64| over the probe's spectral peak."""
    path = os.path.join(GOLDEN, sc.name + "_fp64.npz")
    if not os.path.exists(path):
        return {"fixture": None, "note": f"no FP64 fixture for {sc.name}"}
    fx = np.load(path)
    ref, X64 = fx["probes"], fx["dft"]
    steps = int(fx["steps"])
    rn = np.linalg.norm(ref, axis=1)
    sig = rn >= 1e-3 * rn.max()
    out = {"fixture": os.path.relpath(path, ROOT), "steps": steps, "probes_compared": int(sig.sum())}
    for st in storages:
        sim = api.Simulation.from_scene(sc, st, n_steps=steps)
        sim.run(steps)
        got = sim.probes()[:, :steps].astype(np.float64)
        X = sim.dft()
        sim.close()
        err = np.linalg.norm(got - ref, axis=1)[sig] / rn[sig]
        derr = np.abs(X - X64) / np.abs(X64).max(axis=1, keepdims=True)
        out[st] = {"probe_rel_l2_max": float(f"{err.max():.3e}"), "probe_rel_l2_median": float(f"{np.median(err):.3e}"),
                   "dft_rel_err_max": float(f"{derr.max():.3e}"),
                   "meets_1e-3": bool(err.max() <= 1e-3)}
    return out


def adjoint_leg(checkpoint_every=64):
    """adjoint128 (128x128x64, FP32, 1000 steps, design 32x32x16): adjoint
    wall time over a forward run for the checkpointed (every 64 steps) and
    stored-history paths, and the gradient against the committed FP64
    fixture: relative L2 over all design cells and the 16 seeded cells'
    FP64 central differences (h = 1e-3), as a fraction of max |g|."""
    sc = scenes.adjoint128()
    fx = np.load(os.path.join(GOLDEN, "adjoint128.npz"))
    terms = [(p, f, 1.0) for p in range(len(sc.probes)) for f in range(len(sc.freqs))]
    sim = api.Simulation.from_scene(sc, "f32")
    sim.run(sc.steps)
    sim.reset()
    sim.run(sc.steps)
    fwd = sim.info()["last_run_ms"]
    out = {"workload": sc.name, "steps": sc.steps, "forward_ms": round(fwd, 3)}
    grads = {}
    for K, key in ((checkpoint_every, f"checkpoint_every_{checkpoint_every}"), (0, "stored_history")):
        sim.reset()
        J, g = sim.adjoint_gradient(terms, sc.steps, K)
        inf = sim.info()
        grads[K] = (J, g)
        out[key] = {"adjoint_ms": round(inf["adjoint_total_ms"], 3),
                    "ratio_vs_forward": round(inf["adjoint_total_ms"] / fwd, 3),
                    "peak_device_gb": round(inf["adjoint_device_bytes"] / 1e9, 3)}
    sim.close()
    J, g = grads[checkpoint_every]
    g64 = fx["grad"]
    gmax = float(np.abs(g64).max())
    fd_err = np.abs(g[fx["fd_cells"]] - fx["fd"]) / gmax
    out["checkpointed_equals_stored"] = bool(np.array_equal(g, grads[0][1]))
    out["J_rel_vs_fp64"] = float(f"{abs(J - float(fx['J'])) / float(fx['J']):.3e}")
    out["grad_rel_l2_vs_fp64_adjoint"] = float(f"{np.linalg.norm(g - g64) / np.linalg.norm(g64):.3e}")
    out["fd_fp64_h1e-3_max_err_over_max_grad"] = float(f"{fd_err.max():.3e}")
    out["fd_cells"] = int(fx["fd_cells"].size)
    return out


# ---- live HBM traffic: rocprofv3 counters over a child run of this script --------

def _traffic_probe(workload, steps=3):
    """Child mode under rocprofv3: a plain copy kernel of known size (the
    counter calibration), then `steps` steps of every storage form."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "fdtd_amd", "libhbmcopy.so"))
    lib.hbm_copy_variant_gbs.restype = ctypes.c_double
    lib.hbm_copy_variant_gbs.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
    lib.hbm_copy_variant_gbs(COPY_BYTES, 2, 5)
    sc = scenes.WORKLOADS[workload]()
    for st in ("f32", "f16", "bq16"):
        sim = api.Simulation.from_scene(sc, st, n_steps=steps)
        sim.run(steps)
        sim.close()


COPY_BYTES = 1 << 30


def _kernel_form(name):
    """Storage form of a sweep kernel from its (demangled) name: the first
    template argument is float, the 16-bit float type (printed as _Float16 or
    half), bf16_t or bq16_t."""
    if "hbm_copy" in name:
        return "copy"
    if not any(k in name for k in ("e_kernel", "h_kernel", "e_pair_kernel", "h_pair_kernel")):
        return None
    if "bq16" in name:
        return "bq16"
    if "bf16" in name or "bfloat16" in name:
        return "bf16"
    if "<float" in name or "IfL" in name:
        return "f32"
    return "f16"


def _counter_rows(outdir):
    rows = []
    for path in glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True):
        with open(path, newline="") as f:
            rows.extend(csv.DictReader(f))
    return rows


def live_traffic(workload, steps=3, timeout=240):
    """HBM bytes per time step of each storage form, from two rocprofv3
    --pmc passes (FETCH_SIZE, then WRITE_SIZE: KiB per dispatch) over a child
    run of this script, each counter scaled by what it reads for the copy
    kernel (COPY_BYTES read and written per dispatch).  Returns
    ({form: bytes per step}, info) or (None, reason)."""
    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    sums = {}
    factors = {}
    ndisp = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="fdtd_traffic_")
        try:
            cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "t", "--",
                   sys.executable, os.path.abspath(__file__), "--traffic-probe", "--workload", workload,
                   "--traffic-steps", str(steps)]
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, cwd=ROOT,
                               env=dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp")))
            rows = _counter_rows(d)
            if not rows:
                return None, f"rocprofv3 {counter}: no counter rows (rc {r.returncode}): {r.stderr[-300:]}"
            per = {}
            copy = []
            for row in rows:
                if row.get("Counter_Name") != counter:
                    continue
                form = _kernel_form(row.get("Kernel_Name", ""))
                v = float(row.get("Counter_Value", 0.0))
                if form == "copy":
                    copy.append(v)
                elif form:
                    per[form] = per.get(form, 0.0) + v
                    if counter == "FETCH_SIZE":
                        ndisp[form] = ndisp.get(form, 0) + 1
            if not copy:
                return None, f"rocprofv3 {counter}: no copy-kernel rows"
            factors[counter] = COPY_BYTES / 1024.0 / (sum(copy) / len(copy))
            for form, v in per.items():
                sums[form] = sums.get(form, 0.0) + v * factors[counter] * 1024.0 / steps
        except subprocess.TimeoutExpired:
            return None, f"rocprofv3 {counter}: timed out after {timeout} s"
        finally:
            shutil.rmtree(d, ignore_errors=True)
    info = {"source": "rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE (separate passes) over `bench.py --traffic-probe`, "
                      f"{steps} steps per form, the running library",
            "fetch_factor": round(factors["FETCH_SIZE"], 4), "write_factor": round(factors["WRITE_SIZE"], 4),
            "sweep_dispatches": ndisp}
    return {k: round(v) for k, v in sums.items()}, info


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
    ap.add_argument("--no-accuracy", action="store_true", help="skip the full-length accuracy leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint128 leg")
    ap.add_argument("--no-traffic", action="store_true", help="skip the live rocprofv3 traffic passes")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--traffic-steps", type=int, default=3, help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only time the plain HBM copy kernel (measurement calibration)")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--copy-iters", type=int, default=20)
    ap.add_argument("--copy-variants", action="store_true")
    args = ap.parse_args()
    if args.traffic_probe:
        _traffic_probe(args.workload, args.traffic_steps)
        return
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
        print(f"bench.py runs one GPU: the engine has no multi-GPU path (asked for {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    sc = scenes.WORKLOADS[args.workload]()
    cells = sc.cells
    _progress(f"{sc.name}: FP32, {args.warmup} warmup + {args.steps} timed steps")
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
                     "traffic": None,
                     "kernel": "e_kernel+h_kernel (interior, x-edge, general launches of one step)",
                     "bytes_per_cell": round(bpc, 3), "kernel_ms": round(kms, 5),
                     "copy_ceiling": copy_ceiling_gbs()},
        "timed_loop_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "wall_s": round(f32["wall_s"], 3),
    }
    f16 = None
    if not args.no_fp16:
        _progress(f"{sc.name}: FP16 storage, {args.warmup} warmup + {args.steps} timed steps")
        f16 = run_form(sc, "f16", args.steps, args.warmup, args.profile_steps)
        v16 = cells * args.steps / (f16["event_ms"] * 1e-3) / 1e6
        b16 = b_alg(sc, f16["info"], 2)
        out["fp16"] = {"value": round(v16, 1), "ms_per_step": round(f16["event_ms"] / args.steps, 5),
                       "speedup_vs_fp32": round(v16 / value, 3),
                       "kernel_ms": round(f16["step_ms"], 5), "bytes_per_cell": round(b16, 3),
                       "traffic": None,
                       "frac": (round(b16 * cells / (f16["step_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                if f16["step_ms"] > 0 else None)}
    if not args.no_fp16:
        _progress(f"{sc.name}: block-scaled int16 storage, {args.warmup} warmup + {args.steps} timed steps")
        bq = run_form(sc, "bq16", args.steps, args.warmup, args.profile_steps)
        vbq = cells * args.steps / (bq["event_ms"] * 1e-3) / 1e6
        bbq = b_alg(sc, bq["info"], 2.0 + 4.0 / 24.0, 2)
        out["bq16"] = {"value": round(vbq, 1), "ms_per_step": round(bq["event_ms"] / args.steps, 5),
                       "speedup_vs_fp32": round(vbq / value, 3),
                       "kernel_ms": round(bq["step_ms"], 5), "bytes_per_cell": round(bbq, 3),
                       "frac": (round(bbq * cells / (bq["step_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                if bq["step_ms"] > 0 else None),
                       "storage": "E, H int16 mantissas, one power-of-two exponent byte per 8 x-cells and component "
                                  "(one 4-byte word per block holds x, y, z); psi FP16"}
    if not args.no_accuracy:
        _progress(f"{sc.name}: accuracy leg, every storage form over the workload's {sc.steps} steps")
        out["accuracy_vs_fp64"] = accuracy_vs_fp64(sc)
    if not args.no_adjoint:
        _progress("adjoint128: forward, checkpointed and stored-history adjoints")
        out["adjoint"] = adjoint_leg()
    if not args.no_traffic:
        _progress(f"{sc.name}: live HBM traffic (two rocprofv3 counter passes)")
        t, tinfo = live_traffic(sc.name)
        if t is None:
            out["roofline"]["traffic_note"] = tinfo
        else:
            out["roofline"]["traffic"] = t.get("f32")
            out["roofline"]["traffic_bytes_per_cell"] = round(t["f32"] / cells, 2) if "f32" in t else None
            out["roofline"]["traffic_source"] = tinfo
            for key, form in (("fp16", "f16"), ("bq16", "bq16")):
                if key in out and form in t:
                    out[key]["traffic"] = t[form]
                    out[key]["traffic_bytes_per_cell"] = round(t[form] / cells, 2)
    if not args.no_cpu:
        _progress(f"{sc.name}: CPU oracle sample on the host")
        out["cpu_baseline"], out["cpu_host"] = cpu_baseline(sc)
    print(json.dumps(out), flush=True)


def _progress(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
