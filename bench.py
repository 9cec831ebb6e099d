#!/usr/bin/env python3
"""Throughput of the plain FP32 solver on one MI355X.

Workload (default): hills_400x400x200 -- the curvilinear case: 400x400x200
cells, h = 0.5 km, a terrain-following grid under 20 seeded Gaussian hills
(0-2 km high, sigma 5-20 km, tapered to zero over the CFS-PML shell), Vs
gradient 3.2 -> 4.6 km/s with a correlated +-5% perturbation, Vp = 1.73 Vs,
Gardner density, 12-point CFS-PML on 5 sides, Ricker 1 Hz surface force
(mode C: the 3-D metric).  --workload flat runs the same model on the
spherical grid without hills (mode S, the reference's scheme); the default
run also times mode S and reports it under "flat_spherical".
A step is one full time step (4 RK stages) over all cells; inputs are
resident in HBM before the timed region starts.

Prints ONE JSON line:
  value      cell-updates/s (interior cells x steps / wall time of K steps)
  roofline   one RK stage (k_stage over all cells, the dominant work):
             algorithmic bytes per stage over its mean duration measured
             with HIP events on the solver's stream, against the 8 TB/s HBM
             peak.  Algorithmic bytes per cell and stage: mode S 156 B (the
             average of 120/192/192/120 B over the 4 RK stages of the
             3-register RK4 form: W, media, m, t); mode C 196 B, the
             per-unit figure of the measurement contract (156 B plus a
             10-value curvilinear metric).  The terrain-following scheme
             reads 4 metric values per cell and stage (slopes p, q,
             zeta_r, 1/r): 172 B, reported as frac_scheme_bytes.
             traffic = HBM bytes per stage from rocprofv3 FETCH_SIZE /
             WRITE_SIZE runs, scaled by the same counters on a plain copy
             (summarised by --summarize-pmc into
             profiles/stage_traffic_<workload>.json)
  cpu_baseline  the FP32 C oracle (OpenMP) on every host core this
             process may run on, timed on a bounded sample of the same
             workload (2 full steps); nproc and the CPU model reported
  accuracy   receiver records of the two-layer 100x100x50 workload (1008
             steps) against the committed FP32 / FP64 CPU records
  pipeline   station-pair kernels/hour, plain (one FP32 solve per station,
             SGT through host memory) vs batched + FP16 SGT (one batched
             solve with step graphs and the autotuner, FP16 SGT in HBM), on the
             batched configuration at a tenth of its record lengths: 256 x
             256 x 128, 32 virtual sources of 150-400 steps, SGT every 10
             steps on the 2x decimated grid, 16 pairs; with the kernels'
             relative L2 difference
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic FP32 bytes per cell-update (4 RK stages): mode S 156 values;
# mode C 156 + 4 x 10 metric values (the contract's general curvilinear
# metric); the terrain-following scheme's own minimum 156 + 4 x 4
BYTES_PER_CELL_UPDATE = {"flat": 624.0, "hills": 784.0}
SCHEME_BYTES_PER_CELL_UPDATE = {"flat": 624.0, "hills": 688.0}
WORKLOAD_NAME = {"flat": "gradient_400x400x200", "hills": "hills_400x400x200"}


def traffic_json(workload):
    return os.path.join(ROOT, "profiles", f"stage_traffic_{WORKLOAD_NAME[workload]}.json")
COPY_BYTES = 1 << 30                    # --copy-calibration moves 1 GiB each way per copy


def _kernel_family(name):
    import re
    m = re.search(r"\b(k_\w+)", name)
    return m.group(1) if m else ("copy" if "copyBuffer" in name else name[:40])


def _pmc_means(db):
    """Mean counter value (KiB) per dispatch, by kernel family, from a rocprofv3 results database."""
    import sqlite3
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, dispatch_id, sum(value) from counters_collection "
                       "group by dispatch_id").fetchall()
    acc = {}
    for name, _, v in rows:
        acc.setdefault(_kernel_family(name), []).append(v)
    return {k: sum(v) / len(v) for k, v in acc.items()}, {k: len(v) for k, v in acc.items()}


def summarize_pmc(fetch_db, write_db, copy_fetch_db, copy_write_db, out):
    """HBM bytes per RK stage (k_stage) from FETCH_SIZE / WRITE_SIZE runs.
    Each counter is scaled by bytes-moved / counter-value of a plain 1 GiB device copy."""
    f, nf = _pmc_means(fetch_db)
    w, nw = _pmc_means(write_db)
    cf, _ = _pmc_means(copy_fetch_db)
    cw, _ = _pmc_means(copy_write_db)
    ff = COPY_BYTES / (cf["copy"] * 1024.0)
    fw = COPY_BYTES / (cw["copy"] * 1024.0)
    kern = {}
    for k in ("k_stage",):
        kern[k] = {"fetch_bytes": f[k] * 1024.0 * ff, "write_bytes": w[k] * 1024.0 * fw,
                   "dispatches": nf[k]}
    total = sum(v["fetch_bytes"] + v["write_bytes"] for v in kern.values())
    res = {"bytes_per_launch": total, "launch": "one RK stage = k_stage over all cells",
           "fetch_factor": ff, "write_factor": fw, "kernels": kern}
    with open(out, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps(res))


def summarize_trace(db, ncell, bytes_per_launch_cell):
    """Markdown table of kernel durations by family (and by RK form for the stage kernel)."""
    import re
    import sqlite3
    con = sqlite3.connect(db)
    rows = con.execute("select name, duration from kernels").fetchall()
    acc = {}
    for name, d in rows:
        key = _kernel_family(name)
        m = re.search(r"(k_stage)<([^>]*)>", name)
        if m:
            key += f" form {m.group(2).split(',')[3].strip()}"
        acc.setdefault(key, []).append(d / 1000.0)
    tot = sum(sum(v) for v in acc.values())
    print("| kernel | calls | total us | mean us | share |\n|---|---|---|---|---|")
    for k, v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        print(f"| {k} | {len(v)} | {sum(v):.0f} | {sum(v) / len(v):.1f} | {100 * sum(v) / tot:.1f}% |")
    st = [sum(acc.get(k, [])) for k in acc if k.startswith("k_stage")]
    nst = len([1 for k, v in acc.items() if k.startswith("k_stage") for _ in v])
    if nst:
        mean = sum(st) / nst
        print(f"\nmean stage (k_stage) {mean:.1f} us -> "
              f"{ncell * bytes_per_launch_cell / (mean * 1e-6) / 1e9:.0f} GB/s algorithmic")


def ae_sample(iters=100, batch=16, nblocks=2048):
    """The SGT autoencoder's kernels alone: Adam training on random batches
    of scaled 6 x 16^3 blocks (synthetic, uniform in [-1, 1]) and inference
    of nblocks blocks; wall times and the convolutions' FLOPs."""
    import numpy as np
    from fwant_amd import ae
    from fwant_amd._lib import DeviceArray, device_sync
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (nblocks, 6, 16, 16, 16)).astype(np.float32)
    d = DeviceArray.from_host(x)
    net = ae.AE(seed=5)
    ae.train(net, d, nblocks, iters=2, batch=batch)
    device_sync()
    t = time.perf_counter()
    hist = ae.train(net, d, nblocks, iters=iters, batch=batch, lr=2e-3)
    device_sync()
    t_train = time.perf_counter() - t
    t = time.perf_counter()
    ae.reconstruct(net, d, nblocks)
    device_sync()
    t_inf = time.perf_counter() - t
    # 2 flops per multiply-add of a layer: conv (Co Do^3) x (Ci K^3); a
    # transposed layer the same with input and output swapped
    fwd = sum(2.0 * (co * do ** 3 * ci * k ** 3 if kind == "conv" else ci * di ** 3 * co * k ** 3)
              for kind, ci, co, di, do, k, s, p_, a in ae.LAYERS)
    return {"workload": f"autoencoder {iters} training iterations of batch {batch} (synthetic blocks), "
                        f"inference of {nblocks} blocks", "train_s": t_train, "ms_per_iteration": 1e3 * t_train / iters,
            "inference_s": t_inf, "final_loss": hist[-1], "forward_gflop_per_block": fwd / 1e9,
            "train_gflop_per_iteration": 3 * fwd * batch / 1e9}


def scatter_sample(knt=400, npt=262144, reps=5):
    """The station-pair convolution alone (fwk_scatter_param, alpha/beta):
    6-component strain series of knt samples at npt points (seeded normal
    values), the O(knt^2) sums per point; wall time per call and (n, m)
    pairs per second."""
    import numpy as np
    from fwant_amd import kernel
    from fwant_amd._lib import DeviceArray, device_sync
    rng = np.random.default_rng(0)
    R = DeviceArray.from_host(rng.standard_normal((6, knt, npt)).astype(np.float32))
    S = DeviceArray.from_host(rng.standard_normal((6, knt, npt)).astype(np.float32))
    lam = DeviceArray.from_host(np.full(npt, 3e10, np.float32))
    mu = DeviceArray.from_host(np.full(npt, 3e10, np.float32))
    A, B = kernel.sgt2scatter(R, S, lam, mu, 0.1)
    device_sync()
    A.free()
    B.free()
    t = time.perf_counter()
    for _ in range(reps):
        A, B = kernel.sgt2scatter(R, S, lam, mu, 0.1)
        device_sync()
        A.free()
        B.free()
    t = (time.perf_counter() - t) / reps
    return {"workload": f"fwk_scatter_param: {npt} points, knt {knt}, 6 components", "ms_per_call": 1e3 * t,
            "pairs_per_s": npt * knt * (knt - 1) / 2 / t}


def stream_probe(gib=2, reps=10):
    """Attainable HBM bandwidth on this GPU (fw_hbm_probe): grid-stride
    streaming of two 2 GiB buffers with 16-byte non-temporal accesses,
    read-only and copy (read + write 1:1), over a range of grid sizes."""
    import ctypes as C
    from fwant_amd._lib import lib, kcheck
    out = {"bytes": gib << 30, "reps": reps, "runs": []}
    for blocks in (1024, 2048, 4096, 8192, 16384):
        r, c = C.c_float(), C.c_float()
        kcheck(lib().fw_hbm_probe(gib << 30, blocks, reps, C.byref(r), C.byref(c)), "fw_hbm_probe")
        out["runs"].append({"blocks": blocks, "read_gbs": r.value, "copy_gbs": c.value})
    out["read_gbs_best"] = max(x["read_gbs"] for x in out["runs"])
    out["copy_gbs_best"] = max(x["copy_gbs"] for x in out["runs"])
    return out


def cpu_model():
    """CPU model string of the host (first 'model name' of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pipeline_sample(scale=10):
    """Station-pair kernels/hour on the batched-sources configuration at
    1/scale of its record lengths (default a tenth): 256x256x128 curvilinear
    model (the gradient model under 20 seeded hills, mode C), 32 seeded
    surface stations with record lengths uniform in 150-400 steps (the
    configuration's 1500-4000 over 10), 6-component stress SGT every 10
    steps on the 2x decimated grid (128x128x64), 16 seeded station pairs.
    plain: one FP32 solve per station whose SGT goes to host memory, each
    pair's two SGTs back to the GPU for the integral; batched: all 32 in one
    autotuned, graph-replayed batched solve with FP16 SGT (per-block
    scales) kept in HBM."""
    import numpy as np
    from scipy import signal
    from fwant_amd import models, pipeline
    m = models.hills_256x256x128(nt=4000 // scale)
    rng = np.random.default_rng(11)
    w = 256 - 48
    cells = rng.choice(w * w, size=32, replace=False)
    st = [(int(3 + 24 + v % w), int(3 + 24 + v // w)) for v in cells]
    nsteps = [int(v) * 10 // scale for v in rng.integers(150, 401, size=32)]
    pairs = []
    while len(pairs) < 16:
        a, b = (int(v) for v in rng.choice(32, size=2, replace=False))
        if (a, b) not in pairs and (b, a) not in pairs:
            pairs.append((a, b))
    tinv = 10
    knt = max(nsteps) // tinv
    t = np.arange(1, knt + 1) * m.stept * tinv
    V = np.sin(2 * np.pi * 0.5 * t).astype(np.float32)
    U = np.cos(2 * np.pi * 0.5 * t).astype(np.float32)
    b, a = signal.butter(2, 0.4)
    kw = dict(tinv=tinv, step=2, V=V, U=U, win=(2, 14), M0=1e16, filt=("P2", b, a))
    kp, tp = pipeline.plain(m, st, nsteps, pairs, host_sgt=True, **kw)
    kb, tb = pipeline.batched(m, st, nsteps, pairs, fmt="fp16", **kw)
    errs = [float(np.linalg.norm(y.astype(np.float64) - x) / np.linalg.norm(x))
            for P, B in zip(kp, kb) for x, y in zip(P, B) if np.linalg.norm(x) > 0]
    tot_p = sum(v for k, v in tp.items())
    tot_b = tb["setup"] + tb["steps"] + tb["kernel"]
    plain_h, batch_h = len(pairs) / tot_p * 3600.0, len(pairs) / tot_b * 3600.0
    log(f"pipeline: plain {tp}, batched fp16 {tb}, kernel rel-L2 max {max(errs):.2e} over {len(errs)} kernels")
    return {"workload": f"hills_256x256x128 curvilinear model (mode C), 32 virtual sources (record lengths "
                        f"{min(nsteps)}-{max(nsteps)} steps: the batched configuration's 1500-4000 over {scale}), "
                        "6-component stress SGT every 10 steps on the 2x decimated grid (128x128x64), 16 station pairs",
            "unit": "kernels/hour",
            "plain": dict(value=plain_h, sgt="FP32 via host memory", **tp),
            "batched_fp16": dict(value=batch_h, sgt="FP16 codes with per-16^3-block FP32 scales in HBM", **tb),
            "speedup": batch_h / plain_h,
            "speedup_steps_only": tp["steps"] / tb["steps"],
            "kernel_rel_l2_fp16_vs_fp32_gpu": max(errs), "kernels_compared": len(errs),
            "kernel_error_vs_fp64_oracle": "tests/test_gpu_configs.py (16 pairs: FP16 2.1e-3, BF16 1.6e-2)"}


GOLDEN = os.path.join(ROOT, "tests", "golden", "two_layer_100x100x50.npz")


def accuracy_sample():
    """The GPU's receiver records on the two-layer 100x100x50 workload (the
    full 1008 steps) against the committed FP32 and FP64 CPU records of the
    same inputs (tests/golden); the largest relative L2 over the six
    components that do not vanish at the traction-free surface."""
    import numpy as np
    from fwant_amd import models
    from fwant_amd.solver import Solver
    d = np.load(GOLDEN)
    m = models.two_layer_100x100x50()
    s = Solver(m)
    s.run(int(d["steps"]))
    g = s.receivers().astype(np.float64)
    s.close()
    r32, r64 = d["rec32"].astype(np.float64), d["rec64"].astype(np.float64)
    e32 = e64 = o64 = 0.0
    for f in (0, 1, 2, 3, 4, 6):  # Vx Vy Vz Txx Tyy Txy
        e32 = max(e32, np.linalg.norm(g[..., f] - r32[..., f]) / np.linalg.norm(r32[..., f]))
        e64 = max(e64, np.linalg.norm(g[..., f] - r64[..., f]) / np.linalg.norm(r64[..., f]))
        o64 = max(o64, np.linalg.norm(r32[..., f] - r64[..., f]) / np.linalg.norm(r64[..., f]))
    log(f"accuracy: vs FP32 CPU {e32:.2e}, vs FP64 CPU {e64:.2e} (FP32 CPU vs FP64 {o64:.2e})")
    return {"workload": "two_layer_100x100x50, 1008 steps, 16 surface receivers",
            "rel_l2_vs_fp32_cpu": e32, "rel_l2_vs_fp64_cpu": e64, "fp32_cpu_vs_fp64_cpu": o64}


def time_solver(m, steps, warmup, bytes_per_cell_update):
    """(cell-updates/s, mean stage us, achieved GB/s, stage share of the
    step) of the plain solver on model m: warmup untimed steps, then `steps`
    timed between device synchronisations, then 32 steps with HIP events
    around every stage launch on the solver's stream."""
    from fwant_amd.solver import Solver
    from fwant_amd._lib import device_sync
    s = Solver(m, max_records=steps + warmup + 64)
    ncell = m.ni * m.nj * m.nk
    s.run(warmup, ntime0=0)
    s.sync()
    device_sync()
    t0 = time.perf_counter()
    s.run(steps, ntime0=warmup)
    s.sync()
    device_sync()
    el = time.perf_counter() - t0
    value = ncell * steps / el
    s.timing(True)
    s.run(min(steps, 32), ntime0=warmup + steps)
    stage_ms, total_ms, nst = s.get_timing()
    s.timing(False)
    s.close()
    avg_ms = stage_ms / max(nst, 1)
    achieved = ncell * bytes_per_cell_update / 4 / (avg_ms * 1e-3) / 1e9
    return value, el, avg_ms * 1e3, achieved, stage_ms / total_ms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=16)
    ap.add_argument("--workload", choices=("hills", "flat"), default="hills",
                    help="hills: the curvilinear 400x400x200 case (mode C); flat: the same without hills (mode S)")
    ap.add_argument("--ni", type=int, default=400)
    ap.add_argument("--nj", type=int, default=400)
    ap.add_argument("--nk", type=int, default=200)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-pipeline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-flat", action="store_true", help="do not also time the flat (mode S) variant")
    ap.add_argument("--traffic", type=str, default=None,
                    help="JSON file with measured HBM bytes per RK stage (written by --summarize-pmc)")
    ap.add_argument("--summarize-pmc", nargs=5, metavar=("FETCH_DB", "WRITE_DB", "COPY_FETCH_DB", "COPY_WRITE_DB", "OUT"),
                    help="write the traffic JSON from rocprofv3 --pmc results databases and exit")
    ap.add_argument("--summarize-trace", metavar="DB",
                    help="print a kernel-duration table from a rocprofv3 --kernel-trace database and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (4 times) to calibrate the HBM counters")
    ap.add_argument("--pipeline-only", action="store_true",
                    help="run only the kernels/hour pipeline sample and print its JSON (for profiling it alone)")
    ap.add_argument("--pipeline-scale", type=int, default=10,
                    help="record lengths of the pipeline sample: the configured 1500-4000 steps over this")
    ap.add_argument("--stream-probe", action="store_true",
                    help="measure the attainable HBM bandwidth (read-only sum, copy) and print its JSON")
    ap.add_argument("--scatter-sample", action="store_true",
                    help="run only the station-pair convolution sample and print its JSON")
    ap.add_argument("--ae-sample", action="store_true",
                    help="run only the SGT autoencoder's training and inference sample and print its JSON")
    args = ap.parse_args()
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc)
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace, args.ni * args.nj * args.nk, BYTES_PER_CELL_UPDATE[args.workload] / 4)
        return
    if args.copy_calibration:
        import numpy as np
        from fwant_amd._lib import DeviceArray, device_sync
        a = DeviceArray.from_host(np.ones(1 << 28, np.float32))
        for _ in range(4):
            b = a.copy()
            b.free()
        device_sync()
        log("copy calibration: 4 device copies of 1 GiB (read 1 GiB + write 1 GiB each)")
        return
    if args.pipeline_only:
        print(json.dumps(pipeline_sample(args.pipeline_scale)), flush=True)
        return
    if args.stream_probe:
        print(json.dumps(stream_probe()), flush=True)
        return
    if args.scatter_sample:
        print(json.dumps(scatter_sample()), flush=True)
        return
    if args.ae_sample:
        print(json.dumps(ae_sample()), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(virtual sources shard across GPUs with no exchange; see the sharding design)", file=sys.stderr)
        sys.exit(2)

    from fwant_amd import models

    wl = args.workload
    bpc = BYTES_PER_CELL_UPDATE[wl]
    t = time.time()
    kw = dict(ni=args.ni, nj=args.nj, nk=args.nk)
    m = models.hills_400x400x200(**kw) if wl == "hills" else models.gradient_400x400x200(**kw)
    log(f"model {m.name} {m.ni}x{m.nj}x{m.nk} built in {time.time() - t:.1f}s")
    ncell = m.ni * m.nj * m.nk
    value, el, mean_us, achieved, share = time_solver(m, args.steps, args.warmup, bpc)
    log(f"timed {args.steps} steps in {el:.3f}s -> {value / 1e9:.3f} G cell-updates/s")
    log(f"RK stage (k_stage): mean {mean_us:.1f} us, {achieved:.0f} GB/s algorithmic ({bpc / 4:.0f} B per cell); "
        f"stage kernels {share * 100:.1f}% of step time")
    traffic = None
    tpath = args.traffic or traffic_json(wl)
    if os.path.exists(tpath) and (args.ni, args.nj, args.nk) == (400, 400, 200):
        with open(tpath) as f:
            tj = json.load(f)
        # only a summary of the current stage kernel counts
        if "k_stage" in tj.get("kernels", {}):
            traffic = tj.get("bytes_per_launch")
    flat = None
    if wl == "hills" and not args.no_flat:
        mf = models.gradient_400x400x200(**kw)
        fv, _, fus, fach, _ = time_solver(mf, args.steps, args.warmup, BYTES_PER_CELL_UPDATE["flat"])
        log(f"flat variant (mode S): {fv / 1e9:.3f} G cell-updates/s, stage {fus:.1f} us, {fach:.0f} GB/s")
        flat = {"workload": f"{mf.name}: the same model on the spherical grid without hills (mode S, the "
                            "reference's scheme)",
                "value": fv, "unit": "cell-updates/s", "mean_stage_us": fus,
                "roofline_frac": fach / HBM_PEAK_GBS, "bytes_per_cell_update": BYTES_PER_CELL_UPDATE["flat"]}
        del mf

    cpu = None
    if not args.no_cpu_baseline:
        from oracle.pyoracle import OracleSolver
        cores = len(os.sched_getaffinity(0))
        o = OracleSolver(m, 32, nthreads=cores)
        o.run(1)  # first-touch warm-up outside the timed sample
        t0 = time.perf_counter()
        o.run(args.cpu_steps)
        cel = time.perf_counter() - t0
        cpu = {"value": ncell * args.cpu_steps / cel, "unit": "cell-updates/s", "cores": cores,
               "kind": "port", "nproc": os.cpu_count(), "cpu_model": cpu_model(),
               "sample": f"FP32 C oracle, {args.cpu_steps} full steps of {m.ni}x{m.nj}x{m.nk} "
                         f"after 1 warm-up step ({cel:.1f}s)"}
        log(f"cpu baseline: {cpu['value'] / 1e6:.2f} M cell-updates/s on {cores} threads")
        o.close()

    pipe = None if args.no_pipeline else pipeline_sample()
    acc = None if args.no_accuracy else accuracy_sample()

    out = {
        "metric": "cell-updates/s + % HBM peak; batched+compressed vs plain at ≤1% kernel err",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"{m.name}: plain FP32 solver, {m.ni}x{m.nj}x{m.nk} cells, h=0.5 km, "
                                + ("terrain-following grid under 20 seeded Gaussian hills (0-2 km, sigma 5-20 km; "
                                   "mode C, 3-D metric)" if wl == "hills" else
                                   "flat free surface (spherical reference scheme, mode S)")
                                + ", 12-pt CFS-PML on 5 sides, Ricker 1 Hz force"),
                   "grid": [m.ni, m.nj, m.nk], "dt": m.stept, "mode": "C" if wl == "hills" else "S",
                   "bytes_per_cell_update": bpc, "scheme_bytes_per_cell_update": SCHEME_BYTES_PER_CELL_UPDATE[wl],
                   "hbm_fraction_whole_step": value * bpc / (HBM_PEAK_GBS * 1e9)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_cell_stage": bpc / 4,
                     "frac_scheme_bytes": achieved * SCHEME_BYTES_PER_CELL_UPDATE[wl] / bpc / HBM_PEAK_GBS,
                     "kernel": "RK stage = k_stage (operator sweep, CFS-PML and RK update; velocity and "
                               "stress halves as two wave groups of one block)",
                     "mean_launch_us": mean_us},
        "cpu_baseline": cpu,
        "flat_spherical": flat,
        "pipeline": pipe,
        "accuracy": acc,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
