"""Benchmark: migrated shots/s of the 2-D PSPI workload on one MI355X.

Workload (config 2): nx=1024, dx=10 m; nz=500, dz=5 m; nt=2000, dt=2 ms; 0-80 Hz (641 bins);
4 reference velocities; 1800 m/s + 12 Gaussian blobs; 64 shots 160 m apart; 25 Hz Ricker.  The
records are produced by the library's own demigration of the random reflectivity (untimed
setup); one step migrates and stacks all 64 shots from records already resident in HBM.

Prints one JSON line.  Timing: W warm-up steps, then K timed steps between device
synchronisations; the dominant kernel (the depth march) is timed with HIP events recorded on
the stream it is launched on.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "seismic.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def cpu_baseline(g, vel, wav, sx, m_gpu_fn, cores):
    """Oracle (C port of wem.c, naive FP32 = the reference arithmetic) on a bounded sample:
    one shot, a band of 2*cores frequency bins at full nx/nz.  Also returns the FP64-oracle
    rel-L2 of the GPU image on the same sample."""
    import oracle as O
    from seiswem import WemGeometry

    bins_per_hz = g.dt * (g.padt * g.nt)  # ifmax - ifmin per Hz
    nsamp = 10 * cores  # about 10 s of CPU work
    fmin = 40.0
    fmax = fmin + nsamp / bins_per_hz
    gs = WemGeometry(**{**g.__dict__, "fmin": fmin, "fmax": fmax})
    s_full, s_samp = g.sampling(), gs.sampling()
    nf_full = s_full["ifmax"] - s_full["ifmin"]
    nf_samp = s_samp["ifmax"] - s_samp["ifmin"]
    p = O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, fmin=fmin, fmax=fmax, nref=g.nref,
                    padt=g.padt, padx=g.padx, pspi=g.pspi, pspi_mode=0)
    rng = np.random.default_rng(4)
    d = rng.standard_normal((g.nm, g.nt)).astype(np.float32)
    t0 = time.perf_counter()
    O.wem(p, vel, wav, float(sx), d=d, adj=True, f64=False)
    t_cpu = time.perf_counter() - t0
    shots_per_s = (nf_samp / nf_full) / t_cpu
    m64 = O.wem(p, vel, wav, float(sx), d=d, adj=True, f64=True)
    m_gpu = m_gpu_fn(gs, d)
    err = float(np.linalg.norm(m_gpu - m64) / np.linalg.norm(m64))
    sample = (f"1 shot, {nf_samp} of {nf_full} frequency bins ({fmin:g}-{fmax:g} Hz) at full nx={g.nmx}, nz={g.nz}, "
              f"{t_cpu:.2f} s, scaled by bins")
    return {"value": shots_per_s, "unit": "shots/s", "cores": cores, "kind": "port", "sample": sample}, err, nf_samp


COPY_BYTES = 1 << 30
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "march2d_adj_pspi_nkx2048_64shots_traffic.json")


def traffic_per_launch(nshot):
    """HBM bytes per march launch from the committed FETCH_SIZE / WRITE_SIZE passes (KiB per
    dispatch), each scaled by the same counter's reading for a 1 GiB device copy; None when no
    pass of this workload is on file."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("nshot") != nshot:
        return None
    fetch = t["march_fetch_kib"] * (COPY_BYTES / 1024) / t["copy_fetch_kib"]
    write = t["march_write_kib"] * (COPY_BYTES / 1024) / t["copy_write_kib"]
    return (fetch + write) * 1024


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--freqs-per-block", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed steps, copy a 1 GiB device buffer once (the reference kernel that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not run in this build (shots shard across GPUs "
              "with one image all-reduce, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import seiswem
    from seiswem import synth

    g, vel, refl, wav, sx = synth.config2(nshot=args.shots)
    nshot = int(sx.size)
    plan = seiswem.Plan(g, vel, wav, freqs_per_block=args.freqs_per_block)
    m_true = seiswem.DeviceArray(refl.shape, refl)
    t0 = time.perf_counter()
    d = plan.model(m_true, sx)  # records by demigration (setup, untimed)
    seiswem.synchronize()
    t_model = time.perf_counter() - t0
    m = seiswem.DeviceArray((g.nm, g.nz))
    for _ in range(args.warmup):
        plan.migrate(d, sx, m=m)
    seiswem.synchronize()
    march_ms, launches, flops = 0.0, 0, 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.migrate(d, sx, m=m)
        tm = plan.timing()  # waits for this step's events
        march_ms += tm["march_ms"]
        launches += tm["march_launches"]
        flops += tm["march_flops"]
    seiswem.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = nshot * args.steps / elapsed
    avg_launch_s = march_ms / launches / 1e3
    flops_per_launch = flops / launches
    achieved = flops_per_launch / avg_launch_s / 1e12
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic_per_launch(nshot),
                "kernel": "march2d_adj (PSPI, nkx=2048)", "avg_launch_ms": march_ms / launches,
                "flops_per_launch": flops_per_launch}
    out = {
        "metric": "migrated shots/sec (2D & 3D) at <=1e-4 rel-L2 vs FP64; p50/p99 request latency",
        "value": value, "unit": "shots/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (config-2 generator; records by the library's own demigration)",
        "config": {"workload": "2-D PSPI shot-profile migration, 64 shots, nx=1024 nz=500 nt=2000 nref=4 0-80 Hz",
                   "nshot": nshot, "nkx": 2048, "nf": g.sampling()["ifmax"] - g.sampling()["ifmin"], "nz": g.nz,
                   "pspi_mode": "compat", "precision": "fp32 propagation, fp64 kz, compensated image sums"},
        "roofline": roofline,
        "setup_model_s": t_model,
    }
    if not args.no_cpu:
        cores = min(16, os.cpu_count() or 1)
        os.environ.setdefault("OMP_NUM_THREADS", str(cores))

        def m_gpu_fn(gs, dtr):
            return seiswem.ShotProfileWEM(dtr[None], True, vel=vel, wav=wav, geom=gs, sx=sx[:1])

        cb, err, nf_samp = cpu_baseline(g, vel, wav, float(sx[0]), m_gpu_fn, cores)
        out["cpu_baseline"] = cb
        out["rel_l2_vs_fp64"] = err
        out["rel_l2_sample"] = f"shot 0, {nf_samp} frequency bins, full nx/nz, PSPI compat"
    if args.copy_calibration:
        src = seiswem.DeviceArray((COPY_BYTES // 4,))
        dst = seiswem.DeviceArray((COPY_BYTES // 4,))
        src.zero()
        seiswem.synchronize()
        seiswem.copy_device(dst, src)  # the runtime's copy kernel, timed by the profiler
        seiswem.synchronize()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
