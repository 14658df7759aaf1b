"""Benchmark: migrated shots/s on one MI355X.

Default workload (--workload pspi2d) is the headline one; --workload ssf3d, serving, lsm, angles,
elastic and elastic3d measure the other workloads (one JSON line each).

Workload (config 2): nx=1024, dx=10 m; nz=500, dz=5 m; nt=2000, dt=2 ms; 0-80 Hz (641 bins);
4 reference velocities; 1800 m/s + 12 Gaussian blobs; 64 shots 160 m apart; 25 Hz Ricker.  The
records are produced by the library's own demigration of the random reflectivity (untimed
setup); one step migrates and stacks all 64 shots from records already resident in HBM.

Prints one JSON line.  Timing: W warm-up steps, then K timed steps between device
synchronisations; the dominant kernel (the depth march) is timed with HIP events recorded on
the stream it is launched on.
"""
import argparse
import dataclasses
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


def host_cores():
    """Threads the CPU baseline runs on: OMP_NUM_THREADS when the environment sets the process's
    CPU share, else every CPU this process may run on (sched_getaffinity, i.e. nproc)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(g, vel, wav, sx, m_gpu_fn, cores):
    """Oracle (C port of wem.c, naive FP32 = the reference arithmetic) on a bounded sample:
    one shot, a band of 2*cores frequency bins at full nx/nz.  Also returns the FP64-oracle
    rel-L2 of the GPU image on the same sample."""
    import oracle as O
    from seiswem import WemGeometry

    bins_per_hz = g.dt * (g.padt * g.nt)  # ifmax - ifmin per Hz
    nsamp = 20 * cores  # about 15 s of CPU work
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
    return {"value": shots_per_s, "unit": "shots/s", "cores": cores, "kind": "port", "sample": sample,
            "host_cpus": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "threads_from": "OMP_NUM_THREADS" if os.environ.get("OMP_NUM_THREADS", "").isdigit() else "nproc"}, \
        err, nf_samp


COPY_BYTES = 1 << 30
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "march2d_adj_pspi_nkx2048_64shots_traffic.json")


def traffic_per_launch(nshot, kernel):
    """HBM bytes per march launch from the committed FETCH_SIZE / WRITE_SIZE passes (KiB per
    dispatch), each scaled by the same counter's reading for a 1 GiB device copy; None when no
    pass of this workload is on file."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("nshot") != nshot or kernel not in t.get("kernel", ""):
        return None
    fetch = t["march_fetch_kib"] * (COPY_BYTES / 1024) / t["copy_fetch_kib"]
    write = t["march_write_kib"] * (COPY_BYTES / 1024) / t["copy_write_kib"]
    return (fetch + write) * 1024


def measure_ssf3d(args, steps, warmup, grid=256, nz=0, dot=True, graphs=True):
    """3-D split-step migration of one shot (nx=ny=256, dx=dy=12.5 m; nz=300, dz=10 m; nt=1500,
    dt=4 ms; 3-50 Hz = 565 bins; nkx=nky=512).  The passes are bound by their vector-ALU issue
    (the x-pass keeps the VALU 88 % busy, its state stays in the last-level cache between the
    passes), so the roofline is the FP32 one on the algorithmic flops per (w, iz), 2 F + 4 nm with
    F = 10 nk log2 nk + 6 nk + 8 nm; the HBM figure on 64 nk algorithmic bytes per (w, iz) (two
    wavefields x two passes x read+write of the padded slice) is reported beside it.  Returns the
    measured figures as a dict."""
    import seiswem
    from seiswem import synth

    g, vel, refl, wav, sx, sy = synth.config3(nx=grid, ny=grid)
    if nz:  # fewer depths of the same grid (per-dispatch profiling; every depth is the same work)
        g = dataclasses.replace(g, nz=nz)
        vel, refl = np.ascontiguousarray(vel[:, :nz]), np.ascontiguousarray(refl[:, :nz])
    plan = seiswem.Plan(g, vel, wav)
    plan.set_graphs(graphs)
    mm = seiswem.DeviceArray(refl.shape, refl)
    d = plan.model(mm, sx, sy)
    seiswem.synchronize()
    m = seiswem.DeviceArray((g.nm, g.nz))
    for _ in range(warmup):
        plan.migrate(d, sx, sy, m=m)
    seiswem.synchronize()
    march_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.migrate(d, sx, sy, m=m)
        march_ms += plan.timing()["march_ms"]
    seiswem.synchronize()
    elapsed = time.perf_counter() - t0
    s = g.sampling()
    nf, nk = s["ifmax"] - s["ifmin"], s["nkx"] * s["nky"]
    byts = 64.0 * nk * nf * g.nz
    march_s = march_ms / steps / 1e3
    achieved = byts / march_s / 1e9
    f_ss = 10.0 * nk * np.log2(nk) + 6.0 * nk + 8.0 * g.nm
    flops = (2.0 * f_ss + 4.0 * g.nm) * nf * g.nz
    tflops = flops / march_s / 1e12
    # with graphs, the same shots again with the kernels launched one by one (without graphs no
    # graph is captured at all: profilers run the one-by-one path alone)
    elapsed_eager = elapsed
    if graphs:
        plan.set_graphs(False)
        plan.migrate(d, sx, sy, m=m)
        seiswem.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            plan.migrate(d, sx, sy, m=m)
        seiswem.synchronize()
        elapsed_eager = time.perf_counter() - t1
        plan.set_graphs(True)
    out = {"shots_per_s": steps / elapsed, "ms_per_shot": elapsed / steps * 1e3, "steps": steps, "warmup": warmup,
           "launch": "one captured graph per shot" if graphs else "kernels launched one by one",
           "ms_per_shot_graph": elapsed / steps * 1e3 if graphs else None,
           "ms_per_shot_launched_one_by_one": elapsed_eager / steps * 1e3,
           "workload": f"3-D SSF migration, 1 shot, nx=ny={g.nmx} nz={g.nz} nt=1500 3-50 Hz",
           "nkx": s["nkx"], "nky": s["nky"], "nf": nf, "nz": g.nz,
           "roofline": {"bound": "valu", "achieved": tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": tflops / PEAK_FP32_TFLOPS, "traffic": None, "kernel": "march3d x-pass + y-pass",
                        "algorithmic_flops_per_shot": flops,
                        "note": "algorithmic flops 2 F + 4 nm per (w, iz), F = 10 nk log2 nk + 6 nk + 8 nm; "
                                "the x-pass keeps the VALU 88 % busy (profiles/march3d_ssf_512x512_sq_counters.md)"},
           "hbm": {"achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                   "algorithmic_bytes_per_shot": byts,
                   "note": "64 nk bytes per (w, iz); the kernels skip the pad rows and move half"},
           "dot_test_rel": None}
    if dot:
        rng = np.random.default_rng(7)
        mr = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
        dr = rng.standard_normal((1, g.nm, g.nt)).astype(np.float32)
        lm = plan.model(seiswem.DeviceArray(mr.shape, mr), sx, sy).download()
        lhd = plan.migrate(seiswem.DeviceArray(dr.shape, dr), sx, sy).download()
        a, b = float(np.sum(dr.astype(np.float64) * lm)), float(np.sum(mr.astype(np.float64) * lhd))
        out["dot_test_rel"] = abs(a - b) / abs(b)
    plan.close()
    return out


def bench_ssf3d(args):
    r = measure_ssf3d(args, args.steps, args.warmup, grid=args.grid, nz=args.nz, dot=not args.no_dot,
                      graphs=not args.no_graphs)
    out = {
        "metric": "migrated shots/sec (3D split-step) at <=1e-4 rel-L2 vs FP64", "value": r["shots_per_s"],
        "unit": "shots/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["ms_per_shot"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (3-D generator; record by the library's own demigration)",
        "config": {"workload": r["workload"], "nkx": r["nkx"], "nky": r["nky"], "nf": r["nf"], "nz": r["nz"]},
        "roofline": r["roofline"], "hbm": r["hbm"], "dot_test_rel": r["dot_test_rel"], "launch": r["launch"],
        "ms_per_shot_launched_one_by_one": r["ms_per_shot_launched_one_by_one"],
        "ms_per_shot_graph": r["ms_per_shot_graph"],
    }
    if args.nz:
        out["config"]["workload"] += f" (first {args.nz} depths only)"
    if args.copy_calibration:
        import seiswem
        src = seiswem.DeviceArray((COPY_BYTES // 4,))
        dst = seiswem.DeviceArray((COPY_BYTES // 4,))
        src.zero()
        seiswem.synchronize()
        seiswem.copy_device(dst, src)
        seiswem.synchronize()
    print(json.dumps(out))


def measure_serving(args):
    """Serving workload: Poisson arrivals (default 50 requests/s) of mixed 2-D / 3-D SSF
    migration requests (80 % 2-D: nx 128..1024, nz 200/400, nt 1000..4000; 20 % 3-D: 128^2 or
    256^2 x 200).  Reports p50 / p99 latency (submission -> image ready) at that rate, the
    saturated served throughput (all requests queued at once) and the plain per-request path
    (plan, upload, migrate, download, free; one request at a time) on a sample of the same
    requests."""
    import threading

    import seiswem
    from seiswem import synth

    records = {}
    reqs = [synth.config4_request(i, records=records) for i in range(args.requests)]
    arrivals = synth.poisson_arrivals(args.requests, args.rate)
    srv = seiswem.Server(streams=4, max_inflight=8)
    # warm-up: one request of every band, so the band bases are resident as in steady state
    seen = {}
    for r in reqs:
        seen.setdefault((r[0].nt, r[0].nmy > 1), r)
    for r in seen.values():
        srv.wait(srv.submit(*r))

    def run(times):
        ids = [None] * len(reqs)
        t0 = time.perf_counter()

        def submitter():
            for i, r in enumerate(reqs):
                dt = times[i] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                ids[i] = srv.submit(*r)

        th = threading.Thread(target=submitter)
        th.start()
        lat, done = [], []
        for i in range(len(reqs)):
            while ids[i] is None:
                time.sleep(1e-4)
            _, st = srv.wait(ids[i])
            if st["state"] != seiswem.REQ_DONE:
                raise RuntimeError(st["message"])
            lat.append(st["done_ms"] - st["submit_ms"])
            done.append(st["done_ms"])
        th.join()
        return np.array(lat), time.perf_counter() - t0

    lat, wall = run(arrivals)
    _, wall_sat = run(np.zeros(len(reqs)))
    served = len(reqs) / wall_sat
    srv.close()
    k = min(args.plain_sample, len(reqs))
    t0 = time.perf_counter()
    for r in reqs[:k]:
        seiswem.ShotProfileWEM(r[3], True, vel=r[1], wav=r[2], geom=r[0], sx=r[4], sy=r[5])
    plain = k / (time.perf_counter() - t0)
    n3 = sum(1 for r in reqs if r[0].nmy > 1)
    return {
        "workload": f"serving: {len(reqs)} requests ({n3} 3-D), Poisson {args.rate:g}/s", "requests": len(reqs),
        "requests_3d": n3, "streams": 4, "max_inflight": 8, "warmup_requests": len(seen),
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "mean": float(lat.mean()), "max": float(lat.max())},
        "served_shots_per_s": served, "poisson_run_s": wall, "saturated_run_s": wall_sat,
        "plain_shots_per_s": plain, "plain_sample": k, "served_vs_plain": served / plain,
    }


def measure_survey(args):
    """Serving the shots of one survey: the 2-D PSPI workload's 64 shots (nx 1024, nz 500, nt 2000,
    0-80 Hz, one velocity and wavelet) submitted as 64 single-shot requests, Poisson arrivals at
    --survey-rate (default 100/s) and all at once, on a server that fuses requests of one survey
    queued together into one launch (max_fuse 32) and on one that does not (max_fuse 1); beside
    them the plain per-request path on a sample."""
    import threading

    import seiswem
    from seiswem import synth

    g, vel, _, wav, sx = synth.config2(nshot=64)
    rng = np.random.default_rng(7)
    recs = rng.standard_normal((8, 1, g.nm, g.nt)).astype(np.float32)  # cost does not depend on values
    reqs = [(g, vel, wav, recs[i % 8], sx[i:i + 1], None) for i in range(sx.size)]
    arrivals = synth.poisson_arrivals(len(reqs), args.survey_rate, seed=5)

    def run(srv, times):
        ids = [None] * len(reqs)
        t0 = time.perf_counter()

        def submitter():
            for i, r in enumerate(reqs):
                dt = times[i] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                ids[i] = srv.submit(*r)

        th = threading.Thread(target=submitter)
        th.start()
        lat = []
        for i in range(len(reqs)):
            while ids[i] is None:
                time.sleep(1e-4)
            _, st = srv.wait(ids[i])
            if st["state"] != seiswem.REQ_DONE:
                raise RuntimeError(st["message"])
            lat.append(st["done_ms"] - st["submit_ms"])
        th.join()
        return np.array(lat), time.perf_counter() - t0

    out = {"workload": f"serving one survey: {len(reqs)} single-shot requests of the 2-D PSPI workload "
                       f"(nx 1024, nz 500, nt 2000, 641 bins), Poisson {args.survey_rate:g}/s and all at once"}
    for name, fuse in (("fused", 32), ("unfused", 1)):
        srv = seiswem.Server(streams=4, max_inflight=8, max_fuse=fuse)
        srv.wait(srv.submit(*reqs[0]))  # the band's bases resident, as in steady state
        lat, wall = run(srv, arrivals)
        _, wall_sat = run(srv, np.zeros(len(reqs)))
        st = srv.stats()
        srv.close()
        out[name] = {"max_fuse": fuse, "served_shots_per_s": len(reqs) / wall_sat,
                     "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
                     "poisson_run_s": wall, "saturated_run_s": wall_sat,
                     "fused_launches": st["fused_launches"], "fused_requests": st["fused_requests"]}
    k = min(args.plain_sample, 8)
    t0 = time.perf_counter()
    for r in reqs[:k]:
        seiswem.ShotProfileWEM(r[3], True, vel=r[1], wav=r[2], geom=r[0], sx=r[4])
    out["plain_shots_per_s"] = k / (time.perf_counter() - t0)
    out["plain_sample"] = k
    out["fused_vs_plain"] = out["fused"]["served_shots_per_s"] / out["plain_shots_per_s"]
    out["fused_vs_unfused"] = out["fused"]["served_shots_per_s"] / out["unfused"]["served_shots_per_s"]
    return out


def bench_serving(args):
    r = measure_serving(args)
    out = {
        "metric": "migrated shots/sec (2D & 3D) at <=1e-4 rel-L2 vs FP64; p50/p99 request latency",
        "value": r["served_shots_per_s"], "unit": "shots/s", "n_gpus": 1, "steps": r["requests"],
        "warmup": r["warmup_requests"], "ms_per_step": 1e3 / r["served_shots_per_s"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (per-request seeded velocity; random records)",
        "config": {"workload": r["workload"], "streams": 4, "max_inflight": 8},
        "latency_ms": r["latency_ms"], "poisson_run_s": r["poisson_run_s"], "saturated_run_s": r["saturated_run_s"],
        "plain_shots_per_s": r["plain_shots_per_s"], "plain_sample": r["plain_sample"],
        "served_vs_plain": r["served_vs_plain"],
    }
    if not args.no_survey:
        out["survey"] = measure_survey(args)
    print(json.dumps(out))


def bench_lsm(args):
    """Least-squares migration and precision study: the config-1 geometry with 32 shots
    (sx = 40 + 80 k), 10 CG iterations (ConjugateGradients.jl, mu = 0) with the GPU operator pair,
    model / data / residual resident on the device.  Observed records: FP64 oracle demigration of
    the config-1 reflectivity (computed here on the host cores, checked against the fingerprint of
    the committed FP64 reference).  Variants: FP64 end to end (the oracle; committed result of
    tests/golden/make_cg_fp64.py), FP32 propagation + FP64 kz tables + compensated sums (the
    default), naive FP32."""
    import oracle as O
    import seiswem
    from seiswem import solvers, synth

    ref = np.load(os.path.join(ROOT, "tests", "golden", "cg_fp64_config5.npz"))
    g, vel, refl, wav, _ = synth.config1()
    sx = synth.config5_shots()
    p = O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, fmin=g.fmin, fmax=g.fmax, padt=g.padt,
                    padx=g.padx, pspi=False)
    cores = host_cores()
    t0 = time.perf_counter()
    d_obs = O.shots(p, vel, wav, sx, m=refl, adj=False, f64=True)
    t_oracle = time.perf_counter() - t0
    fp_ok = abs(d_obs.sum() - ref["d_sum"]) <= 1e-9 * abs(ref["d_sumsq"]) ** 0.5 * d_obs.size ** 0.5 and \
        abs((d_obs ** 2).sum() / ref["d_sumsq"] - 1) <= 1e-9
    d_obs = d_obs.astype(np.float32)
    img64 = ref["image"]
    variants = {}
    rng = np.random.default_rng(5)
    mr = rng.standard_normal((g.nm, g.nz)).astype(np.float32)
    dr = rng.standard_normal((sx.size, g.nm, g.nt)).astype(np.float32)
    for name, prec in (("fp32_prop_fp64_kz_compensated", seiswem.PREC_KZ64), ("naive_fp32", seiswem.PREC_FP32)):
        plan = seiswem.Plan(g, vel, wav, precision=prec)
        dd = seiswem.DeviceArray(d_obs.shape, d_obs)
        solvers.ConjugateGradientsDevice(plan, dd, sx, Niter=1)  # warm-up
        seiswem.synchronize()
        t0 = time.perf_counter()
        m, cost = solvers.ConjugateGradientsDevice(plan, dd, sx, Niter=10)
        seiswem.synchronize()
        t_cg = time.perf_counter() - t0
        img = m.download().astype(np.float64)
        lm = plan.model(seiswem.DeviceArray(mr.shape, mr), sx).download()
        lhd = plan.migrate(seiswem.DeviceArray(dr.shape, dr), sx).download()
        a, b = float(np.sum(dr.astype(np.float64) * lm)), float(np.sum(mr.astype(np.float64) * lhd))
        napp = 1 + 2 * (len(cost) - 1)
        variants[name] = {
            "cost": [float(c) for c in cost],
            "image_rel_l2_vs_fp64": float(np.linalg.norm(img - img64) / np.linalg.norm(img64)),
            "dot_test_rel": abs(a - b) / abs(b),
            "cg_s": t_cg, "s_per_iteration": t_cg / (len(cost) - 1),
            "shots_per_s": sx.size * napp / t_cg,
        }
        plan.close()
    variants["fp64_oracle"] = {
        "cost": [float(c) for c in ref["cost"]], "image_rel_l2_vs_fp64": 0.0, "dot_test_rel": None,
        "shots_per_s": sx.size / t_oracle, "cores": cores,
        "note": "committed result of tests/golden/make_cg_fp64.py; throughput from one 32-shot FP64 oracle "
                "demigration on this host",
    }
    main_v = variants["fp32_prop_fp64_kz_compensated"]
    out = {
        "metric": "LSM: migrated shots/sec at <=1e-4 rel-L2 vs FP64 (10 CG iterations)", "value": main_v["shots_per_s"],
        "unit": "shots/s", "n_gpus": 1, "steps": 10, "warmup": 1, "ms_per_step": main_v["s_per_iteration"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (config-1 model; records by FP64 oracle demigration)",
        "config": {"workload": "least-squares migration, config-1 geometry, 32 shots, 10 CG iterations, mu=0"},
        "records_match_reference": bool(fp_ok), "variants": variants,
    }
    print(json.dumps(out))


def plain_single_shot(args, g, vel, wav, sx, d, plan, batched):
    """The single-shot paths beside the batched launch: (a) the plain drop-in, one shot per
    wem() call (host arrays; the plan, its phase tables, the uploads, the march, the download
    and the frees are all inside each call, as one shotwem run per shot does), and (b) one shot
    per launch on a resident plan with the record already on the device."""
    import seiswem

    k = min(args.plain_shots, int(sx.size))
    d_host = d.download()[:k]
    seiswem.ShotProfileWEM(d_host[:1], True, vel=vel, wav=wav, geom=g, sx=sx[:1])  # warm-up
    seiswem.synchronize()
    t0 = time.perf_counter()
    for i in range(k):
        seiswem.ShotProfileWEM(d_host[i:i + 1], True, vel=vel, wav=wav, geom=g, sx=sx[i:i + 1])
    plain = k / (time.perf_counter() - t0)
    recs = [seiswem.DeviceArray((1, g.nm, g.nt), d_host[i:i + 1]) for i in range(k)]
    img = seiswem.DeviceArray((g.nm, g.nz))
    plan.migrate(recs[0], sx[:1], m=img)
    seiswem.synchronize()
    t0 = time.perf_counter()
    for i in range(k):
        plan.migrate(recs[i], sx[i:i + 1], m=img)
    seiswem.synchronize()
    resident = k / (time.perf_counter() - t0)
    return {
        "plain_single_shot": {"shots_per_s": plain, "sample_shots": k,
                              "path": "ShotProfileWEM per shot on host arrays: plan, phase tables, upload, "
                                      "march, download, free in every call"},
        "single_shot_resident_plan": {"shots_per_s": resident, "sample_shots": k,
                                      "path": "one shot per launch, plan and record resident in HBM"},
        "batched_vs_plain_single_shot": batched / plain,
        "batched_vs_single_shot_resident": batched / resident,
    }


def bench_angles(args):
    """Source incidence angles (compute_angles, wem.c:566-992) on the config-2 grid: nx=1024,
    nz=500, nt=2000 (ntfft = nt: 321 frequencies to 80 Hz), one shot per call of the drop-in
    swem_compute_angles (host arrays in and out, plan and volume buffers inside the call), shots
    spread over the line.  CPU baseline: the FP32 oracle on one shot of a reduced band."""
    import seiswem
    from seiswem import synth

    g, vel, _, wav, sx = synth.config2(nshot=args.shots)
    geo = seiswem.WemGeometry(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, fmin=0.0, fmax=80.0,
                              pspi=False)
    for s in sx[: args.warmup]:
        seiswem.compute_angles(vel, wav, geo, float(s))
    k = min(args.steps, sx.size)
    t0 = time.perf_counter()
    for s in sx[:k]:
        angx, _ = seiswem.compute_angles(vel, wav, geo, float(s))
    el = time.perf_counter() - t0
    out = {
        "metric": "source incidence angles: shots/sec (compute_angles, 2-D)", "value": k / el, "unit": "shots/s",
        "n_gpus": 1, "steps": k, "warmup": args.warmup, "ms_per_step": el / k * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (config-2 velocity generator)",
        "config": {"workload": "compute_angles, nx=1024 nz=500 nt=2000 dt=2 ms, 0-80 Hz (321 bins), one shot per call"},
    }
    if not args.no_cpu:
        import oracle as O
        cores = host_cores()
        os.environ.setdefault("OMP_NUM_THREADS", str(cores))
        p = O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, fmin=76.0, fmax=80.0, pspi=False)
        t0 = time.perf_counter()
        O.angles(p, vel, wav, float(sx[0]), f64=False)
        tc = time.perf_counter() - t0
        sp = seiswem.WemGeometry(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, fmin=76.0, fmax=80.0,
                                 pspi=False, padt=1, padx=1).sampling()
        nb = sp["ifmax"] - sp["ifmin"]  # 17 bins
        out["cpu_baseline"] = {"value": nb / tc / 321, "unit": "shots/s", "cores": 1, "kind": "port",
                               "sample": f"1 shot, {nb} of 321 frequency bins (76-80 Hz) at full nx/nz, {tc:.2f} s, "
                                         "scaled by bins (the FP32 oracle is serial over frequencies)"}
    print(json.dumps(out))


def bench_elastic(args):
    """Elastic migration (ewem.c) on the config-2 grid: nx=1024, nz=500, nt=2000, dt=2 ms, 0-80 Hz
    (padt = padx = 2: 641 bins, nkx = 2048), vp = the config-2 velocity, vs = vp / 1.8, synthetic
    three-component records; one shot per call of swem_ewem (host arrays in and out).  CPU
    baseline: the FP32 oracle on one shot of a reduced band."""
    import seiswem
    from seiswem import synth

    g, vp, _, wav, sx = synth.config2(nshot=args.shots)
    vs = (vp / 1.8).astype(np.float32)
    rng = np.random.default_rng(0)
    comps = rng.standard_normal((3, g.nm, g.nt)).astype(np.float32)
    geo = (g.nt, 0.0, g.dt, g.nmx, 0.0, g.dmx, 1, 0.0, g.dmx)

    def shot(s):
        return seiswem.ewem(comps[0], comps[1], comps[2], None, None, None, wav, *geo, float(s), 0.0, g.nz, 0.0,
                            g.dz, 0.0, 0.0, vp, vs, 0.0, 80.0, 2, 2, True)

    geom = seiswem.WemGeometry(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, fmin=0.0, fmax=80.0,
                               pspi=False)
    nb = min(8, sx.size)  # shots per batched call
    batch = [np.broadcast_to(c, (nb, g.nm, g.nt)) for c in comps]

    def batched():
        return seiswem.ShotProfileEWEM(batch, True, vp=vp, vs=vs, wav=wav, geom=geom, sx=sx[:nb])

    for s in sx[: args.warmup]:
        shot(s)
    batched()
    k = min(args.steps, sx.size)
    t0 = time.perf_counter()
    for s in sx[:k]:
        shot(s)
    el = time.perf_counter() - t0
    t0 = time.perf_counter()
    batched()
    eb = time.perf_counter() - t0
    out = {
        "metric": "elastic migration: shots/sec (ewem, 2-D)", "value": nb / eb, "unit": "shots/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": eb * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (config-2 velocity, vs = vp / 1.8, random records)",
        "config": {"workload": f"ewem, nx=1024 nz=500 nt=2000 dt=2 ms, 0-80 Hz (641 bins), nkx=2048, {nb} shots per "
                               "call (swem_eshots, host arrays in and out)"},
        "single_shot_calls": {"shots_per_s": k / el, "ms_per_shot": el / k * 1e3, "path": "swem_ewem, one shot per call"},
    }
    if not args.no_cpu:
        import oracle as O
        cores = host_cores()
        p = O.WemParams(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, fmin=79.0, fmax=80.0, pspi=False)
        sp = seiswem.WemGeometry(nt=g.nt, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nz=g.nz, dz=g.dz, fmin=79.0, fmax=80.0,
                                 pspi=False).sampling()
        nb = sp["ifmax"] - sp["ifmin"]
        t0 = time.perf_counter()
        O.ewem(p, vp, vs, wav, float(sx[0]), ux=comps[0], uz=comps[2], adj=True, f64=False)
        tc = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": nb / tc / 641, "unit": "shots/s", "cores": 1, "kind": "port",
                               "sample": f"1 shot, {nb} of 641 frequency bins (79-80 Hz) at full nx/nz, {tc:.2f} s, "
                                         "scaled by bins (the FP32 oracle is serial over frequencies)"}
    print(json.dumps(out))


def bench_elastic3d(args):
    """3-D elastic migration (ewem.c with nmy > 1) on the 3-D SSF grid: nx=ny=256, dx=dy=12.5,
    nz=300, dz=10, nt=1500, dt=4 ms, 3-50 Hz (ewem.c's band: ntfft 3000, bins 36-200, nkx=nky=512),
    vp = the 3-D generator's velocity, vs = vp / 1.8, synthetic random ux, uz records; one shot per
    call of swem_ewem (host arrays in and out).  CPU baseline: the FP32 oracle on one shot, one
    frequency bin and the first depths at full nx, ny, scaled by bins and depths."""
    import seiswem
    from seiswem import synth

    g, vp, _, wav, sxa, sya = synth.config3()
    vs = (vp / 1.8).astype(np.float32)
    rng = np.random.default_rng(0)
    comps = rng.standard_normal((3, g.nm, g.nt)).astype(np.float32)
    sx, sy = float(sxa[0]), float(sya[0])

    def shot():
        return seiswem.ewem(comps[0], comps[1], comps[2], None, None, None, wav, g.nt, 0.0, g.dt, g.nmx, 0.0, g.dmx,
                            g.nmy, 0.0, g.dmy, sx, sy, g.nz, 0.0, g.dz, 0.0, 0.0, vp, vs, g.fmin, g.fmax, 2, 2, True)

    for _ in range(args.warmup):
        shot()
    k = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(k):
        shot()
    el = time.perf_counter() - t0
    ntfft = int(2 * np.trunc(np.float32(2 * g.nt) / 2))
    ifmin, ifmax = int(np.trunc(np.float32(g.fmin) * np.float32(g.dt) * ntfft)), \
        int(np.trunc(np.float32(g.fmax) * np.float32(g.dt) * ntfft)) + 1
    nf = ifmax - ifmin
    out = {
        "metric": "elastic migration: shots/sec (ewem, 3-D)", "value": k / el, "unit": "shots/s", "n_gpus": 1,
        "steps": k, "warmup": args.warmup, "ms_per_step": el / k * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (3-D generator velocity, vs = vp / 1.8, random records)",
        "config": {"workload": f"3-D ewem, nx=ny=256 nz=300 nt=1500 dt=4 ms, 3-50 Hz ({nf} bins), nkx=nky=512, "
                               "one shot per call (swem_ewem, host arrays in and out)"},
    }
    if not args.no_cpu:
        import oracle as O
        nzs, nts = 1, 64  # one depth; 64-sample records (ntfft 128: the 50 Hz bin alone)
        p = O.WemParams(nt=nts, dt=g.dt, nmx=g.nmx, dmx=g.dmx, nmy=g.nmy, dmy=g.dmy, nz=nzs, dz=g.dz, fmin=50.0,
                        fmax=50.0, padt=2, padx=2, pspi=False)
        nb = 1
        t0 = time.perf_counter()
        O.ewem(p, np.ascontiguousarray(vp[:, :nzs]), np.ascontiguousarray(vs[:, :nzs]), wav[:nts], sx, sy,
               ux=np.ascontiguousarray(comps[0][:, :nts]), uz=np.ascontiguousarray(comps[2][:, :nts]), adj=True,
               f64=False)
        tc = time.perf_counter() - t0
        scale = (nb / nf) * (nzs / g.nz)
        out["cpu_baseline"] = {"value": scale / tc, "unit": "shots/s", "cores": 1, "kind": "port",
                               "sample": f"1 shot, {nb} of {nf} frequency bins (50 Hz) and {nzs} of {g.nz} depths "
                                         f"at full nx, ny (nkx=nky=512), records cut to {nts} samples so the record "
                                         f"transforms stay out of the sample, {tc:.2f} s, scaled by bins and depths "
                                         "(the FP32 oracle is serial)"}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--freqs-per-block", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0, help="2-D kernel variant (SWEM_VARIANT_*, 0 = automatic)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nz", type=int, default=0, help="ssf3d: march only the first NZ depths (profiling)")
    ap.add_argument("--no-dot", action="store_true", help="ssf3d: skip the full-size dot test")
    ap.add_argument("--no-graphs", action="store_true",
                    help="ssf3d: launch each shot's kernels one by one (the timed steps) instead of one captured graph")
    ap.add_argument("--grid", type=int, default=256,
                    help="ssf3d: nx = ny of the 3-D grid (256: the 3-D configuration; others, e.g. 250, take "
                         "the Bluestein kernels)")
    ap.add_argument("--plain-shots", type=int, default=8,
                    help="pspi2d: shots timed one per call on the plain and resident single-shot paths (0: skip)")
    ap.add_argument("--workload", choices=("pspi2d", "ssf3d", "serving", "lsm", "angles", "elastic", "elastic3d"),
                    default="pspi2d")
    ap.add_argument("--no-3d", action="store_true", help="pspi2d: leave out the 3-D part of the line")
    ap.add_argument("--no-serving", action="store_true", help="pspi2d: leave out the serving part of the line")
    ap.add_argument("--steps-3d", type=int, default=3, help="pspi2d: 3-D shots timed for the line's 3-D part")
    ap.add_argument("--requests", type=int, default=120, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=50.0, help="serving: Poisson arrival rate (requests/s)")
    ap.add_argument("--plain-sample", type=int, default=24, help="serving: requests timed on the plain path")
    ap.add_argument("--survey-rate", type=float, default=100.0,
                    help="serving: Poisson arrival rate (requests/s) of the one-survey scenario")
    ap.add_argument("--no-survey", action="store_true", help="serving: leave out the one-survey scenario")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed steps, copy a 1 GiB device buffer once (the reference kernel that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not run in this build (shots would shard across "
              "GPUs with one image all-reduce)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "ssf3d":
        return bench_ssf3d(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "lsm":
        return bench_lsm(args)
    if args.workload == "angles":
        return bench_angles(args)
    if args.workload == "elastic":
        return bench_elastic(args)
    if args.workload == "elastic3d":
        return bench_elastic3d(args)

    import seiswem
    from seiswem import synth

    g, vel, refl, wav, sx = synth.config2(nshot=args.shots)
    nshot = int(sx.size)
    plan = seiswem.Plan(g, vel, wav, freqs_per_block=args.freqs_per_block, variant=args.variant)
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
    packed = args.variant in (0, 6)  # automatic selection takes the packed kernel for PSPI compat
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic_per_launch(nshot, "march2d_adj_packed" if packed else "march2d_adj_pair"),
                "kernel": "march2d_adj_packed<2048,16>" if packed else "march2d_adj (PSPI, nkx=2048)",
                "compute_unit": "FP32 vector ALU (no stage is a matrix product); peak = MI355X FP32 157.3 TFLOP/s",
                "flops_basis": "reference operation count: 2 (1 + nref) transforms per depth and frequency"
                               + ("; the packed kernel executes 3 + nref, and nref where at most four "
                                  "points select the top vref" if packed else ""),
                "avg_launch_ms": march_ms / launches, "flops_per_launch": flops_per_launch}
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
    if args.plain_shots > 0:
        out.update(plain_single_shot(args, g, vel, wav, sx, d, plan, value))
    if not args.no_cpu:
        cores = host_cores()
        os.environ.setdefault("OMP_NUM_THREADS", str(cores))

        def m_gpu_fn(gs, dtr):
            return seiswem.ShotProfileWEM(dtr[None], True, vel=vel, wav=wav, geom=gs, sx=sx[:1], variant=args.variant)

        cb, err, nf_samp = cpu_baseline(g, vel, wav, float(sx[0]), m_gpu_fn, cores)
        out["cpu_baseline"] = cb
        out["rel_l2_vs_fp64"] = err
        out["rel_l2_sample"] = f"shot 0, {nf_samp} frequency bins, full nx/nz, PSPI compat"
    plan.close()
    del d, m, m_true
    # the metric's other two parts: 3-D shots/s (the 3-D configuration, one shot per step) and
    # the serving loop's p50 / p99 latency beside its throughput
    if not args.no_3d:
        out["ssf3d"] = measure_ssf3d(args, steps=args.steps_3d, warmup=1)
    if not args.no_serving:
        out["serving"] = measure_serving(args)
        if not args.no_survey:
            out["serving"]["survey"] = measure_survey(args)
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
