"""Benchmark: batched Monte Carlo Kelvin-Helmholtz ensemble on one MI355X.

Headline workload (one "step" = one SSP-RK3 timestep of every sample):
  2D Euler Kelvin-Helmholtz, 256^2 cells per sample on [0,1]^2, periodic,
  gamma 1.4, WENO2 + HLL3, SSP-RK3, CFL 0.475, FP64 state, 1024 samples
  advanced together; each sample's 40 perturbation parameters are the
  reference sampler's draws (global std::mt19937_64, uniform[0,1]) for that
  sample. Inputs are resident in HBM before the timed region starts.

Reported beside it (extra fields, one GPU): the same batch run one sample at a
time (plain path), the FP32-state variant with its error against the FP64
result, a mixed-size ensemble (128^2/256^2/512^2), the 1024^2 single solve,
the 3D FP32 256^3 x 64 ensemble, the statistics pass and the Sod shock tube;
the roofline of the dominant kernel (the fused stage kernel, timed with HIP
events on the ensemble's stream) and a CPU baseline (the oracle on the host
cores, a bounded sample).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--only main]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baseline's OpenMP threads stay on neighbouring cores (read when the
# oracle's OpenMP runtime starts)
os.environ.setdefault("OMP_PROC_BIND", "close")

HBM_PEAK = 8.0e12
PER_GRAPH = 10


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--sustained-steps", type=int, default=300,
                    help="steps of the sustained window timed after the K timed steps (0: none)")
    ap.add_argument("--samples", type=int, default=1024)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--plain-steps", type=int, default=100)
    ap.add_argument("--extra-steps", type=int, default=40)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--serving-requests", type=int, default=200)
    ap.add_argument("--serving-tmin", type=float, default=0.1)
    ap.add_argument("--serving-tmax", type=float, default=1.0)
    ap.add_argument("--only", choices=["all", "main", "fe-calibration", "serving", "full-uq", "kh3d",
                                       "stats", "single", "sod"],
                    default="all")
    ap.add_argument("--samples3d", type=int, default=64)
    ap.add_argument("--n3d", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rounded(k):
    return max(PER_GRAPH, (k // PER_GRAPH) * PER_GRAPH)


def graph_steps(k):
    """Steps per captured graph for a run of exactly k steps."""
    for g in (PER_GRAPH, 5, 4, 2):
        if k % g == 0:
            return g
    return 1


def run_exact(e, k):
    """Exactly k timesteps on graphs of graph_steps(k) steps."""
    if k > 0:
        e.run_steps(k, graph_steps(k))


def timed_exact(e, k):
    """Wall time of exactly k timesteps, the graph captured beforehand."""
    g = graph_steps(k)
    e.run_steps(0, g)  # capture and instantiate outside the timed region
    e.synchronize()
    t0 = time.perf_counter()
    e.run_steps(k, g)
    e.synchronize()
    return time.perf_counter() - t0


_HOST_CPU = []


def host_cpu():
    """CPU model name and physical core count of this host (/proc/cpuinfo),
    and the cores this process may run on, read on the first call (main()
    calls it first: once OpenMP with OMP_PROC_BIND has run, the calling
    thread is bound to one core and its affinity mask says so)."""
    if _HOST_CPU:
        return _HOST_CPU[0]
    model, cores = "unknown", set()
    phys = core = None
    try:
        for ln in open("/proc/cpuinfo"):
            k, _, v = ln.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name":
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None and core is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None and core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    _HOST_CPU.append((model, len(cores) or (os.cpu_count() or 1), allowed))
    return _HOST_CPU[0]


def kh_scheme(hip, precision, eq=None):
    return hip.make_scheme(eq or hip.EQ_EULER2, hip.FLUX_HLL3, hip.RECON_WENO2, hip.INT_RK3,
                           hip.BC_PERIODIC, gamma=1.4, cfl=0.475,
                           precision=hip.FP64 if precision == "fp64" else hip.FP32)


def make_batch(hip, sizes, precision, X, dim=2):
    from alsvinn_amd import initial_data as ic
    if dim == 2:
        scheme = kh_scheme(hip, precision)
        samples = [hip.make_sample((n, n, 1), (0, 0, 0), (1, 1, 0), 2.0) for n in sizes]
    else:
        scheme = kh_scheme(hip, precision, hip.EQ_EULER3)
        samples = [hip.make_sample((n, n, n), (0, 0, 0), (1, 1, 1), 2.0) for n in sizes]
    e = hip.Ensemble(scheme, samples)
    for i, n in enumerate(sizes):
        u = ic.kelvin_helmholtz_2d(n, X[i]) if dim == 2 else ic.kelvin_helmholtz_3d(n, X[i])
        e.upload(i, u)
    e.apply_boundary()
    e.synchronize()
    return e


def timed_steps(e, steps):
    e.synchronize()
    t0 = time.perf_counter()
    e.run_steps(steps, PER_GRAPH)
    e.synchronize()
    return time.perf_counter() - t0


def throughput(e, cells, steps, warm=PER_GRAPH):
    e.run_steps(rounded(warm), PER_GRAPH)
    e.synchronize()
    steps = rounded(steps)
    secs = timed_steps(e, steps)
    return cells * steps / secs, 1e3 * secs / steps, steps


def cpu_baseline(n, X0, steps):
    """The oracle (restatement of the reference CPU path: OpenMP over rows,
    separate passes) on the host cores, one sample, `steps` steps."""
    from tests import oracle_lib as O
    from alsvinn_amd import initial_data as ic
    s = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC,
                     (n, n, 1), (0, 0, 0), (1, 1, 0), cfl=0.475, end_time=2.0, max_steps=steps)
    u0 = ic.kelvin_helmholtz_2d(n, X0)
    model, physical, allowed = host_cpu()
    t0 = time.perf_counter()
    rc, out, t, st, _ = O.run(s, u0)
    dt = time.perf_counter() - t0
    assert rc == 0 and st == steps
    threads = int(os.environ.get("OMP_NUM_THREADS", allowed))
    return {"value": n * n * steps / dt, "unit": "cell-updates/s", "cores": threads,
            "kind": "port", "cpu_model": model, "host_physical_cores": physical,
            "cores_allowed": allowed, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "sample": "oracle (C++ restatement of the reference cpu platform, OpenMP over rows, "
                      "-O3 -ffp-contract=off) on %d OpenMP threads (OMP_PROC_BIND=%s) of a %s "
                      "host with %d physical cores, 1 sample of the same %d^2 KH WENO2+HLL3 "
                      "SSP-RK3 FP64 workload, %d steps in %.1f s"
                      % (threads, os.environ.get("OMP_PROC_BIND"), model, physical, n, steps, dt)}


def fp32_error(hip, M, n, X, steps):
    """FP32-state batch vs the FP64 batch (bit-identical to the oracle) after
    the same number of steps: relative L1 of rho (first 8 samples) and of the
    ensemble mean / variance of rho."""
    e64 = make_batch(hip, [n] * M, "fp64", X)
    e32 = make_batch(hip, [n] * M, "fp32", X)
    e64.autotune()
    sch = e32.autotune()
    v32, ms32, k = throughput(e32, M * n * n, steps)
    e64.run_steps(rounded(steps) + PER_GRAPH, PER_GRAPH)
    g = 2
    errs = []
    for i in range(min(8, M)):
        a = e64.download(i)[0, 0, g:-g, g:-g]
        b = e32.download(i)[0, 0, g:-g, g:-g]
        errs.append(np.abs(a - b).sum() / np.abs(a).sum())
    m64, v64 = e64.meanvar()
    m32, v32s = e32.meanvar()
    sl = (0, 0, slice(g, -g), slice(g, -g))
    out = {"value": v32, "unit": "cell-updates/s", "ms_per_step": ms32, "steps": k,
           "dtype": "f32 state, f64 CFL/statistics/totals", "schedule": sch,
           "bytes_per_cell_update": 144,
           "hbm_fraction": v32 * 144 / HBM_PEAK,
           "rho_rel_l1_vs_fp64": float(np.mean(errs)),
           "mean_rho_rel_l1_vs_fp64": float(np.abs(m64[sl] - m32[sl]).sum() / np.abs(m64[sl]).sum()),
           "var_rho_rel_l1_vs_fp64": float(np.abs(v64[sl] - v32s[sl]).sum()
                                           / max(np.abs(v64[sl]).sum(), 1e-300)),
           "after_steps": k + PER_GRAPH * 2,
           "workload": "KH 256^2 x %d, FP32 state" % M}
    e64.close()
    e32.close()
    return out


def _pair_error(hip, scheme_kw, samples, u0s, steps=None, mixed=False):
    """Runs the same solves with FP64 and FP32 state; relative L1 of rho (var 0)
    of the FP32 result against the FP64 one (which is bit-identical to the
    oracle), and of the ensemble mean/variance when there are several samples
    of one size (mixed: per grid size, the mean and largest per-sample error
    instead)."""
    out = {}
    res = {}
    for prec in ("fp64", "fp32"):
        sc = hip.make_scheme(**scheme_kw, precision=hip.FP64 if prec == "fp64" else hip.FP32)
        e = hip.Ensemble(sc, samples)
        for i, u in enumerate(u0s):
            e.upload(i, u)
        e.apply_boundary()
        if steps:
            e.run_steps(steps, PER_GRAPH)
        else:
            e.run(32)
        res[prec] = ([e.download(i) for i in range(len(samples))],
                     e.meanvar() if len(samples) > 1 and not mixed else None, e.info(0))
        e.close()
    a, b = res["fp64"][0], res["fp32"][0]
    errs = [np.abs(x[0] - y[0]).sum() / np.abs(x[0]).sum() for x, y in zip(a, b)]
    out["rho_rel_l1"] = float(np.mean(errs))
    if mixed:
        by = {}
        for x, err in zip(a, errs):
            by.setdefault("%dx%d" % x.shape[-2:], []).append(err)
        out["by_grid_with_ghosts"] = {k: {"samples": len(v), "mean": float(np.mean(v)),
                                          "max": float(np.max(v))} for k, v in by.items()}
    if res["fp64"][1] is not None:
        (m64, v64), (m32, v32) = res["fp64"][1], res["fp32"][1]
        out["mean_rho_rel_l1"] = float(np.abs(m64[0] - m32[0]).sum() / np.abs(m64[0]).sum())
        out["var_rho_rel_l1"] = float(np.abs(v64[0] - v32[0]).sum() / max(np.abs(v64[0]).sum(), 1e-300))
    out["steps_fp64"] = res["fp64"][2]["steps"]
    out["steps_fp32"] = res["fp32"][2]["steps"]
    return out


def fp32_errors(hip):
    """FP32-state error against FP64 for each configuration (sizes reduced
    where the full one would take long; stated per entry)."""
    from alsvinn_amd import initial_data as ic, sampling
    out = {}
    kw = dict(equation=hip.EQ_EULER1, flux=hip.FLUX_HLL, reconstruction=hip.RECON_NONE,
              integrator=hip.INT_FE, boundary=hip.BC_NEUMANN, cfl=0.9)
    out["sod_1024_T0.2"] = _pair_error(hip, kw, [hip.make_sample((1024, 1, 1), (-1, 0, 0), (1, 0, 0), 0.2)],
                                       [ic.sod(1024)])
    kw2 = dict(equation=hip.EQ_EULER2, flux=hip.FLUX_HLL3, reconstruction=hip.RECON_WENO2,
               integrator=hip.INT_RK3, boundary=hip.BC_PERIODIC, cfl=0.475)
    X = sampling.stl_mersenne_uniform(8, 40)
    out["kh_1024_100steps"] = _pair_error(hip, kw2, [hip.make_sample((1024, 1024, 1), (0, 0, 0), (1, 1, 0), 2.0)],
                                          [ic.kelvin_helmholtz_2d(1024, X[0])], steps=100)
    out["kh_128x8_T2"] = _pair_error(hip, kw2, [hip.make_sample((128, 128, 1), (0, 0, 0), (1, 1, 0), 2.0)
                                                for _ in range(8)],
                                     [ic.kelvin_helmholtz_2d(128, X[i]) for i in range(8)])
    kw3 = dict(kw2, equation=hip.EQ_EULER3)
    X3 = sampling.stl_mersenne_uniform(2, 400)
    # 256^3: the 100-mode interface perturbation (|s| < 0.003) crosses cell
    # centres only once dx < 0.004, so coarser 3D samples coincide
    out["kh3d_256x2_40steps"] = _pair_error(hip, kw3, [hip.make_sample((256, 256, 256), (0, 0, 0), (1, 1, 1), 2.0)
                                                       for _ in range(2)],
                                            [ic.kelvin_helmholtz_3d(256, X3[i]) for i in range(2)], steps=40)
    kwb = dict(equation=hip.EQ_BURGERS, flux=hip.FLUX_GODUNOV, reconstruction=hip.RECON_ENO2,
               integrator=hip.INT_RK2, boundary=hip.BC_PERIODIC, cfl=0.475)
    out["burgers_16384_T0.1"] = _pair_error(hip, kwb, [hip.make_sample((16384, 1, 1), (0, 0, 0), (1, 0, 0), 0.1)],
                                            [ic.burgers_fbm(16384, 0.5, 3)])
    sizes = [128] * 171 + [256] * 171 + [512] * 170
    Xm = sampling.stl_mersenne_uniform(len(sizes), 40)
    out["mixed_171x128_171x256_170x512_40steps"] = _pair_error(
        hip, kw2, [hip.make_sample((m, m, 1), (0, 0, 0), (1, 1, 0), 2.0) for m in sizes],
        [ic.kelvin_helmholtz_2d(m, Xm[i]) for i, m in enumerate(sizes)], steps=40, mixed=True)
    kwr = dict(kw2, boundary=hip.BC_NEUMANN)
    out["riemann4_256_T0.2"] = _pair_error(hip, kwr, [hip.make_sample((256, 256, 1), (0, 0, 0), (1, 1, 0), 0.2)],
                                           [ic.riemann_four_quadrant(256, ic.riemann_parameters(X[1][:16]))])
    return out


def serving_bench(args):
    """Capacity from a closed batch (every request at t = 0), then a Poisson
    stream at 80% of it; p50/p99 latency (arrival -> result on the host)."""
    from alsvinn_amd import serving
    t_range = (args.serving_tmin, args.serving_tmax)
    cfg = serving.ServerConfig()
    cfg.steps_per_tick = 50
    cfg.keep_results = False
    closed = serving.make_requests(max(8, args.serving_requests // 4), rate=0.0, seed=11,
                                   t_range=t_range)
    srv = serving.Server(cfg)
    srv.serve(closed)
    cap = len(closed) / srv.elapsed
    srv.close()
    stream = serving.make_requests(args.serving_requests, rate=0.8 * cap, seed=12, t_range=t_range)
    srv = serving.Server(cfg)
    srv.serve(stream)
    lat = serving.latency_stats(stream)
    out = {"capacity_req_per_s": cap, "arrival_rate_req_per_s": 0.8 * cap,
           "requests": len(stream), "latency": lat,
           "cell_updates_per_s": srv.cell_updates / srv.elapsed, "elapsed_s": srv.elapsed,
           "end_time_range": list(t_range),
           "workload": "Poisson stream at 80% of measured capacity: 1D periodic Burgers "
                       "4096-65536 cells (fBm, ENO2+Godunov, SSP-RK2) and 2D Euler four-quadrant "
                       "Riemann 256^2-1024^2 (WENO2+HLL3, SSP-RK3, Neumann), FP64"}
    srv.close()
    return out


def sod_solve(hip):
    """Sod 1024 (HLL, forward Euler, CFL 0.9, T = 0.2): a latency-bound
    small solve, timed with the whole-step launch and with the multi-launch
    step (the same 248 steps either way)."""
    from alsvinn_amd import initial_data as ic
    sod = hip.Ensemble(hip.make_scheme(hip.EQ_EULER1, hip.FLUX_HLL, hip.RECON_NONE, hip.INT_FE,
                                       hip.BC_NEUMANN, cfl=0.9),
                       [hip.make_sample((1024, 1, 1), (-1, 0, 0), (1, 0, 0), 0.2)])
    sod.upload(0, ic.sod(1024))
    sod.apply_boundary()
    t0 = time.perf_counter()
    sod.run(50)  # first solve: includes capturing and instantiating the step graph
    t1 = time.perf_counter()
    info = sod.info(0)
    sod.reset_time()
    sod.upload(0, ic.sod(1024))
    sod.apply_boundary()
    sod.synchronize()
    t2 = time.perf_counter()
    sod.run(50)  # same solve again on the instantiated graph
    t3 = time.perf_counter()
    assert sod.info(0)["steps"] == info["steps"]
    whole_step = sod.small_1d()
    # the same solve with the multi-launch step (two launches per step)
    sod.set_small_1d(0)
    sod.reset_time()
    sod.upload(0, ic.sod(1024))
    sod.apply_boundary()
    sod.run(50)
    sod.reset_time()
    sod.upload(0, ic.sod(1024))
    sod.apply_boundary()
    sod.synchronize()
    t4 = time.perf_counter()
    sod.run(50)
    t5 = time.perf_counter()
    assert sod.info(0)["steps"] == info["steps"]
    sod.close()
    out = {"solve_ms": 1e3 * (t1 - t0), "warm_solve_ms": 1e3 * (t3 - t2),
           "steps": info["steps"], "t": info["t"], "whole_step_launch_threads": whole_step,
           "multi_launch_warm_solve_ms": 1e3 * (t5 - t4),
           "workload": "Sod 1024 cells, HLL, FE, CFL 0.9, T=0.2 (latency-bound)"}
    log("  sod: %d steps in %.2f ms (first solve, with graph capture), %.2f ms (again); "
        "multi-launch step %.2f ms" % (info["steps"], 1e3 * (t1 - t0), 1e3 * (t3 - t2),
                                       1e3 * (t5 - t4)))
    return out


def main():
    args = parse()
    host_cpu()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(sample sharding across GPUs is a design note, not code). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    import alsvinn_amd as hip
    from alsvinn_amd import sampling

    if args.only == "serving":
        print(json.dumps(serving_bench(args)), flush=True)
        return

    if args.only == "full-uq":
        # the whole Monte Carlo run: every sample to T = 2, then the statistics
        M, n = args.samples, args.n
        X = sampling.stl_mersenne_uniform(M, 40)
        t0 = time.perf_counter()
        e = make_batch(hip, [n] * M, args.precision, X)
        t1 = time.perf_counter()
        sched = e.autotune()
        t2 = time.perf_counter()
        e.run(50)
        t3 = time.perf_counter()
        mean, var = e.meanvar()
        sf = e.structure_cube(33, 2.0)
        t4 = time.perf_counter()
        steps = np.array([e.info(i)["steps"] for i in range(M)])
        e.close()
        updates = float(n * n * steps.sum())
        print(json.dumps({"workload": "full UQ run KH %d^2 x %d to T=2, %s" % (n, M, args.precision),
                          "setup_s": t1 - t0, "autotune_s": t2 - t1, "solve_s": t3 - t2,
                          "statistics_s": t4 - t3,
                          "steps_min_max": [int(steps.min()), int(steps.max())],
                          "cell_updates_per_s": updates / (t3 - t2),
                          "mean_rho_range": [float(mean[0].min()), float(mean[0].max())],
                          "var_rho_max": float(var[0].max()),
                          "structure_cube_rho_h1_h32": [float(sf[0, 1]), float(sf[0, 32])],
                          "schedule": sched}), flush=True)
        return

    if args.only == "stats":
        # the statistics kernels on the headline batch, after a few steps;
        # each statistic called three times (the first pays the allocations)
        M, n = args.samples, args.n
        X = sampling.stl_mersenne_uniform(M, 40)
        e = make_batch(hip, [n] * M, args.precision, X)
        e.run_steps(PER_GRAPH, PER_GRAPH)
        e.synchronize()
        calls = [("meanvar", lambda: e.meanvar()),
                 ("structure_cube_h33_p2", lambda: e.structure_cube(33, 2.0)),
                 ("structure_basic_h33_x", lambda: e.structure_basic(33, 0, 2.0)),
                 ("onepointmoment_p2", lambda: e.onepoint_moment(2)),
                 ("onepointmoment_p3", lambda: e.onepoint_moment(3)),
                 ("bv_p1", lambda: e.bounded_variation(1))]
        out = {"workload": "statistics of KH %d^2 x %d, %s (device kernels + host copy of the "
                           "result; ms per call, three calls)" % (n, M, args.precision)}
        for name, f in calls:
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                f()
                ts.append(1e3 * (time.perf_counter() - t0))
            out[name + "_ms"] = ts
            log("  %s: %s ms" % (name, ", ".join("%.2f" % t for t in ts)))
        e.close()
        print(json.dumps(out), flush=True)
        return

    if args.only == "sod":
        print(json.dumps(sod_solve(hip)), flush=True)
        return

    if args.only == "single":
        # one KH solve of 1024^2 (the plain GPU path's large configuration)
        # on its own, with the stage kernel's launch time from HIP events
        X1 = sampling.stl_mersenne_uniform(1, 40)
        es = make_batch(hip, [1024], args.precision, X1)
        ssch = es.autotune()
        sv, sms, sk = throughput(es, 1024 * 1024, args.steps, warm=args.warmup)
        kms, launches = es.step_timed(10)
        es.close()
        print(json.dumps({"workload": "KH 2D 1024^2 single solve, %s" % args.precision,
                          "value": sv, "unit": "cell-updates/s", "ms_per_step": sms, "steps": sk,
                          "schedule": ssch, "stage_kernel_avg_ms": kms[0] / launches[0]}),
              flush=True)
        return

    if args.only == "kh3d":
        # 3D KH ensemble (band in y, 10 x 10 modes in x and z) on its own,
        # with the 3D stage kernel's launch time from HIP events
        M3, n3 = args.samples3d, args.n3d
        X3 = sampling.stl_mersenne_uniform(M3, 400)
        e3 = make_batch(hip, [n3] * M3, args.precision, X3, dim=3)
        s3 = e3.autotune()
        v3, ms3, k3 = throughput(e3, M3 * n3 ** 3, args.steps, warm=args.warmup)
        kms, launches = e3.step_timed(3)
        e3.close()
        esz = 8 if args.precision == "fp64" else 4
        per_launch = 3 * 5 * esz * M3 * n3 ** 3
        stage_ms = kms[0] / launches[0]
        print(json.dumps({"workload": "KH 3D %d^3 x %d samples, %s state" % (n3, M3, args.precision),
                          "value": v3, "unit": "cell-updates/s", "ms_per_step": ms3, "steps": k3,
                          "schedule": s3, "bytes_per_cell_update": 9 * 5 * esz,
                          "hbm_fraction": v3 * 9 * 5 * esz / HBM_PEAK,
                          "stage_kernel_avg_ms": stage_ms,
                          "stage_kernel_algorithmic_GBps": per_launch / (stage_ms * 1e-3) / 1e9}),
              flush=True)
        return

    if args.only == "fe-calibration":
        # forward-Euler batch: its elementwise kernel reads two volumes and
        # writes one (known bytes), used to calibrate the HBM byte counters
        from alsvinn_amd import initial_data as ic
        X = sampling.stl_mersenne_uniform(64, 40)
        sc = hip.make_scheme(hip.EQ_EULER2, hip.FLUX_HLL3, hip.RECON_NONE, hip.INT_FE,
                             hip.BC_PERIODIC)
        e = hip.Ensemble(sc, [hip.make_sample((256, 256, 1), (0, 0, 0), (1, 1, 0), 2.0)
                              for _ in range(64)])
        for i in range(64):
            u = ic.kelvin_helmholtz_2d(256, X[i], ghosts=1)
            e.upload(i, u)
        e.apply_boundary()
        e.run_steps(10, 10)
        e.synchronize()
        print(json.dumps({"calibration": "axpy_kernel reads 2 and writes 1 volume of "
                          "64 x 258^2 x 4 doubles per launch",
                          "bytes_read": 2 * 64 * 258 * 258 * 4 * 8,
                          "bytes_written": 64 * 258 * 258 * 4 * 8}))
        return

    M, n = args.samples, args.n
    esize = 8 if args.precision == "fp64" else 4
    steps = max(1, args.steps)
    warm = max(0, args.warmup)
    log("generating %d samples of %d^2 (reference sampler draws)" % (M, n))
    X = sampling.stl_mersenne_uniform(M, 40)
    e = make_batch(hip, [n] * M, args.precision, X)
    sched = e.autotune()
    log("autotuned schedule: %s" % sched)
    run_exact(e, warm)
    e.synchronize()
    log("timed %d steps" % steps)
    secs = timed_exact(e, steps)
    cells = M * n * n
    value = cells * steps / secs
    ms_per_step = 1e3 * secs / steps
    log("batched: %.3f ms/step, %.3e cell-updates/s" % (ms_per_step, value))

    # dominant kernel: the fused stage kernel, HIP events around each launch
    # of 6 more steps right after the timed window (the clock drops under
    # sustained FP64 load, so they are timed before the sustained window)
    kms, launches = e.step_timed(6)
    stage_ms = kms[0] / launches[0]
    bytes_per_stage = 3 * 4 * esize * cells  # read stage input, read U^n, write result
    achieved = bytes_per_stage / (stage_ms * 1e-3)
    log("stage kernel avg %.3f ms, %.1f GB/s algorithmic" % (stage_ms, achieved / 1e9))
    sustained = None
    if args.sustained_steps > 0:
        # the same batch continued for a longer window: the GPU's clock drops
        # under sustained FP64 load, so a short window overstates the rate
        ks = args.sustained_steps
        ssecs = timed_exact(e, ks)
        sustained = {"value": cells * ks / ssecs, "unit": "cell-updates/s", "steps": ks,
                     "ms_per_step": 1e3 * ssecs / ks,
                     "hbm_fraction": cells * ks / ssecs * 3 * 3 * 4 * esize / HBM_PEAK,
                     "note": "the %d steps right after the %d timed steps, same batch and "
                             "schedule" % (ks, steps)}
        log("sustained (%d more steps): %.3f ms/step, %.3e cell-updates/s"
            % (ks, sustained["ms_per_step"], sustained["value"]))

    extra = {}
    if args.only == "all":
        t0 = time.perf_counter()
        mean, var = e.meanvar()
        t1 = time.perf_counter()
        sf = e.structure_cube(33, 2.0)
        t2 = time.perf_counter()
        e.onepoint_moment(3)
        t3 = time.perf_counter()
        e.bounded_variation(1)
        t4 = time.perf_counter()
        extra["statistics"] = {"meanvar_ms": 1e3 * (t1 - t0), "structure_cube_h32_ms": 1e3 * (t2 - t1),
                               "onepointmoment_p3_ms": 1e3 * (t3 - t2), "bv_p1_ms": 1e3 * (t4 - t3),
                               "samples": M, "note": "device kernels + host copy of the result"}
        log("statistics: meanvar %.1f ms, structure_cube(H=33) %.1f ms, m3 %.1f ms, bv_1 %.1f ms"
            % (1e3 * (t1 - t0), 1e3 * (t2 - t1), 1e3 * (t3 - t2), 1e3 * (t4 - t3)))
    e.close()

    plain = None
    if not args.no_plain:
        ep = make_batch(hip, [n], args.precision, X[:1])
        psched = ep.autotune()
        pval, pms, ps = throughput(ep, n * n, args.plain_steps)
        ep.close()
        plain = {"value": pval, "unit": "cell-updates/s", "ms_per_step": pms,
                 "workload": "one sample at a time (batch of 1), same scheme and grid",
                 "schedule": psched, "batched_over_plain": value / pval}
        log("plain: %.3e cell-updates/s, batched/plain = %.1fx" % (pval, value / pval))

    if args.only == "all":
        K = args.extra_steps
        log("fp32 variant")
        extra["fp32_state"] = fp32_error(hip, M, n, X, K)
        log("  %s" % extra["fp32_state"])
        log("mixed sizes")
        sizes = [128] * 171 + [256] * 171 + [512] * 170
        Xm = sampling.stl_mersenne_uniform(len(sizes), 40)
        em = make_batch(hip, sizes, "fp64", Xm)
        msch = em.autotune()
        cm = sum(s * s for s in sizes)
        mv, mms, mk = throughput(em, cm, K)
        em.close()
        extra["mixed_sizes"] = {"value": mv, "unit": "cell-updates/s", "ms_per_step": mms,
                                "steps": mk, "schedule": msch, "hbm_fraction": mv * 288 / HBM_PEAK,
                                "workload": "KH 2D FP64, 512 samples: 171x128^2, 171x256^2, 170x512^2"}
        log("  mixed: %.3e cell-updates/s" % mv)
        log("single 1024^2")
        es = make_batch(hip, [1024], "fp64", X[:1])
        ssch = es.autotune()
        sv, sms, sk = throughput(es, 1024 * 1024, max(K, 60))
        es.close()
        extra["single_1024"] = {"value": sv, "unit": "cell-updates/s", "ms_per_step": sms,
                                "steps": sk, "schedule": ssch, "hbm_fraction": sv * 288 / HBM_PEAK,
                                "workload": "KH 2D 1024^2 single solve, FP64 (plain GPU path)"}
        log("  single 1024^2: %.3e cell-updates/s" % sv)
        log("3D fp32 256^3 x 64")
        X3 = sampling.stl_mersenne_uniform(64, 400)
        e3 = make_batch(hip, [256] * 64, "fp32", X3, dim=3)
        s3 = e3.autotune()
        v3, ms3, k3 = throughput(e3, 64 * 256 ** 3, 10, warm=10)
        e3.close()
        extra["kh3d_fp32"] = {"value": v3, "unit": "cell-updates/s", "ms_per_step": ms3,
                              "steps": k3, "schedule": s3, "bytes_per_cell_update": 180,
                              "hbm_fraction": v3 * 180 / HBM_PEAK,
                              "workload": "KH 3D 256^3 x 64 samples, FP32 state"}
        log("  3d: %.3e cell-updates/s" % v3)
        log("sod")
        extra["sod"] = sod_solve(hip)

    if args.only == "all":
        log("fp32 errors per configuration")
        extra["fp32_vs_fp64"] = fp32_errors(hip)
        log("  %s" % extra["fp32_vs_fp64"])

    if args.only == "all" and args.serving_requests > 0:
        log("serving")
        extra["serving"] = serving_bench(args)
        log("  %s" % extra["serving"])

    cpu = None
    if not args.no_cpu and args.only == "all":
        probe = cpu_baseline(n, X[0], 4)
        steps_cpu = int(max(4, min(4000, args.cpu_seconds * probe["value"] / (n * n))))
        cpu = cpu_baseline(n, X[0], steps_cpu)
        log("cpu baseline: %.3e cell-updates/s on %d threads" % (cpu["value"], cpu["cores"]))

    B = 3 * 3 * 4 * esize
    traffic = None
    prof = os.path.join(ROOT, "profiles", "stage_kernel_kh256x1024_fp64.md")
    if args.precision == "fp64" and M == 1024 and n == 256 and os.path.exists(prof):
        # HBM bytes per stage launch from the calibrated FETCH_SIZE/WRITE_SIZE
        # runs, for the band height the autotuner chose when it was profiled
        by_rows = {}
        for ln in open(prof):
            if ln.startswith("traffic_bytes_per_launch_rows"):
                key, val = ln.split(":")
                by_rows[int(key[len("traffic_bytes_per_launch_rows"):])] = float(val)
            elif ln.startswith("traffic_bytes_per_launch:"):
                traffic = float(ln.split(":")[1])
        traffic = by_rows.get(sched.get("rows"), traffic)
    line = {
        "metric": "cell-updates/s + % HBM peak, batched vs plain (1 GPU); p50/p99 UQ-request latency",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (Kelvin-Helmholtz initial data from the reference sampler's draws)",
        "config": {"workload": "batched Monte Carlo KH 2D Euler %d^2 x %d samples, WENO2+HLL3, "
                               "SSP-RK3, CFL 0.475, periodic" % (n, M),
                   "samples": M, "grid": [n, n], "cells_per_step": cells,
                   "bytes_per_cell_update": B,
                   "hbm_fraction": value * B / HBM_PEAK, "schedule": sched},
        "sustained": sustained,
        "plain": plain,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "algorithmic_bytes_per_launch": bytes_per_stage,
                     "kernel": "stage_kernel_12<{t}, EulerEq<{t},2>, ReconWENO2, FluxHLL3, 2>".format(
                         t="double" if args.precision == "fp64" else "float"),
                     "avg_launch_ms": stage_ms, "launches": launches[0]},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
