"""Benchmark: batched Monte Carlo Kelvin-Helmholtz ensemble on one MI355X.

Workload (one "step" = one SSP-RK3 timestep of every sample of the batch):
  2D Euler Kelvin-Helmholtz, 256^2 cells per sample on [0,1]^2, periodic,
  gamma 1.4, WENO2 + HLL3, SSP-RK3, CFL 0.475, FP64 state, 1024 samples
  advanced together; each sample's 40 perturbation parameters are the
  reference sampler's (global std::mt19937_64, uniform[0,1]) draws for that
  sample. Inputs are resident in HBM before the timed region starts.

Reported: cell-updates/s of the batch (value), the same configuration run one
sample at a time on the GPU (plain path), the roofline of the dominant kernel
(the fused stage kernel, timed with HIP events on the ensemble's stream) and
a CPU baseline (the oracle on the host cores, a bounded sample).

  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
NVAR = 4
STAGES = 3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--samples", type=int, default=1024)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--plain-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-plain", action="store_true")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_batch(hip, nsamples, n, precision, X):
    from alsvinn_amd import initial_data as ic
    scheme = hip.make_scheme(hip.EQ_EULER2, hip.FLUX_HLL3, hip.RECON_WENO2, hip.INT_RK3,
                             hip.BC_PERIODIC, gamma=1.4, cfl=0.475,
                             precision=hip.FP64 if precision == "fp64" else hip.FP32)
    samples = [hip.make_sample((n, n, 1), (0, 0, 0), (1, 1, 0), 2.0) for _ in range(nsamples)]
    e = hip.Ensemble(scheme, samples)
    for i in range(nsamples):
        e.upload(i, ic.kelvin_helmholtz_2d(n, X[i]))
    e.apply_boundary()
    e.synchronize()
    return e


def timed_steps(e, steps, per_graph):
    e.synchronize()
    t0 = time.perf_counter()
    e.run_steps(steps, per_graph)
    e.synchronize()
    return time.perf_counter() - t0


def cpu_baseline(n, X0, steps):
    """The oracle (clean-room restatement of the reference CPU path: OpenMP
    over rows, separate passes) on the host cores, one sample, `steps` steps."""
    from tests import oracle_lib as O
    from alsvinn_amd import initial_data as ic
    s = O.make_setup(O.EQ_EULER2, O.FLUX_HLL3, O.RECON_WENO2, O.INT_RK3, O.BC_PERIODIC,
                     (n, n, 1), (0, 0, 0), (1, 1, 0), cfl=0.475, end_time=2.0, max_steps=steps)
    u0 = ic.kelvin_helmholtz_2d(n, X0)
    t0 = time.perf_counter()
    rc, out, t, st, _ = O.run(s, u0)
    dt = time.perf_counter() - t0
    assert rc == 0 and st == steps
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": n * n * steps / dt, "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": "oracle (C++ restatement of the reference cpu platform, OpenMP, -O3 "
                      "-ffp-contract=off), 1 sample of the same 256^2 KH WENO2+HLL3 SSP-RK3 "
                      "FP64 workload, %d steps, %.1f s" % (steps, dt)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(sample sharding across GPUs is a design note, not code). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    import alsvinn_amd as hip
    from alsvinn_amd import sampling

    M, n = args.samples, args.n
    per_graph = 10
    steps = max(per_graph, (args.steps // per_graph) * per_graph)
    warm = max(per_graph, (args.warmup // per_graph) * per_graph)
    log("generating %d samples of %d^2 (reference sampler draws)" % (M, n))
    X = sampling.stl_mersenne_uniform(M, 40)
    e = make_batch(hip, M, n, args.precision, X)
    sched = e.autotune()
    log("autotuned schedule: %s" % sched)
    log("warmup %d steps" % warm)
    e.run_steps(warm, per_graph)
    e.synchronize()
    log("timed %d steps" % steps)
    secs = timed_steps(e, steps, per_graph)
    cells = M * n * n
    value = cells * steps / secs
    ms_per_step = 1e3 * secs / steps
    log("batched: %.3f ms/step, %.3e cell-updates/s" % (ms_per_step, value))

    # dominant kernel: fused stage kernel, HIP events around each launch
    kms, launches = e.step_timed(6)
    stage_ms = kms[0] / launches[0]
    esize = 8 if args.precision == "fp64" else 4
    bytes_per_stage = 3 * NVAR * esize * cells  # read stage input, read U^n, write result
    achieved = bytes_per_stage / (stage_ms * 1e-3)
    log("stage kernel avg %.3f ms, %.1f GB/s algorithmic" % (stage_ms, achieved / 1e9))
    info0 = e.info(0)
    e.close()

    plain = None
    if not args.no_plain:
        ep = make_batch(hip, 1, n, args.precision, X[:1])
        psched = ep.autotune()
        log("plain schedule: %s" % psched)
        ep.run_steps(per_graph, per_graph)
        ep.synchronize()
        ps = max(per_graph, (args.plain_steps // per_graph) * per_graph)
        psecs = timed_steps(ep, ps, per_graph)
        ep.close()
        pval = n * n * ps / psecs
        plain = {"value": pval, "unit": "cell-updates/s", "ms_per_step": 1e3 * psecs / ps,
                 "workload": "one sample at a time (batch of 1), same scheme and grid",
                 "batched_over_plain": value / pval}
        log("plain: %.3e cell-updates/s, batched/plain = %.1fx" % (pval, value / pval))

    cpu = None
    if not args.no_cpu:
        probe = cpu_baseline(n, X[0], 4)
        steps_cpu = int(max(4, min(4000, args.cpu_seconds * probe["value"] / (n * n))))
        cpu = cpu_baseline(n, X[0], steps_cpu)
        log("cpu baseline: %.3e cell-updates/s on %d threads" % (cpu["value"], cpu["cores"]))

    B = STAGES * 3 * NVAR * esize
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
        "plain": plain,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel": "stage_kernel_12<double, Euler2, WENO2, HLL3>",
                     "avg_launch_ms": stage_ms, "launches": launches[0]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
