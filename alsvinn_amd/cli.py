"""Command-line runner for the reference's configuration files on the hip
platform: the counterpart of alsvinncli (one solve, <fvm> only) and alsuqcli
(all samples of <uq>), alsvinncli/src/main.cpp and alsuqcli/src/main.cpp.

    python -m alsvinn_amd.cli config.xml [--output-dir DIR] [--precision fp64|fp32]

All samples advance together in one ensemble on the GPU. Writers and
statistics run at the reference's save times: the ensemble holds every sample
once its save counter reaches the next save (the state the reference's
FixedIntervalWriter / FixedIntervalStatistics see after that step), the host
writes the snapshots or accumulates the statistics, and the run continues.
Output files are named as the reference names them: the writer's basename, and
<basename>_mean, <basename>_variance, <basename>_structure_cube_<p>,
<basename>_m<p> and <basename>_bv_<p> for the statistics (Setup.cpp:277-295,
StructureCube.cpp:27, OnePointMoment.cpp:21-23, BoundedVariation.cpp:23-26).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import _lib, netcdf
from .config import ConfigError, conserved_volume, read_config, sample_parameters
from .ensemble import Ensemble, make_sample

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _ghosts(scheme):
    return _lib.load().alsvinn_hip_ghost_cells(scheme)


class _StatsOutput:
    """One statistic of the configuration: its files and save axis."""

    def __init__(self, stat, cfg, nvar, out_dir, nsamples):
        self.stat = stat
        self.files = {}
        times = netcdf.save_fractions(stat.number_of_saves, stat.write_initial)
        base = os.path.join(out_dir, stat.writer.basename)
        names = netcdf.variable_names(nvar, 3 if cfg.dims[2] > 1 else 2)
        if stat.name == "meanvar":
            mids = netcdf.midpoints(cfg.lower, cfg.upper, cfg.dims)
            for key in ("mean", "variance"):
                self.files[key] = netcdf.EnsembleWriter(base + "_" + key, cfg.dims, mids, nsamples,
                                                        times, names)
        elif stat.name == "onepointmoment":
            key = "m%d" % int(stat.params["p"])
            self.files[key] = netcdf.EnsembleWriter(
                base + "_" + key, cfg.dims, netcdf.midpoints(cfg.lower, cfg.upper, cfg.dims),
                nsamples, times, names)
        elif stat.name == "bv":
            # one value per variable: the reference's 1x1x1 snapshot volume
            key = "bv_%d" % int(stat.params["p"])
            self.files[key] = netcdf.EnsembleWriter(base + "_" + key, (1, 1, 1),
                                                    [np.zeros(1)] * 3, nsamples, times, names)
        else:
            p = stat.params.get("p", 2.0)
            H = int(stat.params.get("numberOfH", 1))
            h = [np.arange(H, dtype=float), np.zeros(1), np.zeros(1)]
            key = "%s_%f" % (stat.name, p)
            self.files[key] = netcdf.EnsembleWriter(base + "_" + key, (H, 1, 1), h, nsamples, times,
                                                    names)

    def compute(self, e, save_index, results):
        st = self.stat
        if st.name == "meanvar":
            mean, var = e.meanvar()
            g = _grid_ghosts(mean.shape, e)
            self.files["mean"].write(0, save_index, mean, g)
            self.files["variance"].write(0, save_index, var, g)
            results.setdefault("meanvar", []).append((mean, var))
        elif st.name == "onepointmoment":
            key = "m%d" % int(st.params["p"])
            m = e.onepoint_moment(int(st.params["p"]))
            self.files[key].write(0, save_index, m, _grid_ghosts(m.shape, e))
            results.setdefault(key, []).append(m)
        elif st.name == "bv":
            key = "bv_%d" % int(st.params["p"])
            bv = e.bounded_variation(int(st.params["p"]))
            self.files[key].write(0, save_index, bv.reshape(-1, 1, 1, 1), (0, 0, 0))
            results.setdefault(key, []).append(bv)
        else:
            p = st.params.get("p", 2.0)
            H = int(st.params.get("numberOfH", 1))
            if st.name == "structure_cube":
                s = e.structure_cube(H, p)
            else:
                s = e.structure_basic(H, int(st.params.get("direction", 0)), p)
            key = "%s_%f" % (st.name, p)
            self.files[key].write(0, save_index, s.reshape(s.shape[0], 1, 1, H), (0, 0, 0))
            results.setdefault(key, []).append(s)


def _grid_ghosts(shape, e):
    dims6 = e.layout(0)[0]
    return dims6[3], dims6[4], dims6[5]


def run(cfg, out_dir=".", precision=_lib.FP64, log=None):
    """Runs a configuration; returns a summary dict (timings, step counts and
    the statistics computed)."""
    log = log or (lambda *a: None)
    os.makedirs(out_dir, exist_ok=True)
    scheme = cfg.scheme(precision)
    G = _ghosts(scheme)
    params = sample_parameters(cfg) if cfg.has_uq else [{}]
    n = len(params)
    # adjusters: the writer's FixedIntervalWriter first, then the statistics'
    adj = []
    if cfg.writer is not None:
        adj.append(("writer", cfg.writer.number_of_saves, cfg.writer.write_initial))
    stat_saves = sorted({s.number_of_saves for s in cfg.stats})
    if len(stat_saves) > 1:
        raise ConfigError("statistics with different numberOfSaves are not available on the hip "
                          "platform (one save schedule for all statistics)")
    if stat_saves:
        adj.append(("stats", stat_saves[0], any(s.write_initial for s in cfg.stats)))
    samples = [make_sample(cfg.dims, cfg.lower, cfg.upper, cfg.end_time,
                           saves=tuple(a[1] for a in adj)) for _ in range(n)]
    t0 = time.perf_counter()
    e = Ensemble(scheme, samples)
    nvar = e.nvar
    try:
        for i, p in enumerate(params):
            e.upload(i, conserved_volume(cfg, G, p))
        e.apply_boundary()
        e.synchronize()
        t1 = time.perf_counter()
        log("initial data for %d sample(s): %.2f s" % (n, t1 - t0))
        writer = None
        if cfg.writer is not None:
            writer = netcdf.EnsembleWriter(
                os.path.join(out_dir, cfg.writer.basename), cfg.dims,
                netcdf.midpoints(cfg.lower, cfg.upper, cfg.dims), n,
                netcdf.save_fractions(cfg.writer.number_of_saves, cfg.writer.write_initial),
                netcdf.variable_names(nvar, 3 if cfg.dims[2] > 1 else 2))
        stats = [_StatsOutput(s, cfg, nvar, out_dir, n) for s in cfg.stats]
        results = {}
        ghosts = _grid_ghosts(None, e)
        index = {}
        for k, (kind, _, initial) in enumerate(adj):
            index[kind] = 0
            if initial:
                _event(kind, e, n, writer, stats, 0, ghosts, results)
                index[kind] = 1
        limits = [2] * len(adj)  # save counters start at 1; the first save makes them 2
        for k in range(len(adj)):
            e.set_hold(k, limits[k])
        e.autotune()
        while True:
            e.run()
            saved = [e.saves(i) for i in range(n)]
            for k, (kind, _, _) in enumerate(adj):
                if all(s[k] >= limits[k] for s in saved):
                    _event(kind, e, n, writer, stats, index[kind], ghosts, results)
                    index[kind] += 1
                    limits[k] += 1
                    e.set_hold(k, limits[k])
            if not any(e.info(i)["active"] for i in range(n)):
                break
        t2 = time.perf_counter()
        steps = [e.info(i)["steps"] for i in range(n)]
        cells = int(np.prod(cfg.dims))
        summary = {"name": cfg.name, "samples": n, "steps_min_max": [min(steps), max(steps)],
                   "end_time": [e.info(i)["t"] for i in range(min(n, 4))],
                   "setup_s": t1 - t0, "solve_s": t2 - t1,
                   "cell_updates_per_s": cells * sum(steps) / max(t2 - t1, 1e-12),
                   "schedule": e.schedule(), "results": results}
        log("solved %d sample(s) in %.2f s" % (n, t2 - t1))
        return summary
    finally:
        e.close()


def _event(kind, e, n, writer, stats, save_index, ghosts, results):
    if kind == "writer" and writer is not None:
        for i in range(n):
            writer.write(i, save_index, e.download(i), ghosts)
    elif kind == "stats":
        for s in stats:
            s.compute(e, save_index, results)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("config")
    ap.add_argument("--output-dir", default=".")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    args = ap.parse_args(argv)
    cfg = read_config(args.config)
    summary = run(cfg, args.output_dir, _lib.FP64 if args.precision == "fp64" else _lib.FP32,
                  log=lambda *a: print(*a, file=sys.stderr, flush=True))
    summary.pop("results")
    print(json.dumps(summary))
    return 0


if __name__ == "__main__":
    sys.exit(main())
