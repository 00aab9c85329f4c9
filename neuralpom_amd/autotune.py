"""Tune the launch shapes of the point kernels (pom_batch_tune) per workload
and write the table the ensembles read (tuning.json beside this module).

    python -m neuralpom_amd.autotune [--out PATH] [--workloads a,b,...]

Workloads (fp32, forward + adjoint per step):
  seamount64  64 explicit members (32 x 90x82, 24 x 120x110, 8 x 177x161),
              the step replayed from a HIP graph;
  shelf1      one explicit 90x82 member, replayed from a HIP graph;
  channel16   16 semi-implicit 354x322 members of the tidal channel, eager.

For each workload the shapes are swept one group at a time (the barotropic
substep's k_map2, then the 2-D, then the 3-D maps and gathers), keeping the
best of each group (a candidate must beat the current best by more than 1 %,
the run-to-run noise of a step); a step's time is the median of per-step
HIP-event times.  The results do not depend on the shapes (each point's arithmetic is
the same), so the tuner only times."""
from __future__ import annotations

import argparse
import json
import statistics
import time
from pathlib import Path

import torch

from . import _lib as L
from . import tuning as TU

MAP2 = [(1, 768), (2, 768), (4, 768), (8, 768), (4, 512), (4, 1024), (4, 1536)]
FLAT2 = [(1, 768), (2, 768), (4, 768), (2, 1024), (4, 1536)]
FLAT3 = [(1, 768), (2, 1536), (2, 3072), (4, 3072)]
DIFF_OUT = ("u", "v", "t", "s", "el", "ua", "va", "q2", "km", "aam", "rho")
MARGIN = 0.01   # a shape replaces the current best only when faster by more than this (step-time noise)


def _seamount(shapes, dev):
    from . import cases
    from .explicit import Ensemble, StepScalars
    ens = Ensemble(shapes, dtype=torch.float32, device=dev)
    members = [cases.seamount(im, jm, seed=1000 + m, dtype=torch.float32, device=dev)
               for m, (im, jm) in enumerate(shapes)]
    state = {n: ens.pack([mb[0][n] for mb in members], n) for n in L.FIELDS}
    geom = {n: ens.pack([mb[1][n] for mb in members], n) for n in L.GEOM}
    vg = torch.stack([mb[2] for mb in members]).to(dev, torch.float32).contiguous()
    ramp = torch.tensor([min(1.0 / mb[3], 1.0) for mb in members], dtype=torch.float32, device=dev)
    sc = StepScalars(dte=6.0, isplit=30, alpha=0.225, smoth=0.10, horcon=0.05, hmax=4500.0)
    return ens, state, geom, vg, ramp, sc


def _graph_timer(shapes, dev, steps):
    from .explicit import StepGraph
    ens, state, geom, vg, ramp, sc = _seamount(shapes, dev)
    gen = torch.Generator(device="cpu").manual_seed(7)
    cot = {n: (torch.randn(ens.numel(n), generator=gen) * 1e-3).to(dev) for n in DIFF_OUT}

    def run(tuning):
        ens.set_tuning(tuning)
        G = StepGraph(ens, state, geom, sc, ramp, vg, True, cot)
        for _ in range(2):
            G.replay()
        times = []
        for _ in range(steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            G.replay()
            b.record()
            b.synchronize()
            times.append(a.elapsed_time(b))
        del G
        return statistics.median(times)
    return ens, run


def _channel_timer(dev, steps, members=16):
    from . import boundary as B
    from . import cases
    from .implicit import ImplicitEnsemble
    from .implicit_core import POM2K_Core
    fine = (354, 322)
    cores = []
    for m in range(members):
        st, g, vg, period, nodes = cases.tidal_channel(*fine, seed=100 + m, hmin=5.0, hmax=80.0, dtdz=0.05,
                                                       dtype=torch.float32, device=dev)
        c = POM2K_Core(*fine, device=dev, dtype=torch.float32)
        c.configure_runtime(dte=1.0, isplit=30)
        for n, v in {**st, **g}.items():
            setattr(c, n, v)
        c.z, c.zz, c.dz, c.dzz = vg
        c.period, c.horcon, c.double_gyre_wind_stress_pa = period, 0.05, 0.0
        B.install(c, "implicit", nodes=nodes, speed=360.0 / 12.42, radiation=0.0)
        c.time0 = c.time = 0.0
        cores.append(c)
    ens = ImplicitEnsemble([fine] * members, dtype=torch.float32, device=dev)
    g0 = [c._geom() for c in cores]
    geom = {n: ens.pack([gg[n] for gg in g0], n) for n in g0[0]}
    state = {n: ens.pack([c._flat(n).detach() for c in cores], n) for n in L.FIELDS}
    vg = torch.stack([c._vgrid()[0] for c in cores]).contiguous()
    core0 = cores[0]
    core0.time = core0.dti / 86400.0
    sc = core0._scalars_at_step(core0.step_scalars())
    ramp = torch.full((members,), min(core0.time / float(core0.period), 1.0), dtype=torch.float32, device=dev)
    del cores, g0
    ens.prepare(geom, sc)
    gen = torch.Generator(device="cpu").manual_seed(7)
    cot = {n: (torch.randn(ens.numel(n), generator=gen) * 1e-3).to(dev) for n in DIFF_OUT}

    def one():
        out, _ = ens.step_taped(state, geom, sc, ramp, vg, True)
        ens.step_reverse(state, out, geom, sc, ramp, vg, True, dict(cot))

    def run(tuning):
        ens.set_tuning(tuning)
        one()
        torch.cuda.synchronize()
        times = []
        for _ in range(steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            one()
            b.record()
            b.synchronize()
            times.append(a.elapsed_time(b))
        return statistics.median(times)
    return ens, run


WORKLOADS = {
    "seamount64": lambda dev: _graph_timer([(90, 82)] * 32 + [(120, 110)] * 24 + [(177, 161)] * 8, dev, 11),
    "shelf1": lambda dev: _graph_timer([(90, 82)], dev, 20),
    "channel16": lambda dev: _channel_timer(dev, 5),
}


def tune(name, dev, log):
    ens, run = WORKLOADS[name](dev)
    best = dict(TU.DEFAULT)
    default_ms = run(best)
    best_ms = default_ms
    sweep = [[dict(best), default_ms]]
    log(f"{name}: default {default_ms:.3f} ms")
    for prefix, grid in (("map2", MAP2), ("flat2", FLAT2), ("flat3", FLAT3)):
        if name == "channel16" and prefix == "map2":
            continue   # the semi-implicit step has no k_map2 substep
        for bpg, mn in grid:
            cand = dict(best)
            cand[f"{prefix}_blocks_per_group"], cand[f"{prefix}_min_groups"] = bpg, mn
            if cand == best:
                continue
            ms = run(cand)
            sweep.append([dict(cand), ms])
            log(f"{name}: {prefix} ({bpg}, {mn}) {ms:.3f} ms")
            if ms < best_ms * (1.0 - MARGIN):
                best, best_ms = cand, ms
    key = TU.workload_key(ens.shapes, "f32")
    del ens
    torch.cuda.empty_cache()
    return {"workload": name, "key": key, "tuning": best, "ms_per_step": best_ms, "default_ms": default_ms,
            "sweep": sweep}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(TU.TABLE))
    ap.add_argument("--workloads", default=",".join(WORKLOADS))
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    entries = []
    for name in args.workloads.split(","):
        t0 = time.perf_counter()
        entries.append(tune(name, dev, lambda msg: print(msg, flush=True)))
        print(f"{name}: tuned in {time.perf_counter() - t0:.1f} s -> {entries[-1]['tuning']}", flush=True)
    table = {"device": torch.cuda.get_device_name(dev), "measure": "median HIP-event time of a fwd+adjoint step",
             "entries": entries}
    Path(args.out).write_text(json.dumps(table, indent=1) + "\n")
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
