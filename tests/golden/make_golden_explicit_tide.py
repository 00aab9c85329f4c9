"""Write golden vectors for the explicit step under the reference's open-boundary tide adapter.

Runs only in a container that has the reference checkout at /root/reference
(read-only; bytecode writing is disabled before import), with the stubs and
float64 factory wrapping of make_golden.py and its double-gyre explicit core.
The basin is opened into a channel as in make_golden_tide.py: the west and
east edge columns take the depth of their interior neighbours (wet), the
north and south edges stay land, and the reference's own ``areas_masks``
(src/neuralpom/cores/explicit/grid.py:113-166) recomputes areas and masks.
The tide adapter src/neuralpom/cores/case_adapters/coastal_boundary.py is
installed through its own ``install`` with the explicit family (:55-62), which
replaces ``bcond`` by the adapter's (:11-30: the tide at the nodes on every
substep at a substep counter's time, radiation conditions on uaf / vaf, and
bcond 3-6 of case_adapters/kernels/coastal_explicit_baroclinic.py), from a
node file written to a temporary directory: nodes on the first interior
column (0-based i = 1) at every interior j, amplitude 0.3 + 0.02 j m, phase
15 j degrees.  The exterior tracer values tbe tbw sbe sbw tbn tbs sbn sbs
(base.py:181-189) are set to seeded values so that inflow edges read them.

The substep counter ``_tide_substep`` that sets the tide's time is reset to 0
by ``install`` and is otherwise never reset by the core: a run of several
steps in one ``run_time_loop`` call carries it on (label b), while the case
registry's advancement (case_adapters/registry.py:54-59) runs one step per
call with ``time0`` set to the current time and the counter reset (label c).

Output ``tests/golden/explicit_tide_<im>x<jm>.npz``, all float64:
  in/<field>, scalar/<name>, tide/{i, j, amp, phase}   input state and tide nodes
  a/<field>                 state after 1 step from time0 = 0.1 d
  b/<field>, b/substep      after 3 steps in one run_time_loop call, the counter after
  c/<field>, c/substep      after 3 steps of the registry's one-step protocol
The adapter's explicit family is forward-only (the registry refuses to
advance with gradients enabled, registry.py:39-40), so there is no VJP.
"""
from __future__ import annotations

import os
import sys
import tempfile
from pathlib import Path

import numpy as np
import torch

sys.path.insert(0, str(Path(__file__).resolve().parent))
import make_golden as G   # noqa: E402

OUT = Path(__file__).resolve().parent
EXTERIOR = ("tbe", "tbw", "sbe", "sbw", "tbn", "tbs", "sbn", "sbs")


def open_channel(core):
    h = core.h.clone()
    h[0, 1:-1] = h[1, 1:-1]
    h[-1, 1:-1] = h[-2, 1:-1]
    core.h = h
    core.areas_masks()
    core.d = core.h + core.el
    core.dt = core.h + core.et


def build(im, jm, seed, nodes):
    from neuralpom.cores.case_adapters import coastal_boundary as CB

    core = G.build_core(im, jm)
    open_channel(core)
    G.perturb(core, seed)
    with tempfile.TemporaryDirectory(dir=OUT) as tmp:
        path = Path(tmp) / "tide_nodes.dat"
        np.savetxt(path, nodes)
        CB.install(core, "explicit", str(path))
    g = torch.Generator().manual_seed(seed + 11)
    for n in EXTERIOR:
        x = getattr(core, n)
        base = 10.0 if n.startswith("t") else 34.0
        setattr(core, n, base + torch.randn(x.shape, generator=g, dtype=torch.float64))
    core.time0 = core.time = 0.1
    return core


def main(im=14, jm=12, seed=4321):
    G._install_stubs()
    import neuralpom.cores.explicit  # noqa: F401
    import neuralpom.cores.case_adapters.coastal_boundary  # noqa: F401
    originals = G._force_fp64()
    torch.set_default_dtype(torch.float64)
    try:
        tag = f"{im}x{jm}"
        js = np.arange(1, jm - 1)
        nodes = np.stack([np.full(js.shape, 2.0), js + 1.0, 0.3 + 0.02 * js, 15.0 * js], axis=1)
        out = {"tide/i": nodes[:, 0] - 1, "tide/j": nodes[:, 1] - 1, "tide/amp": nodes[:, 2],
               "tide/phase": nodes[:, 3]}
        core = build(im, jm, seed, nodes)
        for k, v in G._tensors(core).items():
            out[f"in/{k}"] = G._np(v)
        out.update(G.scalars(core))
        for label, nstep in (("a", 1), ("b", 3)):
            core = build(im, jm, seed, nodes)
            core.iint = 0
            core.iend = nstep
            with torch.no_grad():
                core.run_time_loop()
            for k, v in G._tensors(core).items():
                out[f"{label}/{k}"] = G._np(v)
            out[f"{label}/substep"] = np.array(core._tide_substep, dtype=np.int64)
        core = build(im, jm, seed, nodes)
        with torch.no_grad():
            for _ in range(3):   # registry.py:54-59
                core.time0 = float(core.time)
                core.iint = 0
                core.iend = 1
                core._tide_substep = 0
                core.run_time_loop()
        for k, v in G._tensors(core).items():
            out[f"c/{k}"] = G._np(v)
        out["c/substep"] = np.array(core._tide_substep, dtype=np.int64)
        out["c/time"] = np.array(float(core.time))
        np.savez_compressed(OUT / f"explicit_tide_{tag}.npz", **out)
    finally:
        for name, fn in originals.items():
            setattr(torch, name, fn)


if __name__ == "__main__":
    main(int(os.environ.get("GOLDEN_IM", "14")), int(os.environ.get("GOLDEN_JM", "12")))
