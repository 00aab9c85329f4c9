"""Write golden vectors for the semi-implicit step under the reference's open-boundary tide adapter.

Runs only in a container that has the reference checkout at /root/reference
(read-only; bytecode writing is disabled before import), with the stubs and
float64 factory wrapping of make_golden.py and the production construction of
make_golden_implicit.py.  The double-gyre basin is opened into a channel: the
west and east edge columns take the depth of their interior neighbours (wet),
the north and south edges stay land, and the reference's own
``areas_masks`` (src/neuralpom/cores/explicit/grid.py:113-166) recomputes
areas and masks.  The tide adapter
src/neuralpom/cores/case_adapters/coastal_boundary.py is installed through its
own ``install`` (implicit family, :60-67) from a node file written to a
temporary directory: nodes on the first interior column (0-based i = 1) at
every interior j, amplitude 0.3 + 0.02 j m, phase 15 j degrees.  The exterior
tracer values tbe tbw sbe sbw tbn tbs sbn sbs (base.py:181-189) are set to
seeded values so that inflow edges read them.

Output ``tests/golden/implicit_tide_<im>x<jm>.npz``, all float64:
  in/<field>, scalar/<name>, tide/{i, j, amp, phase}   input state and tide nodes
  a/<field>, b/<field>      state after 1 and 3 steps from time0 = 0.1 d,
  a/iters, b/iters          the adapter's CG iteration count per step
  cot/<name>, grad/<name>   seeded-cotangent vector-Jacobian product of one
                            step (torch.autograd through the reference step,
                            the adapter's CG included)
"""
from __future__ import annotations

import os
import sys
import tempfile
from pathlib import Path

import numpy as np
import torch

sys.path.insert(0, str(Path(__file__).resolve().parent))
import make_golden as G              # noqa: E402
import make_golden_implicit as GI    # noqa: E402

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
    core._accel_cache_ready = False


def build(im, jm, seed, nodes):
    from neuralpom.cores.case_adapters import coastal_boundary as CB

    core = GI.build_core(im, jm)
    open_channel(core)
    G.perturb(core, seed)
    with tempfile.TemporaryDirectory(dir=OUT) as tmp:
        path = Path(tmp) / "tide_nodes.dat"
        np.savetxt(path, nodes)
        CB.install(core, "implicit", str(path))
    g = torch.Generator().manual_seed(seed + 11)
    for n in EXTERIOR:
        x = getattr(core, n)
        base = 10.0 if n.startswith("t") else 34.0
        setattr(core, n, base + torch.randn(x.shape, generator=g, dtype=torch.float64))
    core.time0 = core.time = 0.1
    history = []
    solve = core.solve_implicit_eta_cg

    def counted(*args, **kwargs):
        out = solve(*args, **kwargs)
        history.append(int(core._validation_eta_iterations))
        return out

    core.solve_implicit_eta_cg = counted
    return core, history


def main(im=14, jm=12, seed=4321):
    G._install_stubs()
    import neuralpom.cores.implicit_p0_fixed  # noqa: F401
    import neuralpom.hybrid.core_pool  # noqa: F401
    import neuralpom.v1.core_spec  # noqa: F401
    import neuralpom.cores.case_adapters.coastal_boundary  # noqa: F401
    originals = G._force_fp64()
    torch.set_default_dtype(torch.float64)
    try:
        tag = f"{im}x{jm}"
        js = np.arange(1, jm - 1)
        nodes = np.stack([np.full(js.shape, 2.0), js + 1.0, 0.3 + 0.02 * js, 15.0 * js], axis=1)
        out = {"tide/i": nodes[:, 0] - 1, "tide/j": nodes[:, 1] - 1, "tide/amp": nodes[:, 2],
               "tide/phase": nodes[:, 3]}
        for label, nstep in (("a", 1), ("b", 3)):
            core, history = build(im, jm, seed, nodes)
            if label == "a":
                for k, v in G._tensors(core).items():
                    out[f"in/{k}"] = G._np(v)
                out.update(GI.scalars(core))
            core.iint = 0
            core.iend = nstep
            with torch.no_grad():
                core.run_time_loop()
            for k, v in G._tensors(core).items():
                out[f"{label}/{k}"] = G._np(v)
            out[f"{label}/iters"] = np.array(history, dtype=np.int64)

        core, _ = build(im, jm, seed, nodes)
        leaves = {}
        for name in GI.VJP_INPUTS:
            leaf = getattr(core, name).detach().clone().requires_grad_(True)
            leaves[name] = leaf
            setattr(core, name, leaf.clone())
        core.iint = 0
        core.iend = 1
        core.run_time_loop()
        g = torch.Generator().manual_seed(seed + 1)
        loss = 0.0
        for name in GI.VJP_OUTPUTS:
            o = getattr(core, name)
            cot = torch.randn(o.shape, generator=g, dtype=torch.float64)
            out[f"cot/{name}"] = G._np(cot)
            loss = loss + (o * cot).sum()
        grads = torch.autograd.grad(loss, [leaves[n] for n in GI.VJP_INPUTS], allow_unused=True)
        for name, gr in zip(GI.VJP_INPUTS, grads):
            out[f"grad/{name}"] = G._np(gr) if gr is not None else np.zeros(leaves[name].shape)
        np.savez_compressed(OUT / f"implicit_tide_{tag}.npz", **out)
    finally:
        for name, fn in originals.items():
            setattr(torch, name, fn)


if __name__ == "__main__":
    main(int(os.environ.get("GOLDEN_IM", "14")), int(os.environ.get("GOLDEN_JM", "12")))
