"""Write golden vectors for the semi-implicit step with the spatial Shapiro
filters on (barotropic, tracer and internal-velocity weights 0.05, the
reference's NEURALPOM_IMPLICIT_SPATIAL_SHAPIRO value,
src/neuralpom/cores/implicit_p0_fixed/core.py:62-119 and
numerics/double_gyre.py:78-184), from the reference.

Runs only in a container that has the reference checkout at /root/reference,
with the stubs and float64 factories of make_golden.py and the construction
of make_golden_implicit.py.  Output ``tests/golden/implicit_shapiro_12x10.npz``
(float64): the perturbed input state ("in/"), the state after 1 step (a) and
3 steps (b) from time0 = 0.1 d with their eta-solve iteration counts, and a
seeded vector-Jacobian product of one step at a tight eta tolerance
("cot/", "grad/").  The same file holds, under "m/", one step with the
Shapiro filters off and the Smagorinsky floor aam_min = AAM_MIN
(implicit_p0_fixed/time_loop.py:186-188), and its VJP ("mcot/", "mgrad/").
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import torch

sys.path.insert(0, str(Path(__file__).resolve().parent))
import make_golden as G   # noqa: E402
import make_golden_implicit as GI   # noqa: E402

OUT = Path(__file__).resolve().parent
GAMMA = 0.05
AAM_MIN = 800.0


def shapiro_core(im, jm):
    core = GI.build_core(im, jm)
    core.spatial_shapiro_enabled = True
    core.barotropic_shapiro = core.tracer_shapiro = core.velocity_shapiro = GAMMA
    return core


def main(im=12, jm=10, seed=1234):
    G._install_stubs()
    import neuralpom.cores.implicit_p0_fixed  # noqa: F401
    import neuralpom.hybrid.core_pool  # noqa: F401
    import neuralpom.v1.core_spec  # noqa: F401
    originals = G._force_fp64()
    torch.set_default_dtype(torch.float64)
    try:
        out = {}
        for label, nstep in (("a", 1), ("b", 3)):
            core = shapiro_core(im, jm)
            G.perturb(core, seed)
            core.time0 = core.time = 0.1
            if label == "a":
                for k, v in G._tensors(core).items():
                    out[f"in/{k}"] = G._np(v)
                out.update(GI.scalars(core))
                out["scalar/shapiro_gamma"] = np.array(GAMMA)
            core.iint = 0
            core.iend = nstep
            with torch.no_grad():
                core.run_time_loop()
            for k, v in G._tensors(core).items():
                out[f"{label}/{k}"] = G._np(v)
            out[f"{label}/iters"] = np.array(core._eta_solver_iteration_history, dtype=np.int64)
        core = shapiro_core(im, jm)
        G.perturb(core, seed)
        core.time0 = core.time = 0.1
        core.eta_solver_rel_tol = 1e-12
        core.eta_solver_tol = 0.0
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
        # ---- Smagorinsky floor, Shapiro filters off
        for grad_pass in (False, True):
            core = GI.build_core(im, jm)
            core.aam_min = AAM_MIN
            G.perturb(core, seed)
            core.time0 = core.time = 0.1
            core.eta_solver_rel_tol = 1e-12
            core.eta_solver_tol = 0.0
            core.iint = 0
            core.iend = 1
            if not grad_pass:
                with torch.no_grad():
                    core.run_time_loop()
                for k, v in G._tensors(core).items():
                    out[f"m/{k}"] = G._np(v)
                out["m/iters"] = np.array(core._eta_solver_iteration_history, dtype=np.int64)
                out["scalar/aam_min"] = np.array(AAM_MIN)
                continue
            leaves = {}
            for name in GI.VJP_INPUTS:
                leaf = getattr(core, name).detach().clone().requires_grad_(True)
                leaves[name] = leaf
                setattr(core, name, leaf.clone())
            core.run_time_loop()
            g = torch.Generator().manual_seed(seed + 2)
            loss = 0.0
            for name in GI.VJP_OUTPUTS:
                o = getattr(core, name)
                cot = torch.randn(o.shape, generator=g, dtype=torch.float64)
                out[f"mcot/{name}"] = G._np(cot)
                loss = loss + (o * cot).sum()
            grads = torch.autograd.grad(loss, [leaves[n] for n in GI.VJP_INPUTS], allow_unused=True)
            for name, gr in zip(GI.VJP_INPUTS, grads):
                out[f"mgrad/{name}"] = G._np(gr) if gr is not None else np.zeros(leaves[name].shape)
        np.savez_compressed(OUT / f"implicit_shapiro_{im}x{jm}.npz", **out)
    finally:
        for name, fn in originals.items():
            setattr(torch, name, fn)


if __name__ == "__main__":
    main()
