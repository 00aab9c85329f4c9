"""Write golden vectors for the semi-implicit (P0-fixed) POM step from the reference.

Runs only in a container that has the reference checkout at /root/reference
(read-only; bytecode writing is disabled before import).  Uses the same
netCDF4 stub and float64 factory wrapping as make_golden.py, and the reference's
own construction protocol (src/neuralpom/hybrid/core_pool.py:30-48) and
production settings (src/neuralpom/v1/core_spec.py:22-33: alpha .333, smoth
.15, Shapiro gamma 0, eta relative tolerance 1e-5, spectral preconditioner).

Outputs (``tests/golden/implicit_*_<im>x<jm>.npz``), all float64:

* ``implicit_steps``  -- perturbed input state ("in/"), then the state after
  1 step (a) and 3 steps (b) from time0 = 0.1 d, 1 step from time0 = 0 (c, the
  internal-mode skip), and 1 step with a tight eta tolerance (t: relative
  1e-12, absolute 0); eta-solve iteration counts per label.
* ``implicit_ops``    -- first-call inputs/outputs of the barotropic stages of
  one step: predictor (+bcond 2), transport/RHS, eta solve (rhs, D, initial
  guess, solution, iterations), post-solve surface BC, corrector (+bcond 2),
  with the 2-D attributes they read; plus the elliptic operator applied to a
  seeded field.
* ``implicit_vjp``    -- seeded-cotangent vector-Jacobian product of one step
  at the tight eta tolerance (torch.autograd through the reference step,
  including its custom eta-solve adjoint).
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np
import torch

sys.path.insert(0, str(Path(__file__).resolve().parent))
import make_golden as G   # noqa: E402  (stubs, fp64 factories, perturbation)

OUT = Path(__file__).resolve().parent

VJP_INPUTS = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab",
              "va", "vab", "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam",
              "et", "etb", "d", "dt", "rho", "egb", "utb", "vtb", "wubot", "wvbot")
VJP_OUTPUTS = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab",
               "va", "vab", "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam",
               "et", "etb", "etf", "elf", "rho", "egb", "utb", "vtb", "wubot", "wvbot", "w")
SCALARS = G.SCALARS + ("eta_solver_rel_tol",)


def build_core(im, jm):
    from neuralpom.hybrid.core_pool import build_initialized_core
    from neuralpom.v1.core_spec import core_spec

    spec = core_spec("implicit_p0_fixed")
    core = build_initialized_core(spec.core_class, torch.device("cpu"), im, jm)
    spec.configure(core)
    return core


def scalars(core):
    return {f"scalar/{k}": np.array(float(getattr(core, k)), dtype=np.float64) for k in SCALARS}


def two_d(core):
    return {k: G._np(v) for k, v in G._tensors(core).items() if v.dim() == 2}


def record_barotropic(core):
    """First calls of the barotropic stages, with the 2-D attributes they read."""
    rec = {}

    def wrap(name):
        fn = getattr(core, name)

        def inner(*args, **kwargs):
            first = name not in rec
            if first:
                pre = two_d(core)
                a = [G._np(x) for x in args if torch.is_tensor(x)]
                kw = {k: G._np(v) for k, v in kwargs.items() if torch.is_tensor(v)}
            out = fn(*args, **kwargs)
            if first:
                outs = out if isinstance(out, tuple) else (out,)
                rec[name] = {"self": pre, "args": a, "kw": kw,
                             "out": [G._np(o) for o in outs if torch.is_tensor(o)],
                             "iters": getattr(core, "_last_eta_solver_iterations", -1)}
            return out
        return inner

    for name in ("_barotropic_predictor_with_bc", "_barotropic_transport_rhs_full", "solve_implicit_eta_cg",
                 "_surface_bc_after_solver", "_barotropic_corrector_with_bc"):
        setattr(core, name, wrap(name))
    return rec


def main(im=12, jm=10, seed=1234):
    G._install_stubs()
    # every module that resets the default dtype at import time is imported
    # before the float64 default is set
    import neuralpom.cores.implicit_p0_fixed  # noqa: F401
    import neuralpom.hybrid.core_pool  # noqa: F401
    import neuralpom.v1.core_spec  # noqa: F401
    originals = G._force_fp64()
    torch.set_default_dtype(torch.float64)
    try:
        tag = f"{im}x{jm}"
        probe = build_core(im, jm)
        bad = [k for k, v in G._tensors(probe).items() if v.dtype != torch.float64]
        if bad or torch.get_default_dtype() != torch.float64:
            raise RuntimeError(f"reference core is not float64: {bad}")
        steps = {}
        for label, time0, nstep, tight in (("a", 0.1, 1, False), ("b", 0.1, 3, False), ("c", 0.0, 1, False),
                                            ("t", 0.1, 1, True)):
            core = build_core(im, jm)
            G.perturb(core, seed)
            core.time0 = core.time = time0
            if tight:
                core.eta_solver_rel_tol = 1e-12
                core.eta_solver_tol = 0.0
            if label == "a":
                for k, v in G._tensors(core).items():
                    steps[f"in/{k}"] = G._np(v)
                steps.update(scalars(core))
            core.iint = 0
            core.iend = nstep
            with torch.no_grad():
                core.run_time_loop()
            for k, v in G._tensors(core).items():
                steps[f"{label}/{k}"] = G._np(v)
            steps[f"{label}/time0"] = np.array(time0)
            steps[f"{label}/nstep"] = np.array(nstep)
            steps[f"{label}/iters"] = np.array(core._eta_solver_iteration_history, dtype=np.int64)
        np.savez_compressed(OUT / f"implicit_steps_{tag}.npz", **steps)

        # ---- barotropic stages of one step ----
        core = build_core(im, jm)
        G.perturb(core, seed)
        core.time0 = core.time = 0.1
        rec = record_barotropic(core)
        core.iint = 0
        core.iend = 1
        with torch.no_grad():
            core.run_time_loop()
        ops = {}
        for name, r in rec.items():
            for k, v in r["self"].items():
                ops[f"{name}/self/{k}"] = v
            for i, v in enumerate(r["args"]):
                ops[f"{name}/arg{i}"] = v
            for k, v in r["kw"].items():
                ops[f"{name}/kw/{k}"] = v
            for i, v in enumerate(r["out"]):
                ops[f"{name}/out{i}"] = v
            ops[f"{name}/iters"] = np.array(r["iters"])
        g = torch.Generator().manual_seed(seed + 7)
        eta = torch.randn((im, jm), generator=g, dtype=torch.float64) * 0.1
        dd = core.h + core.el
        ops["elliptic/eta"] = G._np(eta)
        ops["elliptic/d"] = G._np(dd)
        ops["elliptic/out"] = G._np(core.apply_elliptic_operator(eta, dd))
        ops.update(scalars(core))
        np.savez_compressed(OUT / f"implicit_ops_{tag}.npz", **ops)

        # ---- VJP of one step (tight eta tolerance) ----
        core = build_core(im, jm)
        G.perturb(core, seed)
        core.time0 = core.time = 0.1
        core.eta_solver_rel_tol = 1e-12
        core.eta_solver_tol = 0.0
        leaves = {}
        for name in VJP_INPUTS:
            leaf = getattr(core, name).detach().clone().requires_grad_(True)
            leaves[name] = leaf
            setattr(core, name, leaf.clone())
        core.iint = 0
        core.iend = 1
        core.run_time_loop()
        g = torch.Generator().manual_seed(seed + 1)
        loss = 0.0
        vjp = {}
        for name in VJP_OUTPUTS:
            out = getattr(core, name)
            cot = torch.randn(out.shape, generator=g, dtype=torch.float64)
            vjp[f"cot/{name}"] = G._np(cot)
            loss = loss + (out * cot).sum()
        grads = torch.autograd.grad(loss, [leaves[n] for n in VJP_INPUTS], allow_unused=True)
        for name, gr in zip(VJP_INPUTS, grads):
            vjp[f"grad/{name}"] = G._np(gr) if gr is not None else np.zeros(leaves[name].shape)
        np.savez_compressed(OUT / f"implicit_vjp_{tag}.npz", **vjp)
    finally:
        for name, fn in originals.items():
            setattr(torch, name, fn)


if __name__ == "__main__":
    main(int(os.environ.get("GOLDEN_IM", "12")), int(os.environ.get("GOLDEN_JM", "10")))
