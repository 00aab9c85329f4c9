"""Write golden vectors for the corrector hook, the density-75 projection and
the corrector network from the reference implementation.

Runs only in a container that has the reference checkout at /root/reference
(read-only), with the same import stubs and float64 factories as
make_golden.py.  The core is the reference's explicit double gyre at 12 x 10
with make_golden.py's seeded perturbation (seed 1234).

Output ``tests/golden/corrector_12x10.npz`` (float64):

* ``in/<field>``, ``scalar/<name>``: the perturbed core state the calls
  start from.
* ``jc/<field>``, ``jc/diag``: the state after
  ``apply_joint_correction(core, raw, scales, cap=5)``
  (hybrid/sparse_coupling.py:42-67) for the seeded raw ``jc/raw`` and scales
  ``scales`` (u, v, t, s), and the diagnostics (tracer factor, t budget,
  s budget); the tracer cap binds, the velocity cap does not.
* ``jccot/<name>``, ``jcgrad/<name>``: a seeded vector-Jacobian product of
  that call: cotangents of the 8 corrected fields, rho and the diagnostics,
  gradients of raw, et and the 8 state inputs.
* ``d75<c>/<field>``, ``d75<c>/diag`` for c in (a, b, z): the state after
  ``apply_density75(core, raw, scales)`` (evaluation/density.py:38-64) and
  its diagnostics (ts limit factor, density-component factor, constraint
  residual max; NaN for a key the call does not return); raw ``d75<c>/raw``; case a with a small raw (no cap binds),
  case b with a large one (the joint-correction tracer cap binds; the
  density-component cap does not: after the first cap |x| <= 5, and the 25 %
  shrink of the density component moves a component past 5 only through the
  layer constraint), case z with a zero raw (the early return,
  density.py:44-45).
* ``net/x``, ``net/y``, ``net/p/<name>``: a CorrectorNet of width 4
  (models/corrector_net.py:72-100) with seeded parameters (head included),
  its input and output.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import torch

sys.path.insert(0, str(Path(__file__).resolve().parent))
import make_golden as G  # noqa: E402

STATE = ("u", "ub", "v", "vb", "t", "tb", "s", "sb")
SCALES = {"u": 0.02, "v": 0.02, "t": 0.05, "s": 0.01}


def _core(seed):
    core = G.build_core(12, 10)
    G.perturb(core, seed)
    core.rho = core.dens()[2]
    return core


def main(seed=1234):
    G._install_stubs()
    import neuralpom.cores.explicit  # noqa: F401  (before the factories are wrapped, as make_golden.py)
    G._force_fp64()
    torch.set_default_dtype(torch.float64)
    from neuralpom.evaluation.density import apply_density75
    from neuralpom.hybrid.sparse_coupling import apply_joint_correction
    from neuralpom.models.corrector_net import CorrectorNet

    out = {}
    core = _core(seed)
    for k, v in G._tensors(core).items():
        out[f"in/{k}"] = G._np(v)
    out.update(G.scalars(core))
    out["scales"] = np.array([SCALES[f] for f in ("u", "v", "t", "s")])
    g = torch.Generator().manual_seed(seed + 7)

    # ---- joint correction: forward + VJP
    raw = torch.randn((60, core.im, core.jm), generator=g, dtype=torch.float64)
    raw[:30] *= 0.3          # velocity peak below the cap
    raw[30:] *= 3.0          # tracer peak above it
    out["jc/raw"] = G._np(raw)
    leaves = {"raw": raw.clone().requires_grad_(True), "et": core.et.clone().requires_grad_(True)}
    for f in STATE:
        leaves[f] = getattr(core, f).clone().requires_grad_(True)
    c = _core(seed)
    c.et = leaves["et"]
    for f in STATE:
        setattr(c, f, leaves[f])
    diag = apply_joint_correction(c, leaves["raw"], SCALES, cap=5.0)
    for f in STATE + ("rho",):
        out[f"jc/{f}"] = G._np(getattr(c, f))
    dvals = [diag["ts_limit_factor"], diag["t_budget"], diag["s_budget"]]
    out["jc/diag"] = np.array([float(x.detach()) for x in dvals])
    loss = 0.0
    for f in STATE + ("rho",):
        o = getattr(c, f)
        cot = torch.randn(o.shape, generator=g, dtype=torch.float64)
        out[f"jccot/{f}"] = G._np(cot)
        loss = loss + (o * cot).sum()
    cd = torch.randn(3, generator=g, dtype=torch.float64)
    out["jccot/diag"] = G._np(cd)
    loss = loss + sum(cd[n] * dvals[n] for n in range(3))
    names = list(leaves)
    grads = torch.autograd.grad(loss, [leaves[n] for n in names], allow_unused=True)
    for n, gr in zip(names, grads):
        out[f"jcgrad/{n}"] = G._np(gr) if gr is not None else np.zeros(tuple(leaves[n].shape))

    # ---- density-75 projection: small raw, large raw, zero raw
    for case, amp in (("a", 0.2), ("b", 4.0), ("z", 0.0)):
        c = _core(seed)
        raw = torch.randn((60, c.im, c.jm), generator=g, dtype=torch.float64) * amp
        diag = apply_density75(c, raw, SCALES)
        out[f"d75{case}/raw"] = G._np(raw)
        for f in ("t", "tb", "s", "sb", "rho", "u", "ub", "v", "vb"):
            out[f"d75{case}/{f}"] = G._np(getattr(c, f))
        # the early return leaves the last two keys out: NaN marks them
        out[f"d75{case}/diag"] = np.array([float(diag.get(k, float("nan"))) for k in (
            "ts_limit_factor", "density_component_limit_factor", "density_constraint_residual_max_abs")])

    # ---- CorrectorNet of width 4 with seeded parameters
    torch.manual_seed(seed + 11)
    net = CorrectorNet(hidden_dim=4).double()
    with torch.no_grad():
        for p in net.parameters():
            p.copy_(torch.randn(p.shape, generator=g, dtype=torch.float64) * 0.3)
    net.eval()
    x = torch.randn((2, 101, core.im, core.jm), generator=g, dtype=torch.float64)
    with torch.no_grad():
        y = net(x)
    out["net/x"] = G._np(x)
    out["net/y"] = G._np(y)
    for name, p in net.state_dict().items():
        out[f"net/p/{name}"] = G._np(p)

    np.savez_compressed(G.OUT / "corrector_12x10.npz", **out)


if __name__ == "__main__":
    main()
