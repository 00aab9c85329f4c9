"""Gradients of the HIP explicit step against central finite differences.

The hand-written adjoint is chained over several internal steps through the
``ExplicitStep`` autograd node (backward recomputes each step from its saved
inputs), and the directional derivative it gives, g . delta, is compared with
the central difference (L(x + eps delta) - L(x - eps delta)) / (2 eps) of the
same HIP forward on 32 seeded directions.  The loss is a seeded linear
functional of the prognostic state after the last step.

The start state is a perturbed double-gyre ragged batch spun up by a few HIP
steps.  Directions are relative to each point's own magnitude,
delta = (|x| + 1e-3 max|x|) * N(0, 1), because the closure divides by q2b,
whose values span five decades: a direction scaled by the field maximum moves
the smallest values by percents and the difference quotient then measures
curvature, not the derivative.  Float64 only: in float32 the loss (about 1e4)
is resolved to about 1e-3, and a difference quotient with eps = 1e-3 of that
is rounding noise (the same holds for the oracle's own float32 autograd); the
float32 adjoint is checked against the reference's float64 VJP instead.
"""
import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd.explicit import step_autograd
from tests._gpu_inputs import ensemble_inputs, ramp_of, scalars_of

pytestmark = pytest.mark.gpu

# prognostic fields: the loss reads them after the last step, and the
# directions perturb them (plus the diagnostic inputs the step differentiates)
LOSS_FIELDS = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "q2", "q2b", "q2l", "q2lb",
               "el", "elb", "ua", "uab", "va", "vab", "et", "etb", "egb", "utb", "vtb")


def _start(dtype, shapes=((26, 22), (19, 31)), spin=6):
    from oracle.pomref.cases import double_gyre
    states = []
    for n, (im, jm) in enumerate(shapes):
        S = double_gyre(im, jm)
        gen = torch.Generator().manual_seed(90 + n)
        for f, a in (("t", 0.5), ("el", 0.05), ("u", 0.02), ("v", 0.02), ("ua", 0.01), ("va", 0.01)):
            x = getattr(S, f)
            setattr(S, f, x + a * torch.randn(x.shape, generator=gen, dtype=x.dtype))
        S.d = S.h + S.el
        S.time0 = S.time = 0.25
        states.append(S)
    ens, state, geom, vg = ensemble_inputs(states, dtype)
    sc = scalars_of(states[0])
    ramps = [torch.tensor([ramp_of(S, it) for S in states], dtype=dtype, device="cuda")
             for it in range(1, spin + 8)]
    for it in range(spin):
        state, _ = ens.step(state, geom, sc, ramps[it], vg, True)
    return ens, state, geom, sc, vg, ramps[spin:]


def _loss(ens, state, geom, sc, vg, ramps, nstep, wts):
    for it in range(nstep):
        state, vamax = step_autograd(ens, state, geom, sc, ramps[it], vg, True)
        assert float(vamax) <= sc.vmaxl
    return sum((state[n] * wts[n]).sum() for n in LOSS_FIELDS)


def test_adjoint_matches_central_differences(dtype=torch.float64, eps=1e-8, tol=1e-4, nstep=3):
    print(f"[gpu] central FD vs adjoint, {nstep} steps, dtype={dtype}, eps={eps}", flush=True)
    torch.manual_seed(0)
    ens, x0, geom, sc, vg, ramps = _start(dtype)
    gen = torch.Generator(device="cuda").manual_seed(5)
    wts = {n: torch.randn(x0[n].shape, generator=gen, device="cuda", dtype=dtype) for n in LOSS_FIELDS}
    # pointwise scale: a direction moves every value by about eps of itself
    scale = {n: x0[n].abs() + 1e-3 * x0[n].abs().max() for n in L.DIFF_INPUTS}

    leaves = {n: x0[n].clone().requires_grad_(n in L.DIFF_INPUTS) for n in L.FIELDS}
    loss = _loss(ens, dict(leaves), geom, sc, vg, ramps, nstep, wts)
    grads = torch.autograd.grad(loss, [leaves[n] for n in L.DIFF_INPUTS])
    g = dict(zip(L.DIFF_INPUTS, grads))

    rel = []
    with torch.no_grad():
        for d in range(32):
            gd = torch.Generator(device="cuda").manual_seed(1000 + d)
            delta = {n: scale[n] * torch.randn(x0[n].shape, generator=gd, device="cuda", dtype=dtype)
                     for n in L.DIFF_INPUTS}
            ad = sum(float((g[n].double() * delta[n].double()).sum()) for n in L.DIFF_INPUTS)
            lp = _loss(ens, {n: x0[n] + eps * delta[n] if n in delta else x0[n] for n in L.FIELDS},
                       geom, sc, vg, ramps, nstep, wts)
            lm = _loss(ens, {n: x0[n] - eps * delta[n] if n in delta else x0[n] for n in L.FIELDS},
                       geom, sc, vg, ramps, nstep, wts)
            fd = (float(lp.double()) - float(lm.double())) / (2.0 * eps)
            rel.append(abs(fd - ad) / max(abs(ad), abs(fd), 1e-300))
    worst = max(range(32), key=lambda i: rel[i])
    print(f"[gpu] FD rel err: median {sorted(rel)[16]:.2e}, max {rel[worst]:.2e} (direction {worst})", flush=True)
    assert rel[worst] <= tol, f"direction {worst}: FD vs adjoint rel err {rel[worst]:.2e} > {tol:.0e}"
