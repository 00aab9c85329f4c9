"""The explicit step recorded into a HIP graph (``StepGraph``) against the
same steps run eagerly.

Each replay is checked, with eager work and host synchronisation between the
replays (graph nodes that only the first replay gets right would show here),
for the advanced state and the adjoint's input cotangents.  The graph runs
the same kernels in the same order, so results are bitwise equal.
"""
import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd.explicit import StepGraph
from tests._gpu_inputs import ensemble_inputs, ramp_of, scalars_of

pytestmark = pytest.mark.gpu


def _members():
    from oracle.pomref.cases import double_gyre
    states = []
    for n, (im, jm) in enumerate(((26, 22), (19, 31), (12, 10))):
        S = double_gyre(im, jm)
        gen = torch.Generator().manual_seed(60 + n)
        for f, a in (("t", 0.5), ("el", 0.05), ("u", 0.02), ("v", 0.02), ("ua", 0.01), ("va", 0.01)):
            x = getattr(S, f)
            setattr(S, f, x + a * torch.randn(x.shape, generator=gen, dtype=x.dtype))
        S.d = S.h + S.el
        S.time0 = S.time = 0.01   # ramp still rising: each step has its own ramp
        states.append(S)
    return states


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_graph_replays_match_eager_steps(dtype):
    print(f"[gpu] StepGraph vs eager, 3 members, dtype={dtype}", flush=True)
    states = _members()
    ens, state, geom, vg = ensemble_inputs(states, dtype)
    sc = scalars_of(states[0])
    ramps = [torch.tensor([ramp_of(S, it) for S in states], dtype=dtype, device="cuda") for it in (1, 2, 3, 4)]
    gen = torch.Generator(device="cuda").manual_seed(3)
    cot = {n: torch.randn(ens.numel(n), generator=gen, device="cuda", dtype=dtype)
           for n in ("u", "v", "t", "s", "el", "ua", "va", "q2", "km")}

    G = StepGraph(ens, state, geom, sc, ramps[0], vg, True, cot)
    eager = {n: state[n].clone() for n in L.FIELDS}
    for r in range(3):
        G.ramp.copy_(ramps[r])
        vamax = G.replay()
        # eager reference of the same step, with unrelated eager work in between
        out, vamax_e = ens.step(eager, geom, sc, ramps[r], vg, True)
        grads = ens.step_adj(eager, geom, sc, ramps[r], vg, True, dict(cot))
        torch.cuda.synchronize()
        bad = [n for n in L.FIELDS if not torch.equal(G.state[n], out[n])]
        bad += [f"grad {n}" for n in L.DIFF_INPUTS if not torch.equal(G.grads[n], grads[n])]
        assert not bad, f"replay {r}: differs from the eager step in {bad}"
        assert torch.equal(vamax, vamax_e)
        eager = out
