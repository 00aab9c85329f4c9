"""The split vector-Jacobian product (pom_*_step_taped + pom_*_step_reverse):
the taped forward equals the plain step and the reverse from its kept
intermediates equals the recomputing adjoint, bitwise, on a ragged batch.
The taped outputs are part of the tape: the reverse reads them and leaves
them as they were, and so it does with the output cotangents (read in place
where the step does not add into them)."""
import dataclasses

import pytest
import torch

from neuralpom_amd import _lib as L
from tests._gpu_inputs import ensemble_inputs, ramp_of, scalars_of
from tests.test_gpu_graph import _members

pytestmark = pytest.mark.gpu


def _cot(ens, names, seed, dtype):
    gen = torch.Generator(device="cuda").manual_seed(seed)
    return {n: torch.randn(ens.numel(n), generator=gen, device="cuda", dtype=dtype) for n in names}


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_explicit_taped_and_reverse_equal_step_and_adjoint(dtype):
    print(f"[gpu] taped forward + reverse vs step + step_adj, explicit, 3 members, {dtype}", flush=True)
    states = _members()
    ens, state, geom, vg = ensemble_inputs(states, dtype)
    sc = scalars_of(states[0])
    ramp = torch.tensor([ramp_of(S, 1) for S in states], dtype=dtype, device="cuda")
    cot = _cot(ens, L.FIELDS, 3, dtype)
    cot0 = {n: c.clone() for n, c in cot.items()}
    out_s, va_s = ens.step(state, geom, sc, ramp, vg, True)
    grads_a = ens.step_adj(state, geom, sc, ramp, vg, True, dict(cot))
    out_t, va_t = ens.step_taped(state, geom, sc, ramp, vg, True)
    for n in L.FIELDS:
        assert torch.equal(out_t[n], out_s[n]), n
    assert torch.equal(va_t, va_s)
    grads_r = ens.step_reverse(state, out_t, geom, sc, ramp, vg, True, dict(cot))
    torch.cuda.synchronize()
    bad = [n for n in L.DIFF_INPUTS if not torch.equal(grads_r[n], grads_a[n])]
    assert not bad, bad
    assert float(grads_r["t"].abs().max()) > 0.0
    # the reverse reads the taped outputs and the output cotangents and leaves them as they were
    assert not [n for n in L.FIELDS if not torch.equal(out_t[n], out_s[n])]
    assert not [n for n in L.FIELDS if not torch.equal(cot[n], cot0[n])]


def test_implicit_taped_and_reverse_equal_step_and_adjoint():
    from tests.test_gpu_implicit import _inputs, _state
    print("[gpu] taped forward + reverse vs step + step_adj, implicit, 2 members", flush=True)
    S1, _ = _state("a")
    S2, _ = _state("a")
    gen = torch.Generator().manual_seed(12)
    S2.t = S2.t + 0.3 * torch.randn(S2.t.shape, generator=gen, dtype=S2.t.dtype)
    states = [S1, S2]
    ens, state, geom, vg, sc = _inputs(states, torch.float64)
    sc = dataclasses.replace(sc, eta_max_iter=120)
    ens.prepare(geom, sc)
    ramp = torch.tensor([ramp_of(S, 1) for S in states], dtype=torch.float64, device="cuda")
    cot = _cot(ens, ("u", "t", "el", "ua", "va"), 5, torch.float64)
    it_s = torch.zeros(2, dtype=torch.int32, device="cuda")
    it_t = torch.zeros(2, dtype=torch.int32, device="cuda")
    out_s, va_s = ens.step(state, geom, sc, ramp, vg, True, it_s)
    grads_a = ens.step_adj(state, geom, sc, ramp, vg, True, dict(cot))
    out_t, va_t = ens.step_taped(state, geom, sc, ramp, vg, True, it_t)
    for n in L.FIELDS:
        assert torch.equal(out_t[n], out_s[n]), n
    assert torch.equal(va_t, va_s) and torch.equal(it_t, it_s)
    grads_r = ens.step_reverse(state, out_t, geom, sc, ramp, vg, True, dict(cot))
    torch.cuda.synchronize()
    bad = [n for n in L.DIFF_INPUTS if not torch.equal(grads_r[n], grads_a[n])]
    assert not bad, bad
