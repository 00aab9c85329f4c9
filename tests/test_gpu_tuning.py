"""The launch shapes (pom_batch_tune) change only how blocks are spread over
the grid: forward and adjoint results are bitwise equal under the defaults
and under shapes at both ends of the tuner's range; out-of-range shapes are
rejected and leave the batch's shapes as they were."""
import ctypes as C

import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd import tuning as TU
from tests._gpu_inputs import ensemble_inputs, ramp_of, scalars_of
from tests.test_gpu_graph import _members

pytestmark = pytest.mark.gpu

SHAPES = [
    dict(TU.DEFAULT),
    dict(zip(TU.KEYS, (1, 1, 1, 1, 1, 1))),               # one block per group, no grid-stride
    dict(zip(TU.KEYS, (8, 1, 4, 1, 4, 1))),               # few groups: every group strides over blocks
    dict(zip(TU.KEYS, (2, 64, 3, 5, 7, 3))),
]


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_results_do_not_depend_on_launch_shapes(dtype):
    states = _members()
    ens, state, geom, vg = ensemble_inputs(states, dtype)
    sc = scalars_of(states[0])
    ramp = torch.tensor([ramp_of(S, 1) for S in states], dtype=dtype, device="cuda")
    gen = torch.Generator(device="cuda").manual_seed(4)
    cot = {n: torch.randn(ens.numel(n), generator=gen, device="cuda", dtype=dtype)
           for n in ("u", "v", "t", "s", "el", "ua", "va", "q2", "km")}
    ref = None
    for shapes in SHAPES:
        print(f"[gpu] explicit step + adjoint, 3 members, {dtype}, launch shapes {shapes}", flush=True)
        ens.set_tuning(shapes)
        assert TU.current(ens.lib, ens._h) == shapes
        out, va = ens.step(state, geom, sc, ramp, vg, True)
        grads = ens.step_adj(state, geom, sc, ramp, vg, True, dict(cot))
        torch.cuda.synchronize()
        got = ({n: out[n].clone() for n in L.FIELDS}, va.clone(), {n: grads[n].clone() for n in L.DIFF_INPUTS})
        if ref is None:
            ref = got
            continue
        assert not [n for n in L.FIELDS if not torch.equal(got[0][n], ref[0][n])]
        assert torch.equal(got[1], ref[1])
        assert not [n for n in L.DIFF_INPUTS if not torch.equal(got[2][n], ref[2][n])]


def test_out_of_range_shapes_rejected():
    states = _members()
    ens, *_ = ensemble_inputs(states, torch.float64)
    before = TU.current(ens.lib, ens._h)
    for bad in (0, -3, (1 << 20) + 1):
        t = L.Tuning(*[int(before[k]) for k in TU.KEYS])
        t.flat3_min_groups = bad
        assert ens.lib.pom_batch_tune(ens._h, C.byref(t)) == L.POM_E_ARG
        assert TU.current(ens.lib, ens._h) == before
