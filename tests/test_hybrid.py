"""CorrectorNet structure on the CPU: output shape, the zero head of the
reference (every correction zero until trained), the seeded head, and the
module / parameter names of src/neuralpom/models/corrector_net.py, so that a
state_dict of the reference's network loads into this one."""
import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd.hybrid import CorrectorNet


def test_shapes_and_heads():
    x = torch.randn(1, 101, 26, 22, generator=torch.Generator().manual_seed(0))
    net = CorrectorNet(hidden_dim=8)
    y = net(x)
    assert y.shape == (1, 60, 26, 22)
    assert float(y.abs().max()) == 0.0
    seeded = CorrectorNet(hidden_dim=8, head_std=0.05, generator=torch.Generator().manual_seed(1))
    assert float(seeded(x).abs().max()) > 0.0


def test_parameter_names_follow_the_reference_layout():
    net = CorrectorNet(hidden_dim=64)
    sd = net.state_dict()
    for key, shape in (("inc.double_conv.0.weight", (64, 101, 3, 3)),
                       ("inc.double_conv.1.weight", (64,)),
                       ("inc.double_conv.3.weight", (64, 64, 3, 3)),
                       ("inc.shortcut.0.weight", (64, 101, 1, 1)),
                       ("down1.maxpool_conv.1.double_conv.0.weight", (128, 64, 3, 3)),
                       ("down3.maxpool_conv.1.shortcut.1.bias", (512,)),
                       ("up1.conv.double_conv.0.weight", (384, 768, 3, 3)),
                       ("up1.conv.double_conv.3.weight", (256, 384, 3, 3)),
                       ("up3.conv.shortcut.0.weight", (64, 192, 1, 1)),
                       ("outc.weight", (60, 64, 1, 1)), ("outc.bias", (60,))):
        assert tuple(sd[key].shape) == shape, key
    # GroupNorm groups: min(32, c) reduced to a divisor of c
    assert net.inc.double_conv[1].num_groups == 32
    assert CorrectorNet(hidden_dim=12).inc.double_conv[1].num_groups == 12
    assert CorrectorNet(hidden_dim=40).inc.double_conv[1].num_groups == 20


def _toy_windows(monkeypatch, segment, steps=10):
    """integrate's run structure with a smooth torch map in place of the HIP
    step: returns (gradient of the end state's loss at the start, step calls)."""
    from neuralpom_amd import _lib as L
    from neuralpom_amd import hybrid
    calls = []

    def toy_step(ens, state, geom, vgrid, schedule, it):
        calls.append(it)
        a = 0.1 * it
        out = {n: x + a * torch.sin(x) * state["t"].mean() for n, x in state.items()}
        return out, torch.tensor(float(it))

    monkeypatch.setattr(hybrid, "_step", toy_step)
    gen = torch.Generator().manual_seed(0)
    start = {n: torch.randn(3, generator=gen, dtype=torch.float64).requires_grad_(True) for n in L.FIELDS}
    end, vamax = hybrid.integrate(None, start, None, None, None, 1, steps, segment)
    assert float(vamax) == steps
    loss = sum((v ** 2).sum() for v in end.values())
    g = torch.autograd.grad(loss, [start[n] for n in L.FIELDS])
    return g, len(calls)


def test_integrate_segment_recompute_levels(monkeypatch):
    """Runs recomputed in one or two nested levels give bitwise the gradient
    of plain autograd, with at most one extra forward per level (a level's
    recompute stops once the states it keeps are rebuilt, torch's early stop)."""
    ref, n0 = _toy_windows(monkeypatch, 1)
    assert n0 == 10
    for segment, levels in ((4, 1), ((6, 2), 2), ((4, 3), 2), ((5, 1), 1)):
        g, n = _toy_windows(monkeypatch, segment)
        assert 10 < n <= 10 * (1 + levels), (segment, n)
        for a, b in zip(g, ref):
            assert torch.equal(a, b), segment


@pytest.mark.parametrize("bad", [float("nan"), 1.0e3])
def test_truncated_windows_rejects_nan_or_large_vamax(monkeypatch, bad):
    """A diverged window (vamax NaN or above vmaxl) raises before its loss is
    back-propagated, as the core's time loop does."""
    from neuralpom_amd import hybrid

    def fake_integrate(ens, start, geom, vgrid, schedule, it, steps, segment):
        return start, torch.tensor(bad)

    monkeypatch.setattr(hybrid, "integrate", fake_integrate)
    state = {n: torch.zeros(4) for n in L.FIELDS}
    called = []
    with pytest.raises(FloatingPointError):
        hybrid.truncated_windows(None, state, {}, None, None, 2, 3,
                                 lambda end, w: called.append(w) or torch.zeros((), requires_grad=True),
                                 vmaxl=100.0)
    assert not called
