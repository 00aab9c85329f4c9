"""State movement on the HIP extension (neuralpom_amd.state).

* pack_input: against the reference's construction restated with torch
  permute / cat (src/neuralpom/hybrid/state.py:181-226), with and without
  per-cell statistics; its adjoint against torch autograd of the same;
* inject_state: values, history fallbacks and the zeroed optional fields;
* restrict4: the constant field is preserved, the ring is 0, the restriction
  conserves the weighted integral per level and member, and a brute-force
  loop on a small grid agrees;
* SnapshotRing / copy_arrays: bitwise round trip.
"""
import numpy as np
import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd import state as ST
from tests._helpers import rel_err


def test_pack_contract_constants():
    assert ST.INPUT_CHANNELS == 101
    assert ST.INPUT_PACK_ORDER == ("u", "ub", "v", "vb", "t", "s", "el", "ua", "uab", "va", "vab")


# ---------------------------------------------------------------- GPU
def _ens(shapes, dtype=torch.float64):
    from neuralpom_amd.explicit import Ensemble
    return Ensemble(shapes, dtype=dtype, device="cuda")


def _random_state(ens, seed=0):
    gen = torch.Generator().manual_seed(seed)
    return {n: torch.randn(ens.numel(n), generator=gen, dtype=torch.float64).to("cuda", ens.dtype)
            for n in L.FIELDS}


def _torch_pack(ens, state, means=None, scales=None):
    """The reference's pack (normalise, channels first, concatenate) per member."""
    out = []
    for m in range(len(ens.shapes)):
        pieces = []
        for n in ST.INPUT_PACK_ORDER:
            x = ens.view(state[n], m, n)
            if means is not None:
                x = (x - means[n]) / scales[n]
            pieces.append(x.permute(2, 0, 1) if x.dim() == 3 else x.unsqueeze(0))
        out.append(torch.cat(pieces, 0).reshape(-1))
    return torch.cat(out)


@pytest.mark.gpu
def test_pack_matches_torch_and_adjoint():
    print("[gpu] pack_input ragged float64", flush=True)
    ens = _ens([(12, 10), (9, 7), (13, 6)])
    st = _random_state(ens)
    leaves = {n: st[n].clone().requires_grad_(True) for n in ST.INPUT_PACK_ORDER}
    got = ST.pack_input(ens, leaves)
    ref = _torch_pack(ens, {n: leaves[n] for n in ST.INPUT_PACK_ORDER})
    assert torch.equal(got.detach(), ref.detach())
    gen = torch.Generator().manual_seed(3)
    w = torch.randn(got.shape, generator=gen, dtype=torch.float64).cuda()
    g1 = torch.autograd.grad((got * w).sum(), list(leaves.values()))
    g2 = torch.autograd.grad((ref * w).sum(), list(leaves.values()))
    for n, a, b in zip(ST.INPUT_PACK_ORDER, g1, g2):
        assert torch.equal(a, b), n


@pytest.mark.gpu
def test_pack_with_statistics():
    print("[gpu] pack_input normalised float32", flush=True)
    ens = _ens([(12, 10)] * 3, dtype=torch.float32)
    st = _random_state(ens, 1)
    gen = torch.Generator().manual_seed(5)
    means, scales = {}, {}
    for n in ST.INPUT_PACK_ORDER:
        shape = (12, 10, 16) if n in ST.INPUT_THREED_FIELDS else (12, 10)
        means[n] = torch.randn(shape, generator=gen).cuda()
        scales[n] = (torch.rand(shape, generator=gen) + 0.5).cuda()

    class Norm:
        pass
    norm = Norm()
    norm.means = {n: v.unsqueeze(0) for n, v in means.items()}
    norm.scales = {n: v.unsqueeze(0) for n, v in scales.items()}
    leaves = {n: st[n].clone().requires_grad_(True) for n in ST.INPUT_PACK_ORDER}
    got = ST.pack_input(ens, leaves, norm)
    ref = _torch_pack(ens, leaves, means, scales)
    assert rel_err(got.detach().cpu(), ref.detach().cpu()) < 1e-6
    g1 = torch.autograd.grad(got.sum(), list(leaves.values()))
    g2 = torch.autograd.grad(ref.sum(), list(leaves.values()))
    for n, a, b in zip(ST.INPUT_PACK_ORDER, g1, g2):
        assert rel_err(a.cpu(), b.cpu()) < 1e-6, n
    ragged = _ens([(12, 10), (9, 7)], dtype=torch.float32)
    with pytest.raises(ValueError):
        ST.pack_input(ragged, _random_state(ragged), norm)


@pytest.mark.gpu
def test_inject_state_with_fallbacks():
    print("[gpu] inject_state", flush=True)
    ens = _ens([(12, 10), (9, 7)])
    st = _random_state(ens, 2)
    gen = torch.Generator().manual_seed(6)
    vals = {"u": torch.randn(9, 7, 16, generator=gen, dtype=torch.float64),
            "el": torch.randn(9, 7, generator=gen, dtype=torch.float64),
            "t": torch.randn(9 * 7 * 16, generator=gen, dtype=torch.float64)}
    before0 = {n: ens.view(st[n], 0, n).clone() for n in L.FIELDS}
    ST.inject_state(ens, st, 1, vals)
    assert torch.equal(ens.view(st["u"], 1, "u").cpu(), vals["u"])
    assert torch.equal(ens.view(st["ub"], 1, "ub").cpu(), vals["u"])     # history falls back to current
    assert torch.equal(ens.view(st["elb"], 1, "elb").cpu(), vals["el"])
    assert torch.equal(ens.view(st["t"], 1, "t").cpu().reshape(-1), vals["t"])
    assert float(ens.view(st["drx2d"], 1, "drx2d").abs().max()) == 0.0
    for n in L.FIELDS:   # the other member is untouched
        assert torch.equal(ens.view(st[n], 0, n), before0[n]), n


def _restrict_case(dtype=torch.float64):
    fine_shapes = [(18, 14), (26, 10)]
    fine = _ens(fine_shapes, dtype)
    coarse = _ens([((im - 2) // 4 + 2, (jm - 2) // 4 + 2) for im, jm in fine_shapes], dtype)
    gen = torch.Generator().manual_seed(9)
    geom = {}
    for n in ("art", "h", "fsm", "dum", "dvm"):
        g = torch.rand(fine.numel(n), generator=gen, dtype=torch.float64) + 0.5
        if n in ("fsm", "dum", "dvm"):
            g = (torch.rand(fine.numel(n), generator=gen) > 0.15).double()
        geom[n] = g.to("cuda", dtype)
    el = (0.1 * torch.randn(fine.numel("el"), generator=gen, dtype=torch.float64)).to("cuda", dtype)
    vg = (torch.rand(len(fine_shapes), 4, 16, generator=gen, dtype=torch.float64) + 0.1).to("cuda", dtype)
    return fine, coarse, geom, el, vg


def _brute(fine, coarse, m, f, geom, el, vg, three_d, mask):
    im, jm = fine.shapes[m]
    imc, jmc = coarse.shapes[m]
    F = fine.view(f, m, "t" if three_d else "el").cpu().numpy()
    art = fine.view(geom["art"], m, "art").cpu().numpy()
    msk = fine.view(geom[mask], m, mask).cpu().numpy()
    h = fine.view(geom["h"], m, "h").cpu().numpy()
    e = fine.view(el, m, "el").cpu().numpy()
    dz = vg[m, 2].cpu().numpy()
    out = np.zeros((imc, jmc, 16) if three_d else (imc, jmc))
    for I in range(1, imc - 1):
        for J in range(1, jmc - 1):
            for k in range(16 if three_d else 1):
                sw = sf = 0.0
                for i in range(4 * (I - 1) + 1, 4 * I + 1):
                    for j in range(4 * (J - 1) + 1, 4 * J + 1):
                        w = art[i, j] * msk[i, j] * (h[i, j] + e[i, j]) * (dz[k] if three_d else 1.0)
                        sw += w
                        sf += w * (F[i, j, k] if three_d else F[i, j])
                v = sf / sw if sw > 0 else 0.0
                if three_d:
                    out[I, J, k] = v
                else:
                    out[I, J] = v
    return out


@pytest.mark.gpu
def test_restriction_conserves_and_matches_brute_force():
    print("[gpu] restrict4 18x14, 26x10 -> coarse", flush=True)
    fine, coarse, geom, el, vg = _restrict_case()
    gen = torch.Generator().manual_seed(11)
    t = torch.randn(fine.numel("t"), generator=gen, dtype=torch.float64).cuda()
    ct = ST.restrict4(fine, coarse, t, "t", geom, el, vg)
    e2 = torch.randn(fine.numel("el"), generator=gen, dtype=torch.float64).cuda()
    cel = ST.restrict4(fine, coarse, e2, "el", geom, el, vg)
    u = torch.randn(fine.numel("u"), generator=gen, dtype=torch.float64).cuda()
    cu = ST.restrict4(fine, coarse, u, "u", geom, el, vg)
    for m in range(2):
        assert np.allclose(coarse.view(ct, m, "t").cpu().numpy(), _brute(fine, coarse, m, t, geom, el, vg, True,
                                                                         "fsm"), rtol=1e-13, atol=1e-13)
        assert np.allclose(coarse.view(cel, m, "el").cpu().numpy(),
                           _brute(fine, coarse, m, e2, geom, el, vg, False, "fsm"), rtol=1e-13, atol=1e-13)
        assert np.allclose(coarse.view(cu, m, "u").cpu().numpy(), _brute(fine, coarse, m, u, geom, el, vg, True,
                                                                         "dum"), rtol=1e-13, atol=1e-13)
    # a constant is preserved wherever a block holds weight; the ring is 0
    one = torch.full((fine.numel("t"),), 3.5, dtype=torch.float64, device="cuda")
    c1 = ST.restrict4(fine, coarse, one, "t", geom, el, vg)
    for m in range(2):
        v = coarse.view(c1, m, "t")
        assert float(v[0].abs().max()) == 0.0 and float(v[:, -1].abs().max()) == 0.0
        inner = v[1:-1, 1:-1]
        assert float((inner[inner != 0] - 3.5).abs().max()) < 1e-13
    # conservation: sum over coarse cells of (block weight x value) = fine weighted sum, per level
    for m in range(2):
        im, jm = fine.shapes[m]
        w = (fine.view(geom["art"], m, "art") * fine.view(geom["fsm"], m, "fsm") *
             (fine.view(geom["h"], m, "h") + fine.view(el, m, "el"))).unsqueeze(2) * vg[m, 2].view(1, 1, 16)
        f = fine.view(t, m, "t")
        wi = w[1:-1, 1:-1]
        fi = f[1:-1, 1:-1]
        W = wi.reshape((im - 2) // 4, 4, (jm - 2) // 4, 4, 16).sum((1, 3))
        cf = coarse.view(ct, m, "t")[1:-1, 1:-1]
        lhs = (W * cf).sum((0, 1))
        rhs = (wi * fi).sum((0, 1))
        assert rel_err(lhs.cpu(), rhs.cpu()) < 1e-12


def _torch_restrict(fine, m, f, geom, el, vg, three_d, mask):
    """The restriction of member m as torch ops (differentiable reference)."""
    im, jm = fine.shapes[m]
    w = fine.view(geom["art"], m, "art") * fine.view(geom[mask], m, mask) * (fine.view(geom["h"], m, "h") + el)
    if three_d:
        w = w.unsqueeze(2) * vg[m, 2].view(1, 1, 16)
    wi, fi = w[1:-1, 1:-1], f[1:-1, 1:-1]
    shp = ((im - 2) // 4, 4, (jm - 2) // 4, 4) + ((16,) if three_d else ())
    sw = wi.reshape(shp).sum((1, 3))
    sf = (wi * fi).reshape(shp).sum((1, 3))
    inner = torch.where(sw > 0, sf / torch.where(sw > 0, sw, torch.ones_like(sw)), torch.zeros_like(sw))
    if three_d:   # the coarse edge ring is zero
        return torch.nn.functional.pad(inner.permute(2, 0, 1), (1, 1, 1, 1)).permute(1, 2, 0)
    return torch.nn.functional.pad(inner, (1, 1, 1, 1))


@pytest.mark.gpu
def test_restriction_gradient_matches_autograd():
    """The transpose of the restriction and the cotangent of el through the
    volume weights, against torch autograd of the same restriction."""
    print("[gpu] restrict4 adjoint 18x14, 26x10", flush=True)
    fine, coarse, geom, el, vg = _restrict_case()
    gen = torch.Generator().manual_seed(12)
    for name, mask in (("t", "fsm"), ("u", "dum"), ("el", "fsm")):
        three_d = name in ("t", "u")
        f = torch.randn(fine.numel(name), generator=gen, dtype=torch.float64).cuda().requires_grad_(True)
        e = el.clone().requires_grad_(True)
        out = ST.restrict4(fine, coarse, f, name, geom, e, vg)
        cot = torch.randn(out.shape, generator=gen, dtype=torch.float64).cuda()
        gf, ge = torch.autograd.grad((out * cot).sum(), [f, e])
        f2 = f.detach().clone().requires_grad_(True)
        e2 = el.clone().requires_grad_(True)
        loss = 0.0
        for m in range(len(fine.shapes)):
            ref = _torch_restrict(fine, m, fine.view(f2, m, name), geom, fine.view(e2, m, "el"), vg, three_d, mask)
            assert rel_err(coarse.view(out, m, name).detach().cpu(), ref.detach().cpu()) < 1e-13, name
            loss = loss + (ref * coarse.view(cot, m, name)).sum()
        rf, re = torch.autograd.grad(loss, [f2, e2])
        assert rel_err(gf.cpu(), rf.cpu()) < 1e-12, f"{name} field gradient"
        assert rel_err(ge.cpu(), re.cpu()) < 1e-12, f"{name} el gradient"


@pytest.mark.gpu
def test_restriction_rejects_mismatched_grids():
    fine = _ens([(18, 14)])
    bad = _ens([(7, 5)])   # (18 - 2) / 4 + 2 = 6 rows expected
    geom = {n: torch.ones(fine.numel(n), dtype=torch.float64, device="cuda") for n in ("art", "h", "fsm")}
    with pytest.raises(ValueError):
        ST.restrict4(fine, bad, torch.ones(fine.numel("t"), dtype=torch.float64, device="cuda"), "t", geom,
                     None, torch.ones(1, 4, 16, dtype=torch.float64, device="cuda"))


@pytest.mark.gpu
def test_snapshot_ring_round_trip():
    print("[gpu] snapshot ring", flush=True)
    ens = _ens([(12, 10), (9, 7)], dtype=torch.float32)
    st = _random_state(ens, 4)
    ring = ST.SnapshotRing(st, depth=2)
    k0 = ring.capture(st)
    saved = {n: v.clone() for n, v in st.items()}
    for v in st.values():
        v.mul_(-3.0)
    k1 = ring.capture(st)
    assert (k0, k1) == (0, 1)
    back = ring.restore(k0)
    for n in L.FIELDS:
        assert torch.equal(back[n], saved[n]), n
    ring.restore(k1, into=saved)
    for n in L.FIELDS:
        assert torch.equal(saved[n], st[n]), n
