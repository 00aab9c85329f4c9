"""Pin the oracle (oracle/pomref) to golden vectors written from the reference.

The vectors come from tests/golden/make_golden.py, which runs the reference's
explicit core (src/neuralpom/cores/explicit/*) in float64 on a 12x10x16
double-gyre basin with a seeded perturbation of every dynamic field.  Each
operator is checked on the exact inputs it saw inside a reference step, then
1- and 3-step trajectories, the time0 = 0 internal-mode skip and a seeded VJP.
"""
import numpy as np
import pytest
import torch

from oracle.pomref import explicit as X
from oracle.pomref.state import load_state
from tests._helpers import assert_rel, golden

TAG = "12x10"
TOL = 1e-11


@pytest.fixture(scope="module")
def ops():
    return golden(f"explicit_ops_{TAG}.npz")


@pytest.fixture(scope="module")
def S0(ops):
    S = load_state(ops, "t")
    S.time = S.dti / 86400.0 + 0.1
    S.ramp = min(S.time / S.period, 1.0)
    return S


def A(ops, key):
    return torch.from_numpy(np.array(ops[key]))


def test_advct(ops, S0):
    a = lambda i: A(ops, f"advct/arg{i}")
    advx, advy, curv = X.advct(S0, a(6), a(7), a(9), a(10), a(34), a(5))
    assert_rel(curv, A(ops, "advct/out2"), TOL, "curv")
    assert_rel(advx, A(ops, "advct/out3"), TOL, "advx")
    assert_rel(advy, A(ops, "advct/out4"), TOL, "advy")


def test_baropg(ops, S0):
    s = lambda n: A(ops, f"baropg/self/{n}")
    rho, dx_, dy_ = X.baropg(S0, s("rho"), s("dt"), S0.ramp)
    assert_rel(rho, A(ops, "baropg/out0"), 1e-15, "rho")
    assert_rel(dx_, A(ops, "baropg/out1"), TOL, "drhox")
    assert_rel(dy_, A(ops, "baropg/out2"), TOL, "drhoy")


def test_advave(ops, S0):
    a = lambda i: A(ops, f"advave/arg{i}")
    advua, advva, fluxua, fluxva, tps = X.advave(S0, a(14), a(15), a(10), a(11), a(19), a(9),
                                                 a(3), a(4), a(7))
    assert_rel(advua, A(ops, "advave/out1"), TOL, "advua")
    assert_rel(advva, A(ops, "advave/out2"), TOL, "advva")
    assert_rel(fluxua, A(ops, "advave/out3"), TOL, "fluxua")
    assert_rel(fluxva, A(ops, "advave/out4"), TOL, "fluxva")
    assert_rel(tps, A(ops, "advave/out7"), TOL, "tps")


def test_vertvl(ops, S0):
    a = lambda i: A(ops, f"vertvl/arg{i}")
    w = X.vertvl(S0, a(7), a(8), a(2), a(6), a(11), a(12), a(9), a(10))
    assert_rel(w, A(ops, "vertvl/out2"), TOL, "w")


def test_advq(ops, S0):
    a = lambda i: A(ops, f"advq/arg{i}")
    qf = X.advq(S0, a(0), a(1), a(9), a(10), a(11), a(12), a(5), a(17), a(18))
    assert_rel(qf, A(ops, "advq/out2"), TOL, "q2f")


def test_profq(ops, S0):
    a = lambda i: A(ops, f"profq/arg{i}")
    q2f, q2lf, l, kq, km, kh = X.profq(
        S0, a(10), a(11), a(12), a(13), a(14), a(8), a(9), a(7), a(25), a(26), a(43), a(44),
        a(23), a(39), a(41), a(40), a(42), a(18), a(27))
    for got, k, name in ((l, 6, "l"), (kq, 7, "kq"), (km, 8, "km"), (kh, 9, "kh"),
                         (q2f, 10, "q2f"), (q2lf, 11, "q2lf")):
        assert_rel(got, A(ops, f"profq/out{k}"), TOL, name)


def test_advt1(ops, S0):
    a = lambda i: A(ops, f"advt1/arg{i}")
    fb, f, ff = X.advt1(S0, a(0), a(1), a(2), a(8), a(9), a(14), a(10), a(7), a(16), a(17))
    assert_rel(fb, A(ops, "advt1/out0"), 0, "fb")
    assert_rel(f, A(ops, "advt1/out1"), 0, "f")
    assert_rel(ff, A(ops, "advt1/out3"), TOL, "ff")


def test_proft(ops, S0):
    a = lambda i: A(ops, f"proft/arg{i}")
    f = X.proft(S0, a(0), a(1), a(2), int(ops["proft/arg3"]), a(21), a(10), a(14))
    assert_rel(f, A(ops, "proft/out0"), TOL, "f")


def test_dens(ops, S0):
    s = lambda n: A(ops, f"dens/self/{n}")
    rho = X.dens(S0, s("t"), s("s"), s("rho"))
    assert_rel(rho, A(ops, "dens/out2"), 1e-13, "rho")


def test_advu_advv(ops, S0):
    a = lambda i: A(ops, f"advu/arg{i}")
    uf = X.advu(S0, a(0), a(7), a(12), a(13), a(14), a(10), a(4), a(22), a(23), a(24), a(25), a(5))
    assert_rel(uf, A(ops, "advu/out0"), TOL, "uf")
    a = lambda i: A(ops, f"advv/arg{i}")
    vf = X.advv(S0, a(0), a(7), a(12), a(13), a(14), a(10), a(4), a(22), a(23), a(24), a(25), a(5))
    assert_rel(vf, A(ops, "advv/out0"), TOL, "vf")


def test_profu_profv(ops, S0):
    a = lambda i: A(ops, f"profu/arg{i}")
    uf, wubot, tps = X.profu(S0, a(5), a(14), a(6), a(9), a(7), a(24), a(25))
    assert_rel(uf, A(ops, "profu/out5"), TOL, "uf")
    assert_rel(wubot, A(ops, "profu/out6"), TOL, "wubot")
    assert_rel(tps, A(ops, "profu/out4"), TOL, "tps_u")
    a = lambda i: A(ops, f"profv/arg{i}")
    vf, wvbot, tps = X.profv(S0, a(5), a(22), a(6), a(17), a(24), a(19), a(20))
    assert_rel(vf, A(ops, "profv/out5"), TOL, "vf")
    assert_rel(wvbot, A(ops, "profv/out6"), TOL, "wvbot")
    assert_rel(tps, A(ops, "profv/out4"), TOL, "tps_v")


def test_bconds(ops, S0):
    a = lambda n, i: A(ops, f"bcond{n}/arg{i}")
    s = lambda n, f: A(ops, f"bcond{n}/self/{f}")
    assert_rel(X.bcond1(S0, a(1, 1)), A(ops, "bcond1/out0"), 0, "bcond1")
    uaf, vaf = X.bcond2(S0, a(2, 2), a(2, 3), s(2, "el"), S0.ramp)
    assert_rel(uaf, A(ops, "bcond2/out1"), 1e-15, "bcond2 uaf")
    assert_rel(vaf, A(ops, "bcond2/out2"), 1e-15, "bcond2 vaf")
    uf, vf = X.bcond3(S0, a(3, 4), a(3, 5), s(3, "u"), s(3, "v"))
    assert_rel(uf, A(ops, "bcond3/out3"), 1e-15, "bcond3 uf")
    assert_rel(vf, A(ops, "bcond3/out4"), 1e-15, "bcond3 vf")
    uf, vf = X.bcond4(S0, a(4, 4), a(4, 5), s(4, "t"), s(4, "s"), s(4, "u"), s(4, "v"),
                      s(4, "w"), s(4, "dt"))
    assert_rel(uf, A(ops, "bcond4/out3"), 1e-15, "bcond4 uf")
    assert_rel(vf, A(ops, "bcond4/out4"), 1e-15, "bcond4 vf")
    assert_rel(X.bcond5(S0, a(5, 6)), A(ops, "bcond5/out5"), 0, "bcond5")
    uf, vf = X.bcond6(S0, a(6, 4), a(6, 5), s(6, "q2"), s(6, "q2l"), s(6, "u"), s(6, "v"))
    assert_rel(uf, A(ops, "bcond6/out3"), 1e-15, "bcond6 uf")
    assert_rel(vf, A(ops, "bcond6/out4"), 1e-15, "bcond6 vf")


DYNAMIC = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab", "va", "vab",
           "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam", "et", "etb", "etf", "rho",
           "w", "egf", "egb", "utf", "utb", "vtf", "vtb", "wubot", "wvbot", "dt", "d", "l",
           "uaf", "vaf", "elf", "advua", "advva", "adx2d", "ady2d", "drx2d", "dry2d", "aam2d",
           "advx", "advy", "drhox", "drhoy")


@pytest.mark.parametrize("label,tol", [("a", 1e-10), ("b", 1e-9), ("c", 1e-10)])
def test_steps(label, tol):
    g = golden(f"explicit_steps_{TAG}.npz")
    S = load_state(g, "in")
    S.time0 = float(g[f"{label}/time0"])
    S.time = S.time0
    S.iend = int(g[f"{label}/nstep"])
    assert X.run_time_loop(S)
    assert abs(S.time - float(g[f"{label}/time"])) < 1e-12
    for name in DYNAMIC:
        assert_rel(getattr(S, name), torch.from_numpy(g[f"{label}/{name}"]), tol, name)


VJP_INPUTS = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab",
              "va", "vab", "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam",
              "et", "etb", "d", "dt", "rho", "egb", "utb", "vtb", "wubot", "wvbot")
VJP_OUTPUTS = ("u", "ub", "v", "vb", "t", "tb", "s", "sb", "el", "elb", "ua", "uab",
               "va", "vab", "q2", "q2b", "q2l", "q2lb", "km", "kh", "kq", "aam",
               "et", "etb", "etf", "rho", "egb", "utb", "vtb", "wubot", "wvbot", "w")


def test_vjp():
    g = golden(f"explicit_steps_{TAG}.npz")
    v = golden(f"explicit_vjp_{TAG}.npz")
    S = load_state(g, "in")
    S.time0 = S.time = 0.1
    S.iend = 1
    leaves = {}
    for n in VJP_INPUTS:
        leaves[n] = getattr(S, n).clone().requires_grad_(True)
        setattr(S, n, leaves[n])
    assert X.run_time_loop(S)
    loss = sum((getattr(S, n) * torch.from_numpy(v[f"cot/{n}"])).sum() for n in VJP_OUTPUTS)
    grads = torch.autograd.grad(loss, [leaves[n] for n in VJP_INPUTS], allow_unused=True)
    for n, gr in zip(VJP_INPUTS, grads):
        gr = torch.zeros_like(leaves[n]) if gr is None else gr
        assert_rel(gr, torch.from_numpy(v[f"grad/{n}"]), 1e-8, f"grad {n}")


def test_double_gyre_init():
    """oracle/pomref/cases.py vs the reference's state after construction."""
    from oracle.pomref.cases import double_gyre
    g = golden(f"explicit_init_{TAG}.npz")
    S = double_gyre(12, 10)
    bad = []
    for key in g.files:
        if not key.startswith("t/"):
            continue
        n = key[2:]
        if n in ("io", "jo", "ko", "curv"):
            continue
        ref = torch.from_numpy(np.array(g[key]))
        got = getattr(S, n, None)
        if got is None:
            bad.append(f"{n}: missing")
            continue
        e = (got - ref).abs().max().item() / max(ref.abs().max().item(), 1e-300)
        if e > 1e-12:
            bad.append(f"{n}: {e:.1e}")
    assert not bad, ", ".join(bad)
    assert abs(S.period - float(g["scalar/period"])) < 1e-12
