"""Product case construction (neuralpom_amd/cases.py) vs the oracle's
double-gyre set-up, itself pinned to the reference (test_oracle_golden)."""
import torch

from neuralpom_amd import cases
from oracle.pomref.cases import double_gyre as oracle_dg


def test_double_gyre_matches_oracle():
    st, g, vg, period = cases.double_gyre(12, 10, dtype=torch.float64)
    S = oracle_dg(12, 10)
    for n, v in list(st.items()) + list(g.items()):
        ref = getattr(S, n)
        e = (v - ref).abs().max().item() / max(ref.abs().max().item(), 1e-300)
        assert e < 1e-12, f"{n}: {e}"
    for i, n in enumerate(("z", "zz", "dz", "dzz")):
        assert torch.allclose(vg[i], getattr(S, n), rtol=0, atol=1e-15)
    assert abs(period - S.period) < 1e-12


def test_seamount_is_stable_and_seeded():
    st, g, vg, period = cases.seamount(20, 18, seed=1000, dtype=torch.float64)
    st2, _, _, _ = cases.seamount(20, 18, seed=1001, dtype=torch.float64)
    assert torch.isfinite(st["rho"]).all()
    assert not torch.equal(st["t"], st2["t"])
    # stratification: temperature decreases with depth at the basin centre
    t = st["t"][10, 9, :15]
    assert (t[1:] < t[:-1] + 0.5).all()
    assert g["h"].max() <= 4500.0 and g["h"][10, 9] < 2000.0


def test_seamount_stratification_has_constant_n2_through_the_model_eos():
    """T(z) of the seamount (before its seeded noise) gives N^2 = 1e-5 s^-2 at
    every level pair through the model's own EOS: the oracle's dens (pinned to
    the reference) evaluated at zero pressure, i.e. the potential density,
    rises by rhoref N^2 / g per metre of depth between the levels zz h."""
    from types import SimpleNamespace

    from oracle.pomref import explicit as X
    st, g, vg, period = cases.seamount(20, 18, seed=1000, noise=0.0, dtype=torch.float64)
    for i, j in ((4, 5), (10, 9), (15, 12)):
        t = st["tb"][i, j].view(1, 1, 16)
        s = st["sb"][i, j].view(1, 1, 16)
        depth = -(vg[1] * g["h"][i, j])
        E = SimpleNamespace(kbm1=15, tbias=0.0, sbias=0.0, grav=9.806, rhoref=1035.0,
                            zz=torch.zeros(16, dtype=torch.float64), h=torch.ones(1, 1, dtype=torch.float64),
                            fsm=torch.ones(1, 1, dtype=torch.float64))
        rho = X.dens(E, t, s, torch.zeros_like(t))[0, 0, :15] * 1035.0
        n2 = 9.806 / 1035.0 * (rho[1:] - rho[:-1]) / (depth[1:15] - depth[:14])
        assert torch.allclose(n2, torch.full_like(n2, 1e-5), rtol=1e-9, atol=0), (i, j, n2)
    # the noise is the only seeded part
    st2, _, _, _ = cases.seamount(20, 18, seed=1001, noise=0.0, dtype=torch.float64)
    assert torch.equal(st["tb"], st2["tb"])
