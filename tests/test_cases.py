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
