"""The drop-in core (neuralpom_amd.core.POM2K_Core) against the oracle.

The construction protocol and attribute contract run on the CPU; the time
loop, its autograd / checkpoint composition and the operator overrides
dens() / baropg() run the HIP kernels (marked gpu).  The oracle itself is
pinned to the reference's golden vectors (tests/test_oracle_golden.py).
"""
import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd.core import POM2K_Core
from oracle.pomref import explicit as X
from oracle.pomref.cases import double_gyre as oracle_dg
from tests._helpers import rel_err

IM, JM = 12, 10
DIFF = ("u", "v", "t", "s", "el", "ua", "va")


def _core(device, dtype=torch.float64, **over):
    core = POM2K_Core(IM, JM, device=device, dtype=dtype, settings_overrides=over or None)
    core.configure_runtime(dte=20.0, isplit=30)
    core.run_initialization()
    return core


def test_construction_matches_oracle_double_gyre():
    core = _core("cpu")
    S = oracle_dg(IM, JM)
    for n in L.FIELDS + L.GEOM + ("z", "zz", "dz", "dzz"):
        e = rel_err(getattr(core, n), getattr(S, n))
        assert e < 1e-12, f"{n}: {e}"
    assert core.dti == 600.0 and core.dte2 == 40.0 and core.isp2i == 1.0 / 60.0
    assert abs(core.period - S.period) < 1e-12
    assert core.horcon == 0.05 and core.kbm1 == 15
    # state is plain tensors in attributes (vars(core) capture contract)
    tensors = {k for k, v in vars(core).items() if torch.is_tensor(v)}
    assert set(L.FIELDS) <= tensors


def test_settings_overrides_refresh_derived():
    core = POM2K_Core(IM, JM, device="cpu", settings_overrides={"dte": 10.0, "isplit": 20, "days": 1.0})
    assert core.dti == 200.0 and core.dti2 == 400.0 and core.iend == 432
    with pytest.raises(ValueError):
        POM2K_Core(IM, JM, problem_id=3, device="cpu")


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_time_loop_has_no_cpu_path():
    if not L._LIB_PATH.exists():
        pytest.skip("library not built")
    core = _core("cpu")
    core.iend = 1
    with pytest.raises(L.HipExtensionError):
        core.run_time_loop()


# ---------------------------------------------------------------- GPU
def _perturb(seed=5):
    """Seeded perturbation of the initial double gyre: a state at rest has
    drhox ~ 0 up to rounding, which no relative comparison can resolve."""
    S = oracle_dg(IM, JM)
    gen = torch.Generator().manual_seed(seed)
    for n, a in (("t", 0.5), ("el", 0.05), ("u", 0.02), ("v", 0.02)):
        x = getattr(S, n)
        setattr(S, n, x + a * torch.randn(x.shape, generator=gen, dtype=x.dtype))
    S.d = S.h + S.el
    return S


def _core_from(S, device="cuda"):
    core = _core(device)
    for n in ("t", "el", "u", "v", "d"):
        setattr(core, n, getattr(S, n).to(device).clone())
    return core


@pytest.mark.gpu
@pytest.mark.parametrize("time0", [0.0, 0.25])
def test_run_time_loop_matches_oracle(time0):
    print(f"[gpu] core run_time_loop 12x10 fp64 time0={time0}", flush=True)
    S = _perturb()
    core = _core_from(S)
    core.time0 = core.time = time0
    core.iend = 3
    core.run_time_loop()
    torch.cuda.synchronize()
    S.time0 = S.time = time0
    S.iend = 3
    assert X.run_time_loop(S)
    assert core.iint == 3 and abs(core.time - S.time) < 1e-15 and abs(core.ramp - S.ramp) < 1e-15
    bad = [f"{n}: {rel_err(getattr(core, n).cpu(), getattr(S, n)):.2e}" for n in L.FIELDS
           if not rel_err(getattr(core, n).cpu(), getattr(S, n)) <= 1e-10]
    assert not bad, bad


def _loss(st):
    return sum((getattr(st, n) ** 2).sum() for n in ("u", "v", "t", "el")) + (st.s * st.t).sum()


@pytest.mark.gpu
def test_autograd_and_checkpoint_match_oracle():
    print("[gpu] core autograd + checkpoint 12x10 fp64", flush=True)
    S0 = _perturb(7)
    core = _core_from(S0)
    core.time0 = core.time = 0.25
    core.iend = 2
    leaves = {n: getattr(core, n).clone().requires_grad_(True) for n in DIFF}
    for n, v in leaves.items():
        setattr(core, n, v)
    core.run_time_loop()
    grads = torch.autograd.grad(_loss(core), list(leaves.values()))

    S = S0.copy()
    S.time0 = S.time = 0.25
    S.iend = 2
    oleaves = {n: getattr(S, n).clone().requires_grad_(True) for n in DIFF}
    for n, v in oleaves.items():
        setattr(S, n, v)
    assert X.run_time_loop(S)
    ref = torch.autograd.grad(_loss(S), list(oleaves.values()))
    bad = [f"{n}: {rel_err(g.detach().cpu(), r):.2e}" for n, g, r in zip(DIFF, grads, ref)
           if not rel_err(g.detach().cpu(), r) <= 1e-8]
    assert not bad, bad

    # the checkpointed chunk pattern of src/neuralpom/hybrid/sparse_physics.py:108-149
    core2 = _core_from(S0)
    core2.time0 = core2.time = 0.25
    names = tuple(sorted(k for k, v in vars(core2).items() if torch.is_tensor(v)))
    values = [getattr(core2, k).clone() for k in names]
    for i, k in enumerate(names):
        if k in DIFF:
            values[i].requires_grad_(True)

    def run(*inputs):
        local = _core("cuda")
        for k, v in zip(names, inputs):
            setattr(local, k, v.clone())
        local.time0 = local.time = 0.25
        local.iint, local.iend = 0, 2
        local.run_time_loop()
        return tuple(getattr(local, k) for k in names)

    outs = torch.utils.checkpoint.checkpoint(run, *values, use_reentrant=False, preserve_rng_state=False)
    out = dict(zip(names, outs))
    loss = sum((out[n] ** 2).sum() for n in ("u", "v", "t", "el")) + (out["s"] * out["t"]).sum()
    cg = torch.autograd.grad(loss, [values[names.index(n)] for n in DIFF])
    for n, a, b in zip(DIFF, cg, grads):
        assert torch.equal(a, b), f"checkpointed gradient differs for {n}"


@pytest.mark.gpu
def test_dens_and_baropg_overrides_match_oracle():
    print("[gpu] core dens/baropg 12x10 fp64", flush=True)
    core = _core("cuda")
    gen = torch.Generator().manual_seed(3)
    t0 = (core.t.cpu() + 0.5 * torch.randn(core.t.shape, generator=gen, dtype=torch.float64))
    core.t = t0.to("cuda").requires_grad_(True)
    core.ramp = 0.7
    s_, t_, rho = core.dens()
    rho2, px, py = core.baropg()
    loss = (rho2 * 1.3).sum() + (px ** 2).sum() * 1e6 + (py * 2.0).sum() * 1e3
    g_t, = torch.autograd.grad(loss, [core.t])

    S = oracle_dg(IM, JM)
    tt = t0.clone().requires_grad_(True)
    r = X.dens(S, tt, S.s, S.rho)
    r2, qx, qy = X.baropg(S, r, S.dt, 0.7)
    ref_loss = (r2 * 1.3).sum() + (qx ** 2).sum() * 1e6 + (qy * 2.0).sum() * 1e3
    ref_g, = torch.autograd.grad(ref_loss, [tt])
    for a, b, n in ((rho, r, "rho"), (px, qx, "drhox"), (py, qy, "drhoy")):
        assert rel_err(a.detach().cpu(), b.detach()) <= 1e-12, n
    assert rel_err(g_t.cpu(), ref_g) <= 1e-10
