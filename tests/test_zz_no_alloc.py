"""pom_eta_solve_noalloc (the eta solve that allocates nothing: static
per-GPU scratch, no host synchronisation) against pom_eta_solve on the same
inputs, through the C ABI.

Two calls in a row, the second on a smaller batch (the scratch keeps the
first call's contents), for both preconditioners, in float64 and in float32
(the matrix-core spectral products): the solutions and iteration counts are
bitwise equal, since both run the same kernels and a converged member's
extra iterations in the no-alloc solve do nothing.  A batch whose work
arrays exceed the scratch is rejected before any launch.
"""
import dataclasses

import pytest
import torch

from neuralpom_amd import _lib as L
from neuralpom_amd.implicit import ImplicitEnsemble
from tests.test_gpu_implicit import _scalars

pytestmark = pytest.mark.gpu


def _members(shapes, seed):
    from oracle.pomref import implicit as IM
    from oracle.pomref.cases import double_gyre
    out = []
    for n, (im, jm) in enumerate(shapes):
        S = double_gyre(im, jm)
        IM.configure(S)
        gen = torch.Generator().manual_seed(seed + n)
        S.el = S.el + 0.05 * torch.randn(S.el.shape, generator=gen, dtype=S.el.dtype) * S.fsm
        S.d = S.h + S.el
        out.append(S)
    return out


def _solve_pair(shapes, dtype, precond, seed):
    states = _members(shapes, seed)
    ens = ImplicitEnsemble([(S.im, S.jm) for S in states], dtype=dtype, device="cuda")
    geom = {n: ens.pack([getattr(S, n) for S in states], n) for n in L.GEOM}
    sc = dataclasses.replace(_scalars(states[0]), eta_preconditioner=precond, eta_max_iter=150,
                             eta_rel_tol=1e-8, eta_tol=0.0)
    ens.prepare(geom, sc)
    gen = torch.Generator().manual_seed(seed + 100)
    rhs = ens.pack([1e-3 * torch.randn((S.im, S.jm), generator=gen, dtype=torch.float64) * S.fsm for S in states],
                   "el")
    d = ens.pack([S.d for S in states], "d")
    x0 = ens.pack([S.el for S in states], "el")
    it_ref = torch.zeros(len(states), dtype=torch.int32, device="cuda")
    it_na = torch.full((len(states),), -1, dtype=torch.int32, device="cuda")
    x_ref = ens.eta_solve(geom, sc, rhs, d, x0, it_ref)
    x_na = torch.full_like(x_ref, float("nan"))
    ens.eta_solve_noalloc(geom, sc, rhs, d, x_na, x0, it_na)
    torch.cuda.synchronize()
    return x_ref, x_na, it_ref.cpu().tolist(), it_na.cpu().tolist()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("precond", ["spectral", "jacobi"])
def test_noalloc_solve_equals_workspace_solve(dtype, precond):
    for call, shapes in enumerate((((26, 22), (19, 31), (12, 10)), ((12, 10),))):
        print(f"[gpu] eta solve noalloc vs workspace: call {call}, {shapes}, {dtype}, {precond}", flush=True)
        x_ref, x_na, it_ref, it_na = _solve_pair(shapes, dtype, precond, 40 + 7 * call)
        assert it_na == it_ref, (it_na, it_ref)
        assert max(it_ref) > 2
        assert torch.equal(x_na, x_ref), float((x_na - x_ref).abs().max())


def test_noalloc_rejects_a_batch_beyond_its_scratch():
    print("[gpu] eta solve noalloc: 8 x 354x322 float64 exceeds the static scratch", flush=True)
    shapes = [(354, 322)] * 8
    ens = ImplicitEnsemble(shapes, dtype=torch.float64, device="cuda")
    n = ens.numel("el")
    import ctypes as C
    need = C.c_size_t()
    L.check(ens.lib.pom_eta_workspace(ens._h, C.byref(need)), "pom_eta_workspace")
    assert need.value > L.NOALLOC_SCRATCH_BYTES
    geom = {g: torch.zeros(ens.numel(g), dtype=torch.float64, device="cuda") for g in L.GEOM}
    geom["dx"].fill_(1.0)
    geom["dy"].fill_(1.0)
    geom["art"].fill_(1.0)
    from neuralpom_amd.implicit import ImplicitScalars
    sc = ImplicitScalars(dte=20.0, eta_preconditioner="jacobi")
    ens._prepared = True   # the solve is rejected before the preconditioner cache is read
    z = torch.zeros(n, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(z)
    with pytest.raises(ValueError, match="static scratch"):
        ens.eta_solve_noalloc(geom, sc, z, z + 1.0, x)
    torch.cuda.synchronize()
