"""Batched explicit POM step on the HIP extension, with its adjoint.

``Ensemble`` packs members of (possibly different) grid sizes back to back in
one buffer per field and drives ``pom_explicit_step`` / ``pom_explicit_step_adj``
(include/neuralpom_amd.h).  ``ExplicitStep`` is the ``torch.autograd.Function``
around one internal step: forward runs the HIP step; backward recomputes the
step from the saved inputs (per-step checkpointing) and runs the hand-written
adjoint kernels.  It keeps no reference to a live core object.

Reference: src/neuralpom/cores/explicit/time_loop.py:20-709 (the step),
src/hybrid/sparse_physics.py:108-149 (checkpointed differentiable advance).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import torch

from . import _lib as L

KB = 16


@dataclass
class StepScalars:
    """Scalars of the explicit core (explicit/base.py:46-118, v1/core_spec.py)."""
    dte: float = 20.0
    isplit: int = 30
    alpha: float = 0.225
    smoth: float = 0.10
    horcon: float = 0.05
    grav: float = 9.806
    rhoref: float = 1035.0
    kappa: float = 0.40
    umol: float = 2.e-5
    tprni: float = 0.20
    small: float = 1.e-9
    tbias: float = 0.0
    sbias: float = 0.0
    hmax: float = 2000.0
    rfe: float = 0.0
    rfw: float = 0.0
    rfn: float = 0.0
    rfs: float = 0.0
    ispadv: int = 5
    ntp: int = 2
    nbct: int = 1
    nbcs: int = 1
    vmaxl: float = 100.0
    boundary: int = L.POM_BOUNDARY_CLOSED   # or POM_BOUNDARY_TIDE (open edges, tide nodes)
    # tide policy, explicit family (coastal_boundary.py:9-14): the nodes of
    # substep n take the tide at tide_sec + dte (tide_substep0 + n - 1)
    # seconds, with the argument tide_coef t / 3600
    tide_coef: float = 0.0                  # speed_deg_per_h * pi / 180
    tide_sec: float = 0.0                   # time * 86400 - dti of the step
    tide_substep0: int = 0                  # the adapter's substep counter before the step

    @property
    def dti(self):
        return self.dte * self.isplit

    def struct(self) -> L.Params:
        p = L.Params()
        p.dti = self.dti
        p.dte = self.dte
        p.dti2 = self.dti * 2.0
        p.dte2 = self.dte * 2.0
        p.ispi = 1.0 / self.isplit
        p.isp2i = 1.0 / (2.0 * self.isplit)
        for n in ("alpha", "smoth", "horcon", "grav", "rhoref", "kappa", "umol", "tprni", "small",
                  "tbias", "sbias", "hmax", "rfe", "rfw", "rfn", "rfs", "vmaxl"):
            setattr(p, n, float(getattr(self, n)))
        for n in ("isplit", "ispadv", "ntp", "nbct", "nbcs"):
            setattr(p, n, int(getattr(self, n)))
        p.iext = 0
        p.core = L.POM_CORE_EXPLICIT
        p.boundary = int(self.boundary)
        p.tide_coef = float(self.tide_coef)
        p.tide_sec = float(self.tide_sec)
        p.tide_substep0 = int(self.tide_substep0)
        return p


class Ensemble:
    """Members of (im, jm) grids, all with 16 sigma levels, packed per field."""

    def __init__(self, shapes, dtype=torch.float32, device="cuda"):
        self.lib = L.load()
        if not torch.cuda.is_available():
            raise L.HipExtensionError("no HIP device: the explicit step has no CPU path")
        self.shapes = [(int(a), int(b)) for a, b in shapes]
        self.dtype = dtype
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.off2, self.off3 = [], []
        o2 = o3 = 0
        for im, jm in self.shapes:
            self.off2.append(o2)
            self.off3.append(o3)
            o2 += im * jm
            o3 += im * jm * KB
        self.n2, self.n3 = o2, o3
        mem = (L.Member * len(self.shapes))()
        for m, (im, jm) in enumerate(self.shapes):
            mem[m].im, mem[m].jm, mem[m].off2, mem[m].off3 = im, jm, self.off2[m], self.off3[m]
        code = L.POM_F64 if dtype == torch.float64 else L.POM_F32
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            L.check(self.lib.pom_batch_create(mem, len(self.shapes), code, C.byref(h)), "pom_batch_create")
        self._h = h
        self._ws = {}
        # launch shapes of the nearest tuned workload (tuning.json)
        from . import tuning as TU
        self.tuning_entry, shapes_ = TU.select(self.shapes, "f64" if dtype == torch.float64 else "f32")
        TU.apply(self.lib, self._h, shapes_)

    def set_tuning(self, shapes: dict, entry: str = "manual"):
        """Override the launch shapes (keys of tuning.KEYS)."""
        from . import tuning as TU
        TU.apply(self.lib, self._h, shapes)
        self.tuning_entry = entry

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.lib.pom_batch_destroy(h)
            except Exception:
                pass

    @property
    def cells(self):
        return self.n3

    # ------------------------------------------------------------ buffers
    def numel(self, name):
        return self.n3 if name in L.THREE_D else self.n2

    def zeros(self, name):
        return torch.zeros(self.numel(name), dtype=self.dtype, device=self.device)

    def view(self, flat, m, name):
        im, jm = self.shapes[m]
        if name in L.THREE_D:
            return flat[self.off3[m]:self.off3[m] + im * jm * KB].view(im, jm, KB)
        return flat[self.off2[m]:self.off2[m] + im * jm].view(im, jm)

    def pack(self, per_member, name):
        """Concatenate member tensors of one field into the packed layout."""
        return torch.cat([t.reshape(-1).to(self.device, self.dtype) for t in per_member]).contiguous()

    def _workspace(self, prm: L.Params, adjoint: bool, pool=None):
        """Scratch of one call, taken from ``pool`` (default: the ensemble's own).
        A larger request replaces the pool's buffer; a caller that recorded
        launches on the old one (a HIP graph) passes a pool of its own."""
        n = C.c_size_t()
        L.check(self.lib.pom_explicit_workspace(self._h, C.byref(prm), int(adjoint), C.byref(n)),
                "pom_explicit_workspace")
        pool = self._ws if pool is None else pool
        ws = pool.get(adjoint)
        if ws is None or ws.numel() < n.value:
            ws = torch.empty(n.value, dtype=torch.uint8, device=self.device)
            pool[adjoint] = ws
        return ws, n.value

    def _check(self, tensors, names, nel):
        for t, n in zip(tensors, names):
            if t.dtype != self.dtype or t.device != self.device or not t.is_contiguous() or t.numel() != nel(n):
                raise ValueError(f"field {n}: expected contiguous {self.dtype} on {self.device} "
                                 f"with {nel(n)} elements, got {t.dtype} {t.device} {tuple(t.shape)}")

    # ------------------------------------------------------------ step
    def _outputs(self, state, out):
        if out is None:
            return {n: torch.empty_like(state[n]) for n in L.FIELDS}
        self._check([out[n] for n in L.FIELDS], L.FIELDS, self.numel)
        return out

    def step(self, state, geom, sc: StepScalars, ramp, vgrid, internal=True, ws_pool=None, out=None):
        """One internal step; returns (new state dict, vamax tensor).  ``out``:
        optional buffers (every field, not aliasing ``state``) for the new state."""
        prm = sc.struct()
        ins = [state[n] for n in L.FIELDS]
        gs = [geom[n] for n in L.GEOM]
        self._check(ins, L.FIELDS, self.numel)
        self._check(gs, L.GEOM, self.numel)
        out = self._outputs(state, out)
        vamax = torch.zeros(1, dtype=self.dtype, device=self.device)
        ws, nb = self._workspace(prm, False, ws_pool)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        rc = self.lib.pom_explicit_step(
            self._h, C.byref(prm), ramp.data_ptr(), vgrid.data_ptr(),
            L.ptr_array([t.data_ptr() for t in ins]), L.ptr_array([out[n].data_ptr() for n in L.FIELDS]),
            L.geom_ptrs(geom), int(bool(internal)), vamax.data_ptr(),
            ws.data_ptr(), nb, stream)
        L.check(rc, "pom_explicit_step")
        return out, vamax

    def _adj_buffers(self, cot):
        """Zero-filled gradient views (one allocation) and the cotangent / gradient
        pointer tables of an adjoint call; ``cot`` entries are made contiguous
        in the ensemble's dtype in place."""
        sizes = [self.numel(n) for n in L.DIFF_INPUTS]
        flat = torch.zeros(sum(sizes), dtype=self.dtype, device=self.device)
        grads = dict(zip(L.DIFF_INPUTS, torch.split(flat, sizes)))
        cot_ptrs = []
        for n in L.FIELDS:
            c = cot.get(n)
            if c is not None:
                c = c.contiguous().to(self.dtype)
                cot[n] = c
                cot_ptrs.append(c.data_ptr())
            else:
                cot_ptrs.append(None)
        grad_ptrs = [grads[n].data_ptr() if n in grads else None for n in L.FIELDS]
        return grads, L.ptr_array(cot_ptrs), L.ptr_array(grad_ptrs)

    def step_adj(self, state, geom, sc: StepScalars, ramp, vgrid, internal, cot, ws_pool=None):
        """Gradients of the step inputs for output cotangents ``cot`` (dict, may omit fields)."""
        return self._reverse_call("pom_explicit_step_adj", state, None, geom, sc, ramp, vgrid, internal, cot,
                                  ws_pool)

    def step_taped(self, state, geom, sc: StepScalars, ramp, vgrid, internal=True, ws_pool=None, out=None):
        """``step`` that keeps its intermediates in the adjoint workspace of
        ``ws_pool`` (default: the ensemble's) and in its outputs, for ``step_reverse``."""
        prm = sc.struct()
        ins = [state[n] for n in L.FIELDS]
        gs = [geom[n] for n in L.GEOM]
        self._check(ins, L.FIELDS, self.numel)
        self._check(gs, L.GEOM, self.numel)
        out = self._outputs(state, out)
        vamax = torch.zeros(1, dtype=self.dtype, device=self.device)
        ws, nb = self._workspace(prm, True, ws_pool)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        rc = self.lib.pom_explicit_step_taped(
            self._h, C.byref(prm), ramp.data_ptr(), vgrid.data_ptr(),
            L.ptr_array([t.data_ptr() for t in ins]), L.ptr_array([out[n].data_ptr() for n in L.FIELDS]),
            L.geom_ptrs(geom), int(bool(internal)), vamax.data_ptr(), ws.data_ptr(), nb, stream)
        L.check(rc, "pom_explicit_step_taped")
        return out, vamax

    def step_reverse(self, state, out, geom, sc: StepScalars, ramp, vgrid, internal, cot, ws_pool=None):
        """``step_adj`` of the last ``step_taped`` of the same arguments, from its
        kept intermediates (no forward recompute).  ``state``, ``out`` (the
        taped call's outputs) and ``geom`` must be unchanged since that call, and
        its workspace untouched."""
        return self._reverse_call("pom_explicit_step_reverse", state, out, geom, sc, ramp, vgrid, internal, cot,
                                  ws_pool)

    def _reverse_call(self, fn, state, out, geom, sc, ramp, vgrid, internal, cot, ws_pool):
        prm = sc.struct()
        ins = [state[n] for n in L.FIELDS]
        self._check(ins, L.FIELDS, self.numel)
        grads, cot_ptrs, grad_ptrs = self._adj_buffers(cot)
        ws, nb = self._workspace(prm, True, ws_pool)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        args = [self._h, C.byref(prm), ramp.data_ptr(), vgrid.data_ptr(), L.ptr_array([t.data_ptr() for t in ins])]
        if out is not None:
            self._check([out[n] for n in L.FIELDS], L.FIELDS, self.numel)
            args.append(L.ptr_array([out[n].data_ptr() for n in L.FIELDS]))
        args += [L.geom_ptrs(geom), int(bool(internal)), cot_ptrs, grad_ptrs, ws.data_ptr(), nb, stream]
        L.check(getattr(self.lib, fn)(*args), fn)
        return grads


class ExplicitStep(torch.autograd.Function):
    """One internal step as an autograd node (inputs and outputs in FIELDS order).

    Backward recomputes the forward from the saved inputs (per-step
    checkpointing) and accumulates the input cotangents with the adjoint
    kernels; non-differentiable inputs get None.
    """

    @staticmethod
    def forward(ctx, ens, geom, sc, ramp, vgrid, internal, *inputs):
        state = dict(zip(L.FIELDS, inputs))
        out, vamax = ens.step(state, geom, sc, ramp, vgrid, internal)
        ctx.ens, ctx.geom, ctx.sc, ctx.internal = ens, geom, sc, internal
        ctx.save_for_backward(ramp, vgrid, *inputs)
        ctx.mark_non_differentiable(vamax)
        return tuple(out[n] for n in L.FIELDS) + (vamax,)

    @staticmethod
    def backward(ctx, *gouts):
        ramp, vgrid, *inputs = ctx.saved_tensors
        state = dict(zip(L.FIELDS, inputs))
        cot = {n: g for n, g in zip(L.FIELDS, gouts[:len(L.FIELDS)]) if g is not None}
        grads = ctx.ens.step_adj(state, ctx.geom, ctx.sc, ramp, vgrid, ctx.internal, cot)
        return (None,) * 6 + tuple(grads.get(n) for n in L.FIELDS)


class StepGraph:
    """One explicit internal step (and optionally its adjoint) recorded into
    HIP graphs and replayed.

    Two graphs are recorded over two state buffers: one steps buffer A into
    buffer B, the other B into A, and replays alternate between them, so the
    new state is never copied back.  ``state`` is the buffer holding the
    current state (read it after a replay; it changes with every replay), and
    with ``cot``, ``grads`` holds the input cotangents of the step just taken:
    the forward keeps its intermediates (``step_taped``) and the reverse runs
    from them (``step_reverse``), so the step's forward runs once.  The graphs
    read the static buffers ``ramp`` and, with ``cot``, the output
    cotangents; refill them between replays to change them.  ``on_warm`` is
    called between the warm-up and the recording, ``on_first`` after the first
    of the two graphs is recorded.  The step launches ~100
    kernels per barotropic substep cycle and has no host synchronisation, so
    a replay submits the whole step with one call.  A semi-implicit ensemble
    (``ImplicitEnsemble``) is recorded the same way with ``eta_capture`` set
    in its scalars: its eta solves then launch a fixed number of iterations
    and never synchronise with the host (converged members do nothing, so the
    replay equals the eager step).

    The graphs own their workspace (allocated on the current stream before
    the warm-up and capture, and held for the graphs' lifetime), so later
    eager calls on the same ensemble, whatever scratch they need, never touch
    the memory the recorded kernels use.
    """

    def __init__(self, ens, state, geom, sc, ramp, vgrid, internal=True, cot=None, warmup=1, on_warm=None,
                 on_first=None):
        if hasattr(sc, "eta_capture") and not sc.eta_capture:
            raise ValueError("a semi-implicit step is recorded with eta_capture=True "
                             "(its eta solves otherwise poll the host for convergence)")
        self.ens = ens
        self._ws = {}
        prm = sc.struct()
        ens._workspace(prm, cot is not None, self._ws)
        self._bufs = [{n: state[n].clone() for n in L.FIELDS}, {n: torch.empty_like(state[n]) for n in L.FIELDS}]
        self._cur = 0
        self.geom, self.sc, self.vgrid, self.internal = geom, sc, vgrid, internal
        self.ramp = ramp.clone()
        self.cot = None if cot is None else {n: c.contiguous().to(ens.dtype).clone() for n, c in cot.items()}
        s = torch.cuda.Stream(device=ens.device)
        s.wait_stream(torch.cuda.current_stream(ens.device))
        with torch.cuda.stream(s):   # warm-up outside the graphs (on copies): workspaces are allocated here
            for _ in range(warmup):
                self._body({n: t.clone() for n, t in self._bufs[0].items()},
                           {n: torch.empty_like(t) for n, t in self._bufs[0].items()})
        torch.cuda.current_stream(ens.device).wait_stream(s)
        if on_warm is not None:   # e.g. reset HIP-event tracing before the recorded launches
            torch.cuda.synchronize(ens.device)
            on_warm()
        self._graphs, self._res = [], []
        for k in (0, 1):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._res.append(self._body(self._bufs[k], self._bufs[1 - k]))
            self._graphs.append(g)
            if k == 0 and on_first is not None:   # e.g. stop HIP-event tracing after the first graph
                on_first()
        self.grads, self.vamax = None, None

    def _body(self, st, out):
        grads = None
        if self.cot is None:
            _, vamax = self.ens.step(st, self.geom, self.sc, self.ramp, self.vgrid, self.internal,
                                     ws_pool=self._ws, out=out)
        else:   # the forward keeps its intermediates and the reverse runs from them
            _, vamax = self.ens.step_taped(st, self.geom, self.sc, self.ramp, self.vgrid, self.internal,
                                           ws_pool=self._ws, out=out)
            grads = self.ens.step_reverse(st, out, self.geom, self.sc, self.ramp, self.vgrid, self.internal,
                                          dict(self.cot), ws_pool=self._ws)
        return grads, vamax

    @property
    def state(self):
        """The buffers holding the current state."""
        return self._bufs[self._cur]

    def replay(self):
        """Advance ``state`` by one step (and fill ``grads``); returns the step's max |vaf| (device)."""
        k = self._cur
        self._graphs[k].replay()
        self.grads, self.vamax = self._res[k]
        self._cur = 1 - k
        return self.vamax


def step_autograd(ens, state, geom, sc, ramp, vgrid, internal=True):
    """Differentiable step: returns (new state dict, vamax)."""
    outs = ExplicitStep.apply(ens, geom, sc, ramp, vgrid, internal, *[state[n] for n in L.FIELDS])
    return dict(zip(L.FIELDS, outs[:-1])), outs[-1]


def ramp_value(time_days: float, period_days: float, lramp: bool = True) -> float:
    """time_loop.py:79-88"""
    if not lramp:
        return 1.0
    return min(time_days / float(period_days), 1.0)
