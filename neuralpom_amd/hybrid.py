"""Differentiable hybrid rollout: physics intervals on the drop-in core with the
endpoint corrector applied after each interval.

* ``CorrectorNet``: the residual U-Net corrector of
  src/neuralpom/models/corrector_net.py:7-100 (GroupNorm residual blocks, three
  max-pool downsamplings, bilinear upsampling with skip concatenation, 1x1
  head).  Module and parameter names are the reference's, so its checkpoints
  load with ``load_state_dict``; the head is zero-initialised as there unless
  ``head_std`` asks for a seeded random head.
* ``advance``: one differentiable physics interval (SparsePhysics.advance with
  differentiable=True, src/neuralpom/hybrid/sparse_physics.py:91-149).  The
  per-step recompute the reference gets from torch.utils.checkpoint is built
  into the step's autograd node (``explicit.ExplicitStep``), so the interval
  runs the core's time loop directly.
* ``corrector_input``: the 101-channel NCHW input (HIP pack,
  src/neuralpom/hybrid/state.py:197-226).
* ``coupled_rollout``: intervals of physics, each followed by the corrector
  hook (``coupling.apply_joint_correction``, sparse_coupling.py:42-67).
* ``integrate`` / ``truncated_windows`` / ``restrict_state``: batched members
  over windows of many steps with the window's loss differentiated back to
  its start and the state detached between windows (the daily truncated
  update, src/neuralpom/v1/posttrain/truncated.py:23-33 and
  rollout.py:57-74), the fine members restricted to the forecast grid for the
  loss.  A window can be long: runs of ``segment`` steps are recomputed in
  backward, so the graph holds about steps / segment + segment states.
"""
from __future__ import annotations

import torch
import torch.nn as nn
import torch.nn.functional as F

from torch.utils.checkpoint import checkpoint

from . import _lib as L
from . import coupling
from .explicit import ExplicitStep
from .implicit import ImplicitEnsemble, ImplicitStep
from .state import INPUT_CHANNELS, INPUT_PACK_ORDER, pack_input, restrict4


def _group_norm(channels: int) -> nn.GroupNorm:
    """GroupNorm with the largest group count <= 32 that divides ``channels``."""
    groups = min(32, channels)
    while channels % groups:
        groups -= 1
    return nn.GroupNorm(groups, channels)


class ResBlock(nn.Module):
    """relu(norm(conv3(relu(norm(conv3(x))))) + shortcut(x)), shortcut a 1x1
    conv + norm when the channel count changes (corrector_net.py:14-38)."""

    def __init__(self, cin: int, cout: int, cmid: int | None = None):
        super().__init__()
        cmid = cmid or cout
        self.double_conv = nn.Sequential(
            nn.Conv2d(cin, cmid, 3, padding=1, bias=False), _group_norm(cmid), nn.ReLU(),
            nn.Conv2d(cmid, cout, 3, padding=1, bias=False), _group_norm(cout))
        self.shortcut = nn.Sequential() if cin == cout else nn.Sequential(
            nn.Conv2d(cin, cout, 1, bias=False), _group_norm(cout))
        self.relu = nn.ReLU()

    def forward(self, x):
        return self.relu(self.double_conv(x) + self.shortcut(x))


class Down(nn.Module):
    """2x2 max pool, then a residual block (corrector_net.py:41-49)."""

    def __init__(self, cin: int, cout: int):
        super().__init__()
        self.maxpool_conv = nn.Sequential(nn.MaxPool2d(2), ResBlock(cin, cout))

    def forward(self, x):
        return self.maxpool_conv(x)


class Up(nn.Module):
    """Bilinear x2 upsampling (corners aligned), zero padding to the skip
    tensor's size, concatenation [skip, up] and a residual block
    (corrector_net.py:52-69)."""

    def __init__(self, cin: int, cout: int):
        super().__init__()
        self.up = nn.Upsample(scale_factor=2, mode="bilinear", align_corners=True)
        self.conv = ResBlock(cin, cout, cin // 2)

    def forward(self, x, skip):
        x = self.up(x)
        dy, dx = skip.shape[2] - x.shape[2], skip.shape[3] - x.shape[3]
        x = F.pad(x, [dx // 2, dx - dx // 2, dy // 2, dy - dy // 2])
        return self.conv(torch.cat([skip, x], dim=1))


class CorrectorNet(nn.Module):
    """Residual U-Net from the 101-channel state to the 60-channel increment
    (corrector_net.py:72-100), widths h, 2h, 4h, 8h."""

    def __init__(self, in_channels: int = INPUT_CHANNELS, out_channels: int = coupling.OUTPUT_CHANNELS,
                 hidden_dim: int = 256, head_std: float = 0.0, generator: torch.Generator | None = None):
        super().__init__()
        h = hidden_dim
        self.inc = ResBlock(in_channels, h)
        self.down1 = Down(h, 2 * h)
        self.down2 = Down(2 * h, 4 * h)
        self.down3 = Down(4 * h, 8 * h)
        self.up1 = Up(8 * h + 4 * h, 4 * h)
        self.up2 = Up(4 * h + 2 * h, 2 * h)
        self.up3 = Up(2 * h + h, h)
        self.outc = nn.Conv2d(h, out_channels, 1)
        with torch.no_grad():
            if head_std > 0.0:   # a seeded head: the reference's zero head makes every correction zero
                self.outc.weight.copy_(torch.randn(self.outc.weight.shape, generator=generator) * head_std)
                self.outc.bias.copy_(torch.randn(self.outc.bias.shape, generator=generator) * head_std)
            else:
                self.outc.weight.zero_()
                self.outc.bias.zero_()

    def forward(self, x):
        x1 = self.inc(x)
        x2 = self.down1(x1)
        x3 = self.down2(x2)
        x4 = self.down3(x3)
        y = self.up1(x4, x3)
        y = self.up2(y, x2)
        y = self.up3(y, x1)
        return self.outc(y)


def advance(core, steps: int) -> None:
    """``steps`` internal steps from ``core.time``, differentiable when the
    core's state requires grad (sparse_physics.py:91-149)."""
    steps = int(steps)
    if steps < 1:
        raise ValueError("steps must be positive")
    core.time0 = float(core.time)
    core.iint = 0
    core.iend = steps
    core.run_time_loop()
    if core.iint != steps:
        raise RuntimeError("Physical core returned before the interval completed")


def corrector_input(core, normalizer=None):
    """(1, 101, im, jm) network input of the core's state."""
    ens = core._ensemble()
    packed = pack_input(ens, {n: core._flat(n) for n in INPUT_PACK_ORDER}, normalizer)
    return packed.view(1, INPUT_CHANNELS, core.im, core.jm)


def coupled_rollout(core, net, scales, intervals: int, steps: int, cap: float = 5.0, normalizer=None):
    """``intervals`` times: ``steps`` physics steps, then the corrector's
    increment through the hook.  Returns the hook's diagnostics per interval."""
    diags = []
    for _ in range(int(intervals)):
        advance(core, steps)
        x = corrector_input(core, normalizer)
        raw = net(x.to(next(net.parameters()).dtype))[0].to(core.dtype)
        diags.append(coupling.apply_joint_correction(core, raw, scales, cap))
    return diags


def _step(ens, state, geom, vgrid, schedule, it):
    sc, ramp, internal = schedule(it)
    ins = [state[n] for n in L.FIELDS]
    if isinstance(ens, ImplicitEnsemble):
        outs = ImplicitStep.apply(ens, geom, sc, ramp, vgrid, internal, None, *ins)
    else:
        outs = ExplicitStep.apply(ens, geom, sc, ramp, vgrid, internal, *ins)
    return dict(zip(L.FIELDS, outs[:-1])), outs[-1]


def _run(ens, geom, vgrid, schedule, first, count, segments, *inputs):
    """Steps first .. first + count - 1; ``segments`` (outermost first) are
    the run lengths recomputed in backward, each level inside the one before."""
    state = dict(zip(L.FIELDS, inputs))
    vamax = []
    if not segments:
        for it in range(first, first + count):
            state, va = _step(ens, state, geom, vgrid, schedule, it)
            vamax.append(va)
    else:
        seg, inner = segments[0], segments[1:]
        it, last = first, first + count
        while it < last:
            n = min(seg, last - it)
            ins = [state[f] for f in L.FIELDS]
            if n > 1 and torch.is_grad_enabled() and any(t.requires_grad for t in ins):
                outs = checkpoint(_run, ens, geom, vgrid, schedule, it, n, inner, *ins, use_reentrant=False)
            else:
                outs = _run(ens, geom, vgrid, schedule, it, n, inner, *ins)
            state = dict(zip(L.FIELDS, outs[:-1]))
            vamax.append(outs[-1])
            it += n
    return tuple(state[f] for f in L.FIELDS) + (torch.stack(vamax).amax(),)


def integrate(ens, state, geom, vgrid, schedule, first: int, steps: int, segment=1):
    """Internal steps ``first`` .. ``first + steps - 1`` (1-based step numbers)
    of every member of ``ens`` (an explicit ``Ensemble`` or an
    ``ImplicitEnsemble``); differentiable in ``state``.

    ``schedule(it)`` gives the step's (scalars, ramp tensor [members],
    internal flag), as the core's time loop forms them.  Each step's autograd
    node keeps its input state and recomputes its forward in backward.  With
    ``segment`` = s > 1, runs of s steps are wrapped in
    torch.utils.checkpoint: the graph keeps only each run's input state and
    rebuilds the run's states when its backward comes, so a window of n steps
    holds about n / s + s states instead of n, for one extra forward of the
    window.  A sequence (s1, s2, ...) nests the runs: runs of s1 steps, each
    recomputed as runs of s2 steps, and so on, about n / s1 + s1 / s2 + ... +
    s_last states for one extra forward per level.  The kernels are
    deterministic, so the gradients are bitwise those of ``segment`` = 1.

    Returns (end state, max barotropic speed of the steps as a device scalar)."""
    steps = int(steps)
    if steps < 1:
        raise ValueError("steps must be positive")
    segs = (segment,) if isinstance(segment, int) else tuple(segment)
    if any(int(x) < 1 for x in segs):
        raise ValueError("segment lengths must be positive")
    segs = tuple(int(x) for x in segs if int(x) > 1)
    outs = _run(ens, geom, vgrid, schedule, int(first), steps, segs, *[state[n] for n in L.FIELDS])
    return dict(zip(L.FIELDS, outs[:-1])), outs[-1].detach()


def restrict_state(fine, coarse, state, geom, vgrid, names):
    """The fine members' fields ``names`` on the coarse members: 3-D fields
    volume-weighted with the fine surface elevation (differentiable in the
    field and in el), 2-D fields area-weighted (the 4 x 4 conservative
    restriction, paper/paper_zh.md:625)."""
    out = {}
    for n in names:
        three_d = n in L.THREE_D
        out[n] = restrict4(fine, coarse, state[n], n, geom, state["el"] if three_d else None,
                           vgrid, volume=three_d)
    return out


def truncated_windows(ens, state, geom, vgrid, schedule, windows: int, steps: int, loss_of,
                      first: int = 1, segment: int = 1, vmaxl: float | None = None):
    """``windows`` windows of ``steps`` internal steps, each trained on its own
    (truncated.py:23-33 with daily windows; rollout.py:57-74 detaches the state
    between them).

    Per window: the start state is detached and made a leaf, the window is
    integrated (``integrate`` with ``segment``), ``loss_of(end_state, window)``
    gives a scalar loss whose backward runs at once, accumulating into the
    ``.grad`` of whatever parameters the loss uses (a corrector's), and the
    end state, detached, starts the next window.  ``vmaxl``: raise
    FloatingPointError when a window's barotropic speed exceeds it, as the
    core's time loop does.

    Returns (losses [windows] as floats, per-window start-state gradients
    {field: tensor} of the differentiable fields, final state)."""
    losses, grads = [], []
    it = int(first)
    diff = set(L.DIFF_INPUTS)
    for w in range(int(windows)):
        start = {n: v.detach().requires_grad_(n in diff) for n, v in state.items()}
        end, vamax = integrate(ens, start, geom, vgrid, schedule, it, steps, segment)
        if vmaxl is not None and not float(vamax) <= vmaxl:   # NaN-safe
            raise FloatingPointError(f"Barotropic velocity {float(vamax)} exceeds limit {vmaxl} in window {w}")
        loss = loss_of(end, w)
        loss.backward()
        losses.append(float(loss.detach()))
        grads.append({n: start[n].grad for n in L.DIFF_INPUTS if start[n].grad is not None})
        state = {n: v.detach() for n, v in end.items()}
        it += int(steps)
    return losses, grads, state
