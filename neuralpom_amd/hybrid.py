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
"""
from __future__ import annotations

import torch
import torch.nn as nn
import torch.nn.functional as F

from . import coupling
from .state import INPUT_CHANNELS, INPUT_PACK_ORDER, pack_input


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
