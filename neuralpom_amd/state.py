"""State movement around the step on the HIP extension.

* ``pack_input`` / ``PackInput``: the corrector network's 101-channel NCHW
  input per member (src/neuralpom/hybrid/state.py:181-226: pack order
  u ub v vb t s on all 16 levels, then el ua uab va vab, each normalised
  (x - mean) / scale with the registered per-cell statistics), differentiable
  back to the state.
* ``inject_state``: write a member's state from a dict of tensors with the
  reference's fallbacks (inject_core_state, state.py:228-306) into an
  ensemble's packed buffers.
* ``restrict4``: the volume-weighted conservative 4 x 4 restriction of fine
  members onto the coarse forecast grid (paper/paper_zh.md:625).
* ``SnapshotRing``: device-resident copies of a packed state, one launch per
  capture or restore (capture / restore of src/neuralpom/v1/checkpoint.py:51-95
  without the host round trip).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L

SUPERVISED_KB = 16
INPUT_THREED_FIELDS = ("u", "ub", "v", "vb", "t", "s")
INPUT_TWOD_FIELDS = ("el", "ua", "uab", "va", "vab")
INPUT_PACK_ORDER = INPUT_THREED_FIELDS + INPUT_TWOD_FIELDS
INPUT_CHANNELS = len(INPUT_THREED_FIELDS) * SUPERVISED_KB + len(INPUT_TWOD_FIELDS)

STATE_FALLBACKS = {"ub": "u", "vb": "v", "tb": "t", "sb": "s", "elb": "el", "uab": "ua", "vab": "va",
                   "q2b": "q2", "q2lb": "q2l"}
OPTIONAL_ZERO_FIELDS = ("drx2d", "dry2d")


def _stream(device):
    return torch.cuda.current_stream(device).cuda_stream


def _stats(ens, normalizer):
    """Per-field (mean, scale) device arrays of one member's shape, or (None, None)."""
    if normalizer is None:
        return None, None
    means, scales = [], []
    for n in INPUT_PACK_ORDER:
        m = normalizer.means.get(n) if hasattr(normalizer, "means") else None
        sc = normalizer.scales.get(n) if hasattr(normalizer, "scales") else None
        if m is None:
            means.append(None)
            scales.append(None)
            continue
        im, jm = ens.shapes[0]
        want = im * jm * (SUPERVISED_KB if n in INPUT_THREED_FIELDS else 1)
        m = m.to(ens.device, ens.dtype).contiguous().reshape(-1)
        sc = sc.to(ens.device, ens.dtype).contiguous().reshape(-1)
        if m.numel() != want or sc.numel() != want:
            raise ValueError(f"statistics of {n} do not match the member shape {(im, jm)}")
        means.append(m)
        scales.append(sc)
    return means, scales


def pack_input(ens, state, normalizer=None):
    """(101 x im x jm per member, members back to back) float tensor of the
    ensemble's state; ``normalizer`` has ``means`` / ``scales`` dicts of
    per-cell statistics (the reference's StateNormalizer) or is None."""
    means, scales = _stats(ens, normalizer)
    return PackInput.apply(ens, means, scales, *[state[n] for n in INPUT_PACK_ORDER])


def member_planes(ens, packed, m, channels=INPUT_CHANNELS):
    """View of member m's (channels, im, jm) planes of a packed NCHW tensor."""
    im, jm = ens.shapes[m]
    o = channels * ens.off2[m]
    return packed[o:o + channels * im * jm].view(channels, im, jm)


class PackInput(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ens, means, scales, *fields):
        fl = [f.contiguous() for f in fields]
        ens._check(fl, INPUT_PACK_ORDER, ens.numel)
        out = torch.empty(INPUT_CHANNELS * ens.n2, dtype=ens.dtype, device=ens.device)
        mp = L.ptr_array([m.data_ptr() if m is not None else None for m in means]) if means else None
        sp = L.ptr_array([s.data_ptr() if s is not None else None for s in scales]) if scales else None
        L.check(ens.lib.pom_pack_input(ens._h, L.ptr_array([f.data_ptr() for f in fl]), mp, sp, out.data_ptr(),
                                       _stream(ens.device)), "pom_pack_input")
        ctx.ens, ctx.scales = ens, scales
        return out

    @staticmethod
    def backward(ctx, g):
        ens = ctx.ens
        grads = [torch.zeros(ens.numel(n), dtype=ens.dtype, device=ens.device) if ctx.needs_input_grad[3 + i]
                 else None for i, n in enumerate(INPUT_PACK_ORDER)]
        sp = L.ptr_array([s.data_ptr() if s is not None else None for s in ctx.scales]) if ctx.scales else None
        L.check(ens.lib.pom_pack_input_adj(ens._h, sp, g.contiguous().data_ptr(),
                                           L.ptr_array([x.data_ptr() if x is not None else None for x in grads]),
                                           _stream(ens.device)), "pom_pack_input_adj")
        return (None, None, None) + tuple(grads)


def inject_state(ens, state, m, values):
    """Write member m's slice of the packed ``state`` (dict of flat tensors)
    from ``values`` (name -> (im, jm[, 16]) tensor), with the reference's
    fallbacks: a missing history level takes its current level, drx2d / dry2d
    are zeroed when absent (state.py:18-28, 86, 228-306)."""
    for name, flat in state.items():
        src = values.get(name)
        if src is None and name in STATE_FALLBACKS:
            src = values.get(STATE_FALLBACKS[name])
        dst = ens.view(flat, m, name)
        if src is None:
            if name in OPTIONAL_ZERO_FIELDS:
                dst.zero_()
            continue
        src = src.detach().to(dst.device, dst.dtype)
        if src.shape != dst.shape:
            if src.numel() != dst.numel():
                continue   # a field of another shape is skipped, as the reference does
            src = src.reshape(dst.shape)
        dst.copy_(src)


def restrict4(fine, coarse, field, name, fine_geom, fine_el=None, fine_vgrid=None, volume=True):
    """Conservative 4 x 4 restriction of the fine batch's ``field`` (flat, the
    layout of field ``name``) onto the coarse batch.  Weights art * mask
    [* (h + el) * dz]; mask is dum for u / ua / ub / uab, dvm for v, fsm else.
    Differentiable in ``field`` and, with volume weights, in ``fine_el``."""
    three_d = name in L.THREE_D
    mask = {"u": "dum", "ub": "dum", "ua": "dum", "uab": "dum",
            "v": "dvm", "vb": "dvm", "va": "dvm", "vab": "dvm"}.get(name, "fsm")
    geom = (fine_geom["art"].detach(), fine_geom[mask].detach(), fine_geom["h"].detach())
    vg = fine_vgrid.detach() if fine_vgrid is not None else None
    return Restrict4.apply(fine, coarse, int(three_d), bool(volume), geom, vg, field.contiguous(), fine_el)


class Restrict4(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fine, coarse, three_d, volume, geom, vg, field, el):
        out = torch.empty(coarse.n3 if three_d else coarse.n2, dtype=coarse.dtype, device=coarse.device)
        el_c = el.contiguous() if el is not None else None
        ptrs = [t.data_ptr() if t is not None else None for t in (field,) + tuple(geom) + (el_c, vg)]
        L.check(fine.lib.pom_restrict4(fine._h, coarse._h, three_d, int(volume), *ptrs, out.data_ptr(),
                                       _stream(fine.device)), "pom_restrict4")
        ctx.fine, ctx.coarse, ctx.three_d, ctx.volume = fine, coarse, three_d, volume
        ctx.has_el = el is not None
        ctx.save_for_backward(field, el_c if el is not None else field, vg if vg is not None else field, *geom)
        ctx.has_vg = vg is not None
        return out

    @staticmethod
    def backward(ctx, g):
        field, el, vg, art, mask, h = ctx.saved_tensors
        el = el if ctx.has_el else None
        vg = vg if ctx.has_vg else None
        gf = torch.zeros_like(field) if ctx.needs_input_grad[6] else None
        gel = torch.zeros_like(el) if (el is not None and ctx.needs_input_grad[7]) else None
        if gf is None and gel is None:
            return (None,) * 8
        gf_buf = gf if gf is not None else torch.zeros_like(field)
        ptrs = [t.data_ptr() if t is not None else None for t in (field, art, mask, h, el, vg)]
        L.check(ctx.fine.lib.pom_restrict4_adj(ctx.fine._h, ctx.coarse._h, ctx.three_d, int(ctx.volume), *ptrs,
                                               g.contiguous().data_ptr(), gf_buf.data_ptr(),
                                               gel.data_ptr() if gel is not None else None,
                                               _stream(ctx.fine.device)), "pom_restrict4_adj")
        return (None,) * 6 + (gf, gel)


def copy_arrays(dst, src):
    """dst[i].copy_(src[i]) for device tensors of equal byte size, one launch."""
    if len(dst) != len(src):
        raise ValueError("dst and src lists differ in length")
    for d, s in zip(dst, src):
        if d.numel() * d.element_size() != s.numel() * s.element_size() or not d.is_contiguous() \
                or not s.is_contiguous():
            raise ValueError("copy_arrays needs contiguous tensors of equal byte size")
    if not dst:
        return
    sizes = (C.c_int64 * len(dst))(*[d.numel() * d.element_size() for d in dst])
    lib = L.load()
    L.check(lib.pom_copy_arrays(len(dst), L.ptr_array([d.data_ptr() for d in dst]),
                                L.ptr_array([s.data_ptr() for s in src]), sizes, _stream(dst[0].device)),
            "pom_copy_arrays")


class SnapshotRing:
    """``depth`` device-resident copies of a packed state dict; ``capture``
    writes the next slot and returns its index, ``restore`` copies a slot back
    into a state dict (or returns fresh copies)."""

    def __init__(self, template, depth=2):
        if depth < 1:
            raise ValueError("depth must be positive")
        self.names = tuple(template)
        self.slots = [{n: torch.empty_like(template[n]) for n in self.names} for _ in range(depth)]
        self.next = 0

    def capture(self, state):
        k = self.next
        copy_arrays([self.slots[k][n] for n in self.names], [state[n].contiguous() for n in self.names])
        self.next = (k + 1) % len(self.slots)
        return k

    def restore(self, k, into=None):
        if into is None:
            into = {n: torch.empty_like(self.slots[k][n]) for n in self.names}
        copy_arrays([into[n] for n in self.names], [self.slots[k][n] for n in self.names])
        return into
