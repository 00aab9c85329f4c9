"""MI355X-native differentiable Princeton Ocean Model step (NeuralPOM's hot path).

The compute path is libneuralpom_amd.so (hand-written HIP kernels for gfx950,
C ABI in include/neuralpom_amd.h); the Python layer mirrors the reference's
dynamical-core interface (src/neuralpom/cores/explicit/core.py:17-61).
"""
__all__ = ["explicit"]
