"""fastsolver_amd: HODLR factor-and-solve on one MI355X (gfx950).

Native pieces (built in-tree by ``python -m fastsolver_amd.build``):
  libfastsolver_amd.so      HIP kernels, level schedule and C-ABI (include/fastsolver_amd.h)
  libfastsolver_amd_gen.so  host-side input generation (include/fastsolver_amd_gen.h)
"""
__version__ = "0.1.0"

from . import gen  # noqa: F401
