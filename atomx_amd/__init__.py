"""atomx_amd — MI355X-native Slater-integral engine for atomx's B-spline radial engine.

Drop-in modules (same names and signatures as the reference's):
  atomx_amd.yk         <- atomx.numerics.yk
  atomx_amd.slater     <- atomx.numerics.bspline.slater
  atomx_amd.slater_yk  <- atomx.numerics.bspline.slater_yk
  atomx_amd.bspline    <- atomx.numerics.bspline.{knots,basis,evaluate} (grid tables)
Engine-level:
  atomx_amd.tensor     batched all-quartet R^k tensor (device-resident plan)

All compute runs in libatomx_slater.so (hand-written HIP for gfx950) through
the C ABI in include/atomx_slater.h; there is no CPU fallback.
"""

__version__ = "0.1.0"

from . import bspline  # noqa: F401  (host geometry tables, no GPU needed)
