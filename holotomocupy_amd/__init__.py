"""holotomocupy_amd -- MI355X-native operator stack under the holotomography
Bilinear-Hessian solver (Fresnel D/DT, B-spline shift family, USFFT Radon,
fused BH gradient), drop-in for the corresponding holotomocupy classes.

Device arrays are torch tensors on a ROCm GPU; compute runs in hand-written
gfx950 HIP kernels behind the C ABI of libhtamd.so.  There is no CPU fallback.
"""
__version__ = "0.1.0"

from .propagation import Propagation  # noqa: F401
from .shift import Shift  # noqa: F401
from .tomo import Tomo  # noqa: F401
from .chunking import Chunking  # noqa: F401
