"""vlp2d_amd: gfx950 multigrid Dirichlet solver for the 5-point Poisson /
Laplace problem on a rectangle, behind the C ABI of include/vlp2d_mg.h.

The numerical work runs in libvlp2d_mg.so (hand-written HIP kernels).  This
package holds the Python side of the boundary: the ctypes binding
(``_lib``), a host mirror of the reference solver-method interface
(``scheme.MultiGridSchemeHIP``, after VisualLaplacePoisson2D/Model/CPU/
IScheme.cs:16-45), and the synthetic problem generators (``problems``).
"""
from ._lib import VLP2DError, load_library  # noqa: F401
from .scheme import MultiGridSchemeHIP, Options  # noqa: F401
