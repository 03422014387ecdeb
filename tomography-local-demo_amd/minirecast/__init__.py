"""minirecast on AMD Instinct MI355X: streaming FDK/FBP slice reconstruction with
hand-written gfx950 HIP kernels behind the reference's Python API."""
from .geometry import ParallelGeometry, ConeVecGeometry, circular_cone_vectors
from .recon import SliceReconstructor, HAVE_NUMBA

__version__ = "0.1.0"
