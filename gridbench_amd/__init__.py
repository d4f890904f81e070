"""gridbench_amd -- MI355X (gfx950) implementation of GridBench's hot path,
the domain-wall-fermion Wilson hopping term (Hand DhopSite over Ls slices),
behind the C-ABI of libgbx.so (include/gbx.h)."""
from .lattice import *  # noqa: F401,F403
from .lattice import __all__  # noqa: F401
from ._lib import LIB_PATH  # noqa: F401
