"""MI355X-native heat-method geodesics and cortical wiring metrics.

Drop-in backend for FastCorticalWiring's distance-engine adapter
(distance_engines.py:54-69) plus a GPU counterpart of its per-source wiring
loop (core_analysis.py:1441-1672). Compute runs in libfcw.so (hand-written
gfx950 HIP kernels behind a C ABI, include/fcw.h); Python does the once-per-mesh
host preparation and result layout.
"""

__version__ = "0.1.0"

from . import meshes, mesh_prep  # noqa: F401  (numpy only; no GPU needed to import)

__all__ = ["meshes", "mesh_prep", "engine", "analysis", "distance_engines"]
