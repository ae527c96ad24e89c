"""MI355X-native DoonEngine voxel pipeline.

The product is the C-ABI library libdoon.so (include/doon/doon.h): the
reference's chunk/map engine API over hand-written gfx950 HIP kernels for the
draw cast, the visible-chunk pass and the per-voxel lighting pass. This package
holds its sources (csrc/), the build recipe (build.py), ctypes bindings
(_lib.py, volume.py) and the seeded scene generators (scenes.py).
"""
from ._lib import load as load_library  # noqa: F401
