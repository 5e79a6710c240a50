"""MI355X-native hot path of the point-cloud ray-marching renderer.

librm_hip.so (HIP, gfx950) holds every kernel behind a C ABI (include/rm.h); this package holds
the seeded scene generators and a Python mirror of the reference interface used by the tests
and the benchmark. The reference-shaped C++ host (CLI, PLY/JSON/PNG/.npy I/O, radii cache)
lives in renderer_ray_marching_amd/host/.
"""
from . import scenes  # noqa: F401

__all__ = ["scenes"]
