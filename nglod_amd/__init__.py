"""MI355X-native NGLOD neural-SDF query and sphere-tracing path.

Drop-in classes (same names and arguments as sdf-net/lib): OctreeSDF,
FeatureVolume, SphereTracer, RenderBuffer, Renderer, look_at.  Compute runs
in libnglod_hip.so (hand-written gfx950 kernels behind a C ABI).
"""
from .octree_sdf import FeatureVolume, OctreeSDF, default_args
from .tracer import BaseTracer, RenderBuffer, SphereTracer, setparam
from .renderer import Renderer, look_at, normalized_grid

__all__ = [
    "FeatureVolume", "OctreeSDF", "default_args", "BaseTracer", "RenderBuffer", "SphereTracer", "setparam",
    "Renderer", "look_at", "normalized_grid",
]
