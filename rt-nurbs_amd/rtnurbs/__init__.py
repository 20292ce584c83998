"""rtnurbs: MI355X-native NURBS ray tracing (per-pixel BVH traversal with
Newton-Raphson rational Bezier patch intersection), Python host over the
C ABI of librtnurbs.so (rtnurbs.lib)."""
from . import scenes  # noqa: F401  (pure numpy)
