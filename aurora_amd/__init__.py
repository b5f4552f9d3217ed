"""aurora_amd: the Aurora renderer's render core (NMH build, traversal, direct
lighting, path loop) as hand-written HIP kernels for MI355X (gfx950), behind a
C ABI (include/aurora_amd.h) that mirrors the reference's kernel interface."""
from . import scenes  # noqa: F401

__all__ = ["scenes", "native", "renderer"]
