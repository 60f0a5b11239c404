"""vidyut3d_amd -- MI355X (gfx950) implementation of Vidyut's per-timestep
hot path behind a C ABI (include/vidyut3d_amd.h)."""
__version__ = "0.1.0"
