"""parallem_amd — MI355X (gfx950) landscape-evolution timestep.

The per-timestep hot path of PARALLEM (flow direction with plateau and sink
routing, runoff and flow accumulation, erosion/deposition/weathering and
surface updates) as hand-written HIP kernels behind the C-ABI in
include/lemhip.h, loaded here through ctypes (parallem_amd.lemhip).
"""
from . import synth  # noqa: F401

__all__ = ["synth", "lemhip"]
