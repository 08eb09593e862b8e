"""x2amd — host side of the MI355X path-tracing device behind XPU2's broker/device interface.

The device itself is the C-ABI library libx2dev.so (include/x2dev.h) built from ../csrc for gfx950; this package
mirrors the reference's GpuLane interface over it (lane.GpuLane) and makes the synthetic workloads (scenes)."""
from . import abi, scenes
from .lane import GpuLane, render_frame

__all__ = ["abi", "scenes", "GpuLane", "render_frame"]
