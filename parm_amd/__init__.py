"""parm_amd -- MI355X (gfx950) device path for ParM's per-timestep MD hot path.

``parm_amd.capi``  ctypes binding of the C ABI (include/parm_hip.h).
``parm_amd.d3`` / ``parm_amd.d2``  the reference's class names (AtomVec,
OriginBox, NeighborList, LJAttractRepulse, Repulsion, CollectionVerlet, ...)
over that ABI, for 3D and 2D.
"""
__all__ = ["capi"]
