"""3D module with the reference's names (the SWIG module pyparm.d3 / sim3d):
OriginBox, AtomVec, NeighborList, EpsSigExpAtom, Repulsion, IEpsSigCutAtom,
LJAttractRepulse, CollectionVerlet, CollectionDamped, ... on the MI355X device
path.  Needs the built extension parm_amd/_d3*.so (no CPU fallback)."""
from ._d3 import *  # noqa: F401,F403
