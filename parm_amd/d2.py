"""2D module with the reference's names (the SWIG module pyparm.d2 / sim2d)
on the MI355X device path.  Needs the built extension parm_amd/_d2*.so."""
from ._d2 import *  # noqa: F401,F403
