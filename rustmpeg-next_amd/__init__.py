"""rustmpeg-next_amd: MI355X (gfx950) accelerator for VVC reconstruction stages.

Python mirror of the accelerator surface over librmh (include/rmh.h).
"""
from . import abi  # noqa: F401
