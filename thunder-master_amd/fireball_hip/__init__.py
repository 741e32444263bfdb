"""MI355X (gfx950) two- and three-centre H/S assembler for FIREBALL.

``lib.Assembler`` drives the HIP library through its C ABI
(include/fireball_hip.h); ``synth`` builds synthetic species, Fdata tables
and structures; ``fdata_io`` writes them in the reference's on-disk formats.
"""
from . import synth  # noqa: F401
