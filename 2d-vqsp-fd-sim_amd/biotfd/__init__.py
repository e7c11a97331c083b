"""biotfd - MI355X-native drop-in for the Biot velocity-stress time loop.

Host-side Python mirror of the reference's time-loop interface: input/output
file formats (:mod:`biotfd.formats`), synthetic model preparation
(:mod:`biotfd.models`) and the ctypes binding of the C-ABI library
(:mod:`biotfd.lib`).
"""
from . import formats, models  # noqa: F401
