"""MI355X-native implementation of the Tritok boundary-element hot path.

The compute lives in libtritok_amd.so (HIP for gfx950 + host C++), declared in
include/tritok_bem.h.  This package is the Python mirror of that C-ABI."""
from . import _lib  # noqa: F401  (fails loudly when the library is not built)
from .host import Host  # noqa: F401
