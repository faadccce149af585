"""MI355X-native implementation of the Tritok boundary-element hot path.

The compute lives in libtritok_amd.so (HIP for gfx950 + host C++), declared in
include/tritok_bem.h.  This package is the Python mirror of that C-ABI."""
from . import _lib  # noqa: F401  (fails loudly when the library is not built)
from .host import Host  # noqa: F401


def device_info():
    """Name, compute-unit count and HBM bytes of the current device (tt_device_info)."""
    import ctypes as C
    f = _lib._sig("tt_device_info", _lib.i32, C.c_char_p, _lib.i32, C.POINTER(_lib.i32), C.POINTER(_lib.i64))
    name = C.create_string_buffer(256)
    cu, hbm = _lib.i32(), _lib.i64()
    _lib.check(f(name, 256, C.byref(cu), C.byref(hbm)), "tt_device_info")
    return {"name": name.value.decode(errors="replace"), "cus": cu.value, "hbm_bytes": hbm.value}
