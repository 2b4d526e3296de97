"""pytest plugin of the sanitizer run (run.sh): the asset and image tests
call the AddressSanitizer + UBSan build of the parsers (libhr_host.so, the
same image_io.cpp / assets.cpp sources as libhr.so) through the asset
module's library hook, and the oracle tests load the sanitizer build of the
oracle (HR_ORACLE_LIB, read by oracle/pyoracle.py)."""
import ctypes as C
import os


def _host_lib(path):
    from vulkanhybridraytracer_amd import abi
    L = C.CDLL(path)
    vp, u32, sz, st = C.c_void_p, C.c_uint32, C.c_size_t, C.c_int
    sigs = {
        "hr_image_load": (st, [C.c_char_p, C.POINTER(abi.hr_image), C.c_char_p, sz]),
        "hr_image_decode": (st, [vp, sz, C.POINTER(abi.hr_image), C.c_char_p, sz]),
        "hr_image_free": (None, [C.POINTER(abi.hr_image)]),
        "hr_asset_load": (st, [C.c_char_p, u32, C.POINTER(vp), C.c_char_p, sz]),
        "hr_asset_view_get": (st, [vp, C.POINTER(abi.hr_asset_view)]),
        "hr_asset_destroy": (None, [vp]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


def pytest_configure(config):
    path = os.environ.get("HR_SAN_HOST_LIB")
    if not path:
        raise RuntimeError("HR_SAN_HOST_LIB is not set: run tests/sanitize/run.sh")
    from vulkanhybridraytracer_amd import assets
    lib = _host_lib(path)
    assets._L = lambda: lib
    config.addinivalue_line("markers", "sanitized: ran with the sanitizer builds")
