"""Locating and loading the native libraries (ctypes).

Loading fails loudly when a library is missing: there is no Python fallback
for any native entry point.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
HOST_LIB = os.path.join(LIB_DIR, "libglv_host.so")
HIP_LIB = os.path.join(LIB_DIR, "libgrappolo_hip.so")

_cache = {}


def _load(path, what):
    if path in _cache:
        return _cache[path]
    if not os.path.exists(path):
        raise RuntimeError(
            f"{what} not built: {path} is missing (run __graft_entry__.build() or "
            f"`make -C grappolo_amd/csrc`)")
    # local scope: the engine exports the reference's C++ names (drop-in), which
    # must not interpose on another library that defines them too
    lib = ctypes.CDLL(path)
    _cache[path] = lib
    return lib


def host_lib():
    return _load(HOST_LIB, "libglv_host.so")


def hip_lib():
    if HIP_LIB not in _cache:
        # torch first: the HIP runtime it loads is the one this library binds
        # to (loaded the other way round, a second runtime comes up that sees
        # no device once torch has initialised its own)
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    return _load(HIP_LIB, "libgrappolo_hip.so")
