"""MI355X-native KinectFusion / voxel-hashing fusion path.

Native libraries live in ``natureuserinterfacestudio_amd/lib`` (built in tree by
``_build.py``).  Loading fails loudly when they are missing: there is no CPU
fallback for any product entry point.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
# NUIF_LIB_DIR selects another build of the three libraries (the schedule tuner's variants)
LIB_DIR = os.environ.get("NUIF_LIB_DIR") or os.path.join(_PKG, "lib")

_libs = {}


class NativeLibraryMissing(RuntimeError):
    pass


def _load(name):
    if name in _libs:
        return _libs[name]
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` first")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    _libs[name] = lib
    return lib


def libnuif():
    """The HIP kernel library (C ABI of include/nuif.h)."""
    return _load("libnuif.so")


def libkinfu():
    """The C++ engine library above the C ABI."""
    libnuif()
    return _load("libnuif_kinfu.so")


def libsynth():
    """The synthetic workload generator."""
    return _load("libnuif_synth.so")


def build():
    from . import _build

    _build.build_all()
