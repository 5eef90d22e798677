"""CPU checks of the boundary: the built library loads without a GPU and
exports exactly the entry points include/insilico_hip.h declares."""
import ctypes
import os
import re

from insilico_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "insilico_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ins_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(abi.LIB_PATH)
    names = _declared()
    assert len(names) >= 29
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(abi.SYMBOLS) == names


def test_binding_loads_and_reports_errors():
    L = abi.load()
    assert L.ins_strerror(-4) == b"index out of range"
    assert L.ins_strerror(0) == b"ok"
    # null arguments are rejected without touching a device
    assert L.ins_create(None, 0, 0) == -1
    assert L.ins_destroy(None) == -1
    assert L.ins_set_mesh(None, 3, 1, None, 1, 8, None) == -1


def test_apps_built_and_print_usage():
    """The example C++ drivers link against the HIP library; without
    arguments they print their usage and touch no device."""
    import subprocess
    apps = os.path.join(ROOT, "insilico_amd", "apps")
    for name in ("poisson", "linear_elastic", "compressible"):
        exe = os.path.join(apps, name)
        assert os.path.exists(exe), exe
    for name in ("linear_elastic", "compressible"):
        out = subprocess.run([os.path.join(apps, name)], capture_output=True, text=True, timeout=60)
        assert out.returncode == 0 and "Usage" in out.stdout
