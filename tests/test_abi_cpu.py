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
    for name in ("poisson", "linear_elastic", "compressible", "compressible_q2", "diffusion", "serve"):
        exe = os.path.join(apps, name)
        assert os.path.exists(exe), exe
    for name in ("linear_elastic", "compressible", "compressible_q2", "diffusion"):
        out = subprocess.run([os.path.join(apps, name)], capture_output=True, text=True, timeout=60)
        assert out.returncode == 0 and "Usage" in out.stdout


def test_cpp_layer_q2_numbering_matches_host_setup(tmp_path, golden):
    """The drop-in layer's generateField (degree 2) numbers the DoF objects
    of square_20.smf exactly as the Python host setup does (which
    reproduces reference/03-doFHandler/sparsity.2.ref.dat), and puts the
    same support points; compiled with g++, run without a device."""
    import subprocess
    import numpy as np
    from insilico_amd import problems as P
    src = tmp_path / "q2num.cpp"
    src.write_text(r'''
#include <cstdio>
#include "insilico/insilico.hpp"
#include "insilico/post.hpp"
int main(int argc, char** argv) {
    insilico::Mesh m = insilico::readSMF(argv[1], 2);
    insilico::Field f = insilico::generateField(m, 2, 1);
    std::printf("%zu\n", f.numNodes());
    for (std::size_t i = 0; i < f.conn.size(); i++) std::printf("%d\n", f.conn[i]);
    for (std::size_t i = 0; i < f.support.size(); i++) std::printf("%.17g\n", f.support[i]);
    return 0;
}
''')
    exe = tmp_path / "q2num"
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "insilico_amd", "include")]
    subprocess.run(["g++", "-O1", "-std=c++17"] + inc + [str(src), "-o", str(exe),
                    "-L" + os.path.join(ROOT, "insilico_amd"), "-linsilico_hip",
                    "-Wl,-rpath," + os.path.join(ROOT, "insilico_amd")], check=True, timeout=120)
    smf = os.path.join(golden, "square_20.smf")
    out = subprocess.run([str(exe), smf], capture_output=True, text=True, timeout=60, check=True)
    vals = out.stdout.split()
    n = int(vals[0])
    xyz, conn = P.read_smf(smf, 2)
    fconn, fxyz = P.q2_field(xyz, conn, 2)
    assert n == fxyz.shape[0]
    c = np.array([int(v) for v in vals[1:1 + fconn.size]]).reshape(fconn.shape)
    assert np.array_equal(c, fconn)
    sp = np.array([float(v) for v in vals[1 + fconn.size:]]).reshape(fxyz.shape)
    assert np.array_equal(sp, fxyz)
