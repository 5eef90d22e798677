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
    for name in ("poisson", "linear_elastic", "compressible", "compressible_q2", "mixed_poisson",
                 "diffusion", "serve"):
        exe = os.path.join(apps, name)
        assert os.path.exists(exe), exe
    for name in ("linear_elastic", "compressible", "compressible_q2", "mixed_poisson", "diffusion"):
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


def test_cpp_vtk_writer_reproduces_reference_layout(tmp_path, golden):
    """base::io::vtk::LegacyWriter of the drop-in layer: the grid of
    square_20.smf and the disp / sigma data of
    reference/06-elastic/quad.020.Dref.vtk, fed back as values, come out as
    the reference's file line for line (only the free-text title line
    differs); the Q2 boundary faces and the node values of a degree-2 field
    follow the host setup.  Compiled with g++, run without a device."""
    import subprocess
    import numpy as np
    ref_path = os.path.join(golden, "quad.020.Dref.vtk")
    with open(ref_path) as fh:
        ref = fh.read().splitlines()
    disp = [float(v) for v in "\n".join(ref).split("VECTORS disp float")[1].split("CELL_DATA")[0].split()]
    sig = [float(v) for v in "\n".join(ref).split("TENSORS sigma float")[1].split()]
    vals = tmp_path / "vals.txt"
    vals.write_text(" ".join(repr(v) for v in disp) + "\n" + " ".join(repr(v) for v in sig) + "\n")
    src = tmp_path / "vtkw.cpp"
    src.write_text(r'''
#include <cstdio>
#include <fstream>
#include <sstream>
#include "insilico/insilico.hpp"
#include "insilico/post.hpp"
int main(int argc, char** argv) {
    insilico::Mesh m = insilico::readSMF(argv[1], insilico::smfDimension(argv[1]));
    std::ifstream in(argv[2]);
    std::string l1, l2;
    std::getline(in, l1);
    std::getline(in, l2);
    // disp: 3 values per node in the file, the field holds dim = 2
    insilico::Field f = insilico::generateField(m, 1, 2);
    std::istringstream s1(l1), s2(l2);
    for (std::size_t n = 0; n < m.numNodes(); n++) {
        double a, b, c;
        s1 >> a >> b >> c;
        f.value[2 * n] = a;
        f.value[2 * n + 1] = b;
    }
    std::vector<double> sigma;
    double v;
    while (s2 >> v) sigma.push_back(v);
    std::ofstream out(argv[3]);
    base::io::vtk::LegacyWriter w(out);
    w.writeUnstructuredGrid(m);
    base::io::vtk::writePointData(w, m, f, "disp");
    base::io::vtk::writeCellData(w, sigma, 9, "sigma");
    // Q2 boundary faces: count, then the nodes of the first face
    insilico::Field q2 = insilico::generateField(m, 2, 2);
    auto faces = insilico::boundaryFieldFaces(m, q2);
    std::printf("%zu %zu\n", faces.size(), faces[0].size());
    for (std::size_t i = 0; i < q2.value.size(); i++) q2.value[i] = (double)i;
    auto nv = base::post::evaluateFieldAtNodes(m, q2);
    for (double x : nv) std::printf("%.17g\n", x);
    std::printf("%s\n", base::io::baseName("dir/quad.020.smf", ".smf").c_str());
    std::printf("%s\n", base::io::leadingZeros(5).c_str());
    return 0;
}
''')
    exe = tmp_path / "vtkw"
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "insilico_amd", "include")]
    subprocess.run(["g++", "-O1", "-std=c++17"] + inc + [str(src), "-o", str(exe),
                    "-L" + os.path.join(ROOT, "insilico_amd"), "-linsilico_hip",
                    "-Wl,-rpath," + os.path.join(ROOT, "insilico_amd")], check=True, timeout=120)
    smf = os.path.join(golden, "square_20.smf")
    outp = tmp_path / "out.vtk"
    res = subprocess.run([str(exe), smf, str(vals), str(outp)], capture_output=True, text=True,
                         timeout=60, check=True)
    got = outp.read_text().splitlines()
    assert len(got) == len(ref)
    diff = [(i, a, b) for i, (a, b) in enumerate(zip(got, ref)) if a != b and i != 1]
    assert not diff, diff[:5]
    lines = res.stdout.split("\n")
    nf, fn = map(int, lines[0].split())
    assert (nf, fn) == (80, 3)
    from insilico_amd import problems as P
    xyz, conn = P.read_smf(smf, 2)
    fconn, _ = P.q2_field(xyz, conn, 2)
    nodev = np.array([float(x) for x in lines[1:1 + 2 * xyz.shape[0]]]).reshape(-1, 2)
    want = np.zeros_like(nodev)
    want[conn[:, :4].ravel()] = np.stack([2 * fconn[:, :4].ravel(), 2 * fconn[:, :4].ravel() + 1], 1)
    assert np.array_equal(nodev, want)
    assert lines[1 + 2 * xyz.shape[0]] == "dir/quad.020"
    assert lines[2 + 2 * xyz.shape[0]] == "0005"


def test_sgf_ingest_and_structured_vtk(tmp_path, golden):
    """SGF ingest (base/io/sgf/Reader.hpp): reference/01-mesh/squareGrid.sgf
    and cubeGrid.sgf read as the Q1 meshes of the unit square / cube (nodes
    lexicographic, cells lexicographic, vertices hierarchic), in Python and
    in the C++ layer; the C++ structured-grid writer at resolution 2
    reproduces reference/01-mesh/squareGrid.ref.vtk line for line (title
    aside).  Compiled with g++, run without a device."""
    import subprocess
    import numpy as np
    from insilico_amd import problems as P
    for name, dim in (("squareGrid.sgf", 2), ("cubeGrid.sgf", 3)):
        xyz, conn, sizes = P.read_sgf(os.path.join(golden, name))
        assert sizes == (10,) * dim
        uxyz, uconn = P.unit_cube(dim, 10)
        assert np.array_equal(conn, uconn)
        assert np.abs(xyz - uxyz).max() < 1e-12
    src = tmp_path / "sgf.cpp"
    src.write_text(r'''
#include <cstdio>
#include <fstream>
#include "insilico/insilico.hpp"
#include "insilico/post.hpp"
int main(int argc, char** argv) {
    std::array<int, 3> sizes;
    insilico::Mesh m = insilico::readSGF(argv[1], &sizes);
    std::ofstream out(argv[2]);
    base::io::vtk::LegacyWriter w(out);
    w.writeStructuredGrid(m, sizes, 2);
    std::printf("%d %zu %zu %d %d %d\n", m.dim, m.numNodes(), m.numElements(), sizes[0], sizes[1], sizes[2]);
    for (std::size_t i = 0; i < m.conn.size(); i++) std::printf("%d\n", m.conn[i]);
    return 0;
}
''')
    exe = tmp_path / "sgf"
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "insilico_amd", "include")]
    subprocess.run(["g++", "-O1", "-std=c++17"] + inc + [str(src), "-o", str(exe),
                    "-L" + os.path.join(ROOT, "insilico_amd"), "-linsilico_hip",
                    "-Wl,-rpath," + os.path.join(ROOT, "insilico_amd")], check=True, timeout=120)
    outp = tmp_path / "squareGrid.vtk"
    res = subprocess.run([str(exe), os.path.join(golden, "squareGrid.sgf"), str(outp)],
                         capture_output=True, text=True, timeout=60, check=True)
    head = res.stdout.split("\n")[0].split()
    assert head == ["2", "121", "100", "10", "10", "0"]
    _, uconn = P.unit_cube(2, 10)
    assert [int(v) for v in res.stdout.split("\n")[1:1 + uconn.size]] == uconn.ravel().tolist()
    with open(os.path.join(golden, "squareGrid.ref.vtk")) as fh:
        ref = fh.read().splitlines()
    got = outp.read_text().splitlines()
    assert len(got) == len(ref)
    diff = [(i, a, b) for i, (a, b) in enumerate(zip(got, ref)) if a != b and i != 1]
    assert not diff, diff[:5]
