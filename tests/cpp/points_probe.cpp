// Prints the host-side evaluation points of the drop-in C++ layer for a
// mesh read from an SMF file: every element's quadrature points x(xi_q)
// (what bodyForceComputation and a conductivity function see) and every
// boundary face's surface points x(eta_q) with their unit normals (what
// neumannForceComputation's function sees), for a field of the given
// degree and npts Gauss points per direction.  Test helper: the parity test
// compares them with the same points formed independently in Python.
// Usage: points_probe mesh.smf dim degree npts
#include <cstdio>
#include <cstdlib>

#include "insilico/driver.hpp"

int main(int argc, char* argv[])
{
    if (argc != 5) return 2;
    try {
        const int dim = std::atoi(argv[2]), degree = std::atoi(argv[3]), npts = std::atoi(argv[4]);
        const insilico::Mesh m = insilico::readSMF(std::string(argv[1]), dim);
        const insilico::Field f = insilico::generateField(m, degree, 1);
        std::vector<double> w, xi, ws, eta;
        insilico::detail_::quadratureRule(dim, npts, w, xi);
        insilico::detail_::quadratureRule(dim - 1, npts, ws, eta);
        std::printf("E %zu %zu\n", m.numElements(), w.size());
        for (std::size_t e = 0; e < m.numElements(); e++)
            for (std::size_t q = 0; q < w.size(); q++) {
                const auto x = insilico::detail_::q1Point(m, e, &xi[q * dim]);
                std::printf("%.17g %.17g %.17g\n", x[0], x[1], x[2]);
            }
        const auto faces = insilico::boundaryFieldFaces(m, f);
        std::printf("F %zu %zu %zu\n", faces.size(), ws.size(), faces.empty() ? 0 : faces[0].size());
        for (const auto& fc : faces) {
            for (int32_t v : fc) std::printf("%d ", v);
            std::printf("\n");
        }
        for (const auto& fc : faces)
            for (std::size_t q = 0; q < ws.size(); q++) {
                std::array<double, 3> x, n;
                insilico::detail_::facePoint(m, f, fc, &eta[q * (dim - 1)], x, n);
                std::printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", x[0], x[1], x[2], n[0], n[1],
                            n[2]);
            }
    } catch (const insilico::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
        return 1;
    }
    return 0;
}
