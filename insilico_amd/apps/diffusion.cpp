// Time-dependent heat equation with heat::DiffusionDriver on the drop-in
// layer (as reference/05-mixedPoisson/diffusionWithDriver.cpp drives it):
// u_t - Lap u = 0 on the unit square, u = 0 on the boundary, u0 = sin(pi x)
// sin(pi y) at the nodes, BDF(order) with its start-up, advanceInTime on the
// device.  After the last step it prints the largest nodal deviation from the
// discrete solution y_n u0 (y_n from the BDF recurrence with the generalized
// eigenvalue of the Q1 mode) and the L2 error against the continuous
// solution exp(-2 pi^2 t) sin(pi x) sin(pi y) (errorComputation<0>).
// Usage: diffusion N order numSteps stepSize
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "insilico/insilico.hpp"
#include "insilico/post.hpp"

template <unsigned Q>
static int run(int N, int numSteps, double dt)
{
    using namespace insilico;
    Mesh mesh = unitCube(2, N, N);
    Field field(mesh.numNodes(), 1);
    constrainBoundary(mesh, field, [](const std::array<double, 3>&, std::size_t node, Field& f) {
        f.constrainValue(node, 0, 0.0);
    });
    numberDoFsConsecutively(field);
    std::vector<double> u0(mesh.numNodes());
    for (std::size_t n = 0; n < mesh.numNodes(); n++) {
        u0[n] = std::sin(M_PI * mesh.xyz[2 * n]) * std::sin(M_PI * mesh.xyz[2 * n + 1]);
        field.value[n] = u0[n];
    }
    heat::DiffusionDriver<base::time::BDF<Q>> diffusion(mesh, field);
    // BDF weights with start-up (base/time/BDF.hpp, base/time/MultiStep.hpp)
    static const double W[2][3] = {{1.0, -1.0, 0.0}, {1.5, -2.0, 0.5}};
    const double h = 1.0 / N;
    const double lam = 2 * 6 / (h * h) * (1 - std::cos(M_PI * h)) / (2 + std::cos(M_PI * h));
    std::vector<double> y{1.0};
    for (int step = 0; step < numSteps; step++) {
        diffusion.advanceInTime((unsigned)step, dt);
        const int q = (Q == 1 || step == 0) ? 1 : 2;
        const double* a = W[q - 1];
        double rhs = 0.;
        for (int s = 1; s <= q; s++) rhs -= a[s] * y[y.size() - s];
        y.push_back(rhs / (a[0] + dt * lam));
        std::printf("%4d %10.6f %3d\n", step, (step + 1) * dt, diffusion.lastIterations());
    }
    double dev = 0.;
    for (std::size_t n = 0; n < mesh.numNodes(); n++)
        dev = std::max(dev, std::abs(field.value[n] - y.back() * u0[n]));
    const double t = numSteps * dt;
    base::Quadrature<2, base::QUAD> quadrature;
    const double l2 = base::post::errorComputation<0>(
        quadrature, mesh, field, SineProductSolution{std::exp(-2 * M_PI * M_PI * t), {M_PI, M_PI, 0.}});
    std::printf("# discrete deviation %.6e\n# y_n %.17g\n# L2 error %.17g\n", dev, y.back(), l2);
    return 0;
}

int main(int argc, char* argv[])
{
    if (argc != 5) {
        std::printf("Usage: %s N order numSteps stepSize\n", argv[0]);
        return 0;
    }
    const int N = std::atoi(argv[1]), order = std::atoi(argv[2]), steps = std::atoi(argv[3]);
    const double dt = std::atof(argv[4]);
    try {
        return order == 2 ? run<2>(N, steps, dt) : run<1>(N, steps, dt);
    } catch (const insilico::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
        return 1;
    }
}
