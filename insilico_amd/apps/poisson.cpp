// Poisson problem -Lap u = 2 pi^2 sin(pi x) sin(pi y) on the unit square,
// u = 0 on the boundary, Q1 on an N x N mesh, written against the drop-in API
// in the order of the reference's heat::PoissonDriver (heat/PoissonDriver.hpp
// :185-210) and reference/04-heat/dirichlet.cpp:120-173.
// Usage: poisson [N [rtol]] (rtol of the PCG, default 1e-10; one below
// FP64 round-off makes cgSolve report INS_ERR_NOCONV, exit status 2).
// Prints: numDoFs iterations relres max|u_h - u| L2(u_h - u)
#include <cstdio>
#include <cstdlib>

#include "insilico/insilico.hpp"
#include "insilico/post.hpp"

int main(int argc, char* argv[])
{
    using namespace insilico;
    const int N = argc > 1 ? std::atoi(argv[1]) : 64;
    const double rtol = argc > 2 ? std::atof(argv[2]) : 1e-10;
    try {
        Mesh mesh = unitCube(2, N, N);
        Field field(mesh.numNodes(), 1);
        constrainBoundary(mesh, field, [](const std::array<double, 3>&, std::size_t node, Field& f) {
            f.constrainValue(node, 0, 0.0);
        });
        const std::size_t numDoFs = numberDoFsConsecutively(field);

        base::Quadrature<2, base::QUAD> quadrature;   // 2*fieldDeg, PoissonDriver.hpp:191
        base::FieldBinder<Mesh, Field> fieldBinder(mesh, field);
        typedef base::FieldBinder<Mesh, Field>::TupleBinder<1, 1>::Type FTB;
        // heat::PoissonDriver's kernel (heat/PoissonDriver.hpp:78, :89-104)
        typedef mat::thermal::IsotropicConstant Material;
        heat::Static<Material, FTB> kernel{Material(1.0)};

        base::solver::Hip solver(numDoFs);
        base::asmb::stiffnessMatrixComputation<FTB>(quadrature, solver, fieldBinder, kernel);
        base::asmb::bodyForceComputation<FTB>(
            quadrature, solver, fieldBinder,
            base::asmb::SineProductForce{2. * M_PI * M_PI, {M_PI, M_PI, 0.}});
        solver.finishAssembly();
        const int iters = solver.cgSolve(rtol);
        base::dof::setDoFsFromSolver(solver, field);

        double maxErr = 0.;
        for (std::size_t n = 0; n < mesh.numNodes(); n++) {
            const double x = mesh.xyz[2 * n], y = mesh.xyz[2 * n + 1];
            maxErr = std::max(maxErr, std::abs(field.value[n] - std::sin(M_PI * x) * std::sin(M_PI * y)));
        }
        const double l2 = base::post::errorComputation<0>(
            quadrature, mesh, field, SineProductSolution{1.0, {M_PI, M_PI, 0.}});
        std::printf("%zu %d %.6e %.6e %.6e\n", numDoFs, iters, solver.lastRelres(), maxErr, l2);
    } catch (const Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
        return e.status == INS_ERR_NOCONV ? 2 : 1;
    }
    return 0;
}
