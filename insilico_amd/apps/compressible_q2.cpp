// reference/06-elastic/compressible.cpp on the drop-in layer: a block of
// compressible neo-Hookean material with a degree-2 field (geomDeg 1,
// fieldDeg 2, kernelDegEstimate 3) on the unit square/cube, x = 0 clamped
// and u_x prescribed on x = 1 (reference/06-elastic/PulledSheet.hpp,
// displacement control), load steps with the (step+1)/step ramp and full
// Newton: a fresh solver per iteration, residual forces and the tangent
// with incremental lifting, |F| = solver.norm(), stop when |F| < tol * E,
// cgSolve, addToDoFsFromSolver, |x| = solver.norm(), stop when |x| < tol.
// Usage: compressible_q2 dim N E nu pull loadSteps maxIter tolerance
// Prints the reference's table: step iter |F| |x|
#include <cstdio>
#include <cstdlib>

#include "insilico/insilico.hpp"

int main(int argc, char* argv[])
{
    using namespace insilico;
    if (argc != 9) {
        std::printf("Usage: %s dim N E nu pull loadSteps maxIter tolerance\n", argv[0]);
        return 0;
    }
    const int dim = std::atoi(argv[1]);
    const int N = std::atoi(argv[2]);
    const double E = std::atof(argv[3]), nu = std::atof(argv[4]), pull = std::atof(argv[5]);
    const int loadSteps = std::atoi(argv[6]), maxIter = std::atoi(argv[7]);
    const double tolerance = std::atof(argv[8]);
    try {
        Mesh mesh = unitCube(dim, N, N, N);
        Field field = generateField(mesh, 2, dim);   // base::dof::generate<FEBasis>
        const double firstPull = pull / static_cast<double>(loadSteps);
        const double coordTol = 1.e-5;
        constrainBoundary(mesh, field, [&](const std::array<double, 3>& x, std::size_t node, Field& f) {
            if (std::abs(x[0] - 0.) < coordTol)
                for (int d = 0; d < dim; d++)
                    if (f.isActive(node, d)) f.constrainValue(node, d, 0.0);
            if (std::abs(x[0] - 1.) < coordTol && f.isActive(node, 0)) f.constrainValue(node, 0, firstPull);
        });
        const std::size_t numDofs = numberDoFsConsecutively(field);
        std::printf("# Number of dofs %zu\n", numDofs);
        std::printf("#  Step  Iter  |F|          |x|  \n");

        typedef mat::hypel::NeoHookeanCompressible Material;
        typedef base::FieldBinder<Mesh, Field>::TupleBinder<1, 1>::Type FTB;
        base::FieldBinder<Mesh, Field> fieldBinder(mesh, field);
        solid::HyperElastic<Material, FTB> hyperElastic(Material(mat::Lame::lambda(E, nu),
                                                                 mat::Lame::mu(E, nu)));
        base::solver::Hip solver(numDofs);
        for (int step = 0; step < loadSteps; step++) {
            const double pullFactor = (step == 0 ? static_cast<double>(step + 1)
                                                 : static_cast<double>(step + 1) / static_cast<double>(step));
            scaleConstraints(field, pullFactor);
            unsigned iter = 0;
            while ((int)iter < maxIter) {
                solver.reset();
                if (dim == 2) {
                    base::Quadrature<3, base::QUAD> quadrature;
                    base::asmb::computeResidualForces<FTB>(quadrature, solver, fieldBinder, hyperElastic);
                    base::asmb::stiffnessMatrixComputation<FTB>(quadrature, solver, fieldBinder, hyperElastic);
                } else {
                    base::Quadrature<3, base::HEX> quadrature;
                    base::asmb::computeResidualForces<FTB>(quadrature, solver, fieldBinder, hyperElastic);
                    base::asmb::stiffnessMatrixComputation<FTB>(quadrature, solver, fieldBinder, hyperElastic);
                }
                solver.finishAssembly();
                const double conv1 = solver.norm();
                if (conv1 < tolerance * E) {
                    std::printf("%3d %5u %15.6g\n", step, iter, conv1);
                    break;
                }
                solver.cgSolve(1e-12);
                base::dof::addToDoFsFromSolver(solver, field);
                const double conv2 = solver.norm();
                std::printf("%3d %5u %15.6g %15.6g\n", step, iter, conv1, conv2);
                iter++;
                if (conv2 < tolerance) break;
            }
            if ((int)iter == maxIter)
                std::printf("# (WW) Step %d has not converged within %d iterations \n", step, maxIter);
        }
    } catch (const Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
        return 1;
    }
    return 0;
}
