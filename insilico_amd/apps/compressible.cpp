// Compressible neo-Hookean block under displacement control, the load-step
// and Newton loop of reference/06-elastic/compressible.cpp:245-335 with the
// boundary conditions of reference/06-elastic/PulledSheet.hpp (x = 0 fixed,
// u_x prescribed on x = 1), on an N^3 Q1 hex mesh.  Each Newton iteration
// assembles the internal-force residual (FP64) and the tangent with the
// incremental lifting g - u on the device, solves with Jacobi-PCG and adds the
// increment to the DoFs; |F| and |x| are the reference's solver.norm()
// (|b|_2 / N).
// Usage: compressible N loadSteps pull maxIter tolerance [mu lambda] [fp32]
// Prints one line per Newton iteration: step iter |F| |x|
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "insilico/insilico.hpp"

int main(int argc, char* argv[])
{
    using namespace insilico;
    if (argc < 6) {
        std::printf("Usage: %s N loadSteps pull maxIter tolerance [mu lambda] [fp32]\n", argv[0]);
        return 0;
    }
    const int N = std::atoi(argv[1]);
    const int loadSteps = std::atoi(argv[2]);
    const double pull = std::atof(argv[3]);
    const int maxIter = std::atoi(argv[4]);
    const double tolerance = std::atof(argv[5]);
    const double muIn = argc > 7 ? std::atof(argv[6]) : 1.0;
    const double lambdaIn = argc > 7 ? std::atof(argv[7]) : 10.0;
    const bool fp32 = argc > 8 && std::strcmp(argv[8], "fp32") == 0;
    try {
        Mesh mesh = unitCube(3, N, N, N);
        Field field(mesh.numNodes(), 3);
        const double firstPull = pull / static_cast<double>(loadSteps);
        const double coordTol = 1.e-5;
        constrainBoundary(mesh, field, [&](const std::array<double, 3>& x, std::size_t node, Field& f) {
            if (std::abs(x[0] - 0.) < coordTol)
                for (int d = 0; d < 3; d++) f.constrainValue(node, d, 0.0);
            if (std::abs(x[0] - 1.) < coordTol) f.constrainValue(node, 0, firstPull);
        });
        const std::size_t numDofs = numberDoFsConsecutively(field);
        std::printf("# Number of dofs %zu\n", numDofs);

        base::Quadrature<2, base::HEX> quadrature;
        base::FieldBinder<Mesh, Field> fieldBinder(mesh, field);
        typedef base::FieldBinder<Mesh, Field>::TupleBinder<1, 1>::Type FTB;
        typedef mat::hypel::NeoHookeanCompressible Material;
        solid::HyperElastic<Material, FTB> hyperElastic(Material(lambdaIn, muIn));

        base::solver::Hip solver(numDofs);
        if (fp32) solver.setPrecision(INS_FP32);
        for (int step = 0; step < loadSteps; step++) {
            const double pullFactor = (step == 0 ? static_cast<double>(step + 1)
                                                 : static_cast<double>(step + 1) / static_cast<double>(step));
            scaleConstraints(field, pullFactor);
            int iter = 0;
            while (iter < maxIter) {
                solver.reset();
                base::asmb::computeResidualForces<FTB>(quadrature, solver, fieldBinder, hyperElastic);
                base::asmb::stiffnessMatrixComputation<FTB>(quadrature, solver, fieldBinder, hyperElastic);
                solver.finishAssembly();
                const double conv1 = solver.norm();
                if (conv1 < tolerance) {
                    std::printf("%3d %5d %15.6e\n", step, iter, conv1);
                    break;
                }
                solver.cgSolve(1e-10, fp32 ? INS_SOLVE_MIXED : INS_SOLVE_FP64);
                base::dof::addToDoFsFromSolver(solver, field);
                const double conv2 = solver.norm();
                std::printf("%3d %5d %15.6e %15.6e\n", step, iter, conv1, conv2);
                iter++;
                if (conv2 < tolerance) break;
            }
            if (iter == maxIter)
                std::printf("# (WW) Step %d has not converged within %d iterations\n", step, maxIter);
        }
        double umax = 0.;
        for (std::size_t n = 0; n < mesh.numNodes(); n++) umax = std::max(umax, std::abs(field.value[3 * n + 1]));
        std::printf("# max |u_y| %.9e\n", umax);
    } catch (const Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
        return 1;
    }
    return 0;
}
