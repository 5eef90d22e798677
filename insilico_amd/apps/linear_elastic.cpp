// Linear elasticity with the elastostatic fundamental solution as boundary
// data, as reference/06-elastic/linearElastic.cpp:80-216: Q1 StVenant,
// E = 1000, nu = 0.25, all boundary DoFs constrained to U(x; y, e_1) with the
// source y = (-0.1, ...), quadrature degree 3, CG solve; writes base.vtk next
// to the mesh (displacement at the nodes, Cauchy stress at the element
// centroids, linearElastic.cpp:48-77) and prints the L2 error of the
// displacement against the fundamental solution
// (base::post::errorComputation<0>), both evaluated on the device.
// Usage: linear_elastic mesh.smf dim
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

#include "insilico/insilico.hpp"
#include "insilico/post.hpp"

int main(int argc, char* argv[])
{
    using namespace insilico;
    if (argc != 3) {
        std::printf("Usage:  %s  mesh.smf dim\n", argv[0]);
        return 0;
    }
    const int dim = std::atoi(argv[2]);
    const double E = 1000.0, nu = 0.25;
    try {
        Mesh mesh = readSMF(argv[1], dim);
        const double lambda = mat::Lame::lambda(E, nu), mu = mat::Lame::mu(E, nu);
        const double src[3] = {-0.1, -0.1, -0.1};
        const double dir[3] = {0., 1., 0.};
        auto solFun = [&](const double* x) { return fundSolElastoStatic(dim, lambda, mu, x, src, dir); };

        Field field(mesh.numNodes(), dim);
        constrainBoundary(mesh, field, [&](const std::array<double, 3>& x, std::size_t node, Field& f) {
            const std::array<double, 3> U = solFun(x.data());
            for (int d = 0; d < dim; d++) f.constrainValue(node, d, U[d]);
        });
        const std::size_t numDofs = numberDoFsConsecutively(field);

        base::Quadrature<3, base::HEX> quadrature;  // kernelDegEstimate = 3
        base::FieldBinder<Mesh, Field> fieldBinder(mesh, field);
        typedef base::FieldBinder<Mesh, Field>::TupleBinder<1, 1>::Type FTB;
        typedef mat::hypel::StVenant Material;
        solid::HyperElastic<Material, FTB> hyperElastic(Material(lambda, mu));

        base::solver::Hip solver(numDofs);
        base::asmb::stiffnessMatrixComputation<FTB>(quadrature, solver, fieldBinder, hyperElastic);
        solver.finishAssembly();
        solver.cgSolve(1e-13);
        base::dof::setDoFsFromSolver(solver, field);

        const Material material(lambda, mu);
        {
            const std::string vtkFile = base::io::baseName(argv[1], ".smf") + ".vtk";
            std::ofstream vtk(vtkFile.c_str());
            if (!vtk) throw Error(INS_ERR_ARG, "cannot write " + vtkFile);
            base::io::vtk::LegacyWriter vtkWriter(vtk);
            vtkWriter.writeUnstructuredGrid(mesh);
            base::io::vtk::writePointData(vtkWriter, mesh, field, "disp");
            base::io::vtk::writeCellData(
                vtkWriter, solid::cauchy(mesh, field, material, base::shapeCentroid()), 9, "sigma");
        }
        FundamentalSolutionElastoStatic ref{lambda, mu, {src[0], src[1], src[2]}, {dir[0], dir[1], dir[2]}};
        std::printf("%g\n", base::post::errorComputation<0>(quadrature, mesh, field, ref));
    } catch (const Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
        return 1;
    }
    return 0;
}
