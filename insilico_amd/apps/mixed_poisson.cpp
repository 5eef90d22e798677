// reference/05-mixedPoisson/mixedPoisson.cpp on the drop-in layer: Poisson's
// equation with a degree-2 field (geomDeg 1, fieldDeg 2, kernelDegEstimate
// 3) on an SMF mesh, the constructed solution u = cos(3x) cos(5y)
// (ReferenceSolution.hpp with alpha = 3, 5, 4): u prescribed at the
// boundary DoFs with some x_d = 0, the co-normal derivative grad u . n as
// Neumann data on the boundary, f = -Lap u (GivenData.hpp); direct solve
// (choleskySolve: a PCG to 1e-13 here).  Writes base.vtk next to the mesh
// (temperature at the nodes, flux = grad u_h at the element centroids) and
// prints the number of DoFs, the L2 and H1 errors and the mesh volume, all
// evaluated on the device.
// Usage: mixed_poisson file.smf
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include "insilico/insilico.hpp"
#include "insilico/post.hpp"

int main(int argc, char* argv[])
{
    using namespace insilico;
    if (argc != 2) {
        std::printf("Usage:  %s file.smf \n\n", argv[0]);
        return 0;
    }
    const std::string smfFile = argv[1];
    const std::string baseName = base::io::baseName(smfFile, ".smf");
    try {
        const int dim = smfDimension(smfFile);
        Mesh mesh = readSMF(smfFile, dim);
        const std::array<double, 3> alpha{3., 5., 4.};
        std::array<double, 3> k{0., 0., 0.};
        double amp = 0.;
        for (int d = 0; d < dim; d++) {
            k[d] = alpha[d];
            amp += alpha[d] * alpha[d];
        }
        Field field = generateField(mesh, 2, 1);
        const double coordTol = 1.e-5;
        constrainBoundary(mesh, field, [&](const std::array<double, 3>& x, std::size_t node, Field& f) {
            double value = 1.;
            for (int d = 0; d < dim; d++) value *= std::cos(alpha[d] * x[d]);
            bool onDirichletBdr = false;
            for (int d = 0; d < dim; d++)
                if (std::abs(x[d] - 0.) < coordTol) onDirichletBdr = true;
            if (onDirichletBdr && f.isActive(node, 0)) f.constrainValue(node, 0, value);
        });
        const std::size_t numDofs = numberDoFsConsecutively(field);
        std::printf("Number of dofs %zu\n", numDofs);

        typedef base::FieldBinder<Mesh, Field>::TupleBinder<1, 1>::Type FTB;
        base::FieldBinder<Mesh, Field> fieldBinder(mesh, field);
        base::solver::Hip solver(numDofs);
        if (dim == 2) {
            base::Quadrature<3, base::QUAD> quadrature;
            base::asmb::bodyForceComputation<FTB>(quadrature, solver, fieldBinder,
                                                  base::asmb::CosineProductForce{amp, k});
        } else {
            base::Quadrature<3, base::HEX> quadrature;
            base::asmb::bodyForceComputation<FTB>(quadrature, solver, fieldBinder,
                                                  base::asmb::CosineProductForce{amp, k});
        }
        // the whole boundary: contributions to constrained rows drop out
        base::asmb::neumannForceComputation(solver, fieldBinder, boundaryFieldFaces(mesh, field),
                                            base::asmb::CosineProductFlux{1.0, k});
        heat::Laplace<FTB> laplace(1.);
        if (dim == 2) {
            base::Quadrature<3, base::QUAD> quadrature;
            base::asmb::stiffnessMatrixComputation<FTB>(quadrature, solver, fieldBinder, laplace);
        } else {
            base::Quadrature<3, base::HEX> quadrature;
            base::asmb::stiffnessMatrixComputation<FTB>(quadrature, solver, fieldBinder, laplace);
        }
        solver.finishAssembly();
        solver.choleskySolve();
        base::dof::setDoFsFromSolver(solver, field);

        {
            const std::string vtkFile = baseName + ".vtk";
            std::ofstream vtk(vtkFile.c_str());
            if (!vtk) throw Error(INS_ERR_ARG, "cannot write " + vtkFile);
            base::io::vtk::LegacyWriter vtkWriter(vtk);
            vtkWriter.writeUnstructuredGrid(mesh);
            base::io::vtk::writePointData(vtkWriter, mesh, field, "temperature");
            base::io::vtk::writeCellData(
                vtkWriter, base::post::evaluateFieldGradientAtCells(mesh, field, base::shapeCentroid()),
                dim, "flux");
        }
        const CosineProductSolution refSol{1.0, k};
        double volume = 0.;
        if (dim == 2) {
            base::Quadrature<3, base::QUAD> quadrature;
            std::printf("L2-error = %g\n", base::post::errorComputation<0>(quadrature, mesh, field, refSol));
            std::printf("H1-error = %g\n", base::post::errorComputation<1>(quadrature, mesh, field, refSol));
            base::asmb::simplyIntegrate<FTB>(quadrature, volume, fieldBinder, base::kernel::Measure{});
        } else {
            base::Quadrature<3, base::HEX> quadrature;
            std::printf("L2-error = %g\n", base::post::errorComputation<0>(quadrature, mesh, field, refSol));
            std::printf("H1-error = %g\n", base::post::errorComputation<1>(quadrature, mesh, field, refSol));
            base::asmb::simplyIntegrate<FTB>(quadrature, volume, fieldBinder, base::kernel::Measure{});
        }
        std::printf("Volume of mesh: %g\n", volume);
    } catch (const Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
        return 1;
    }
    return 0;
}
