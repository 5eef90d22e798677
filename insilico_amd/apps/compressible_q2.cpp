// reference/06-elastic/compressible.cpp on the drop-in layer: a block of
// compressible neo-Hookean material with a degree-2 field (geomDeg 1,
// fieldDeg 2, kernelDegEstimate 3) on an SMF mesh of the unit square/cube,
// x = 0 clamped (reference/06-elastic/PulledSheet.hpp) and on x = 1 either
// u_x prescribed (dispControlled 1) or the traction t_x = traction (y - 1/2)
// (dispControlled 0); load steps with the (step+1)/step ramp of the
// constraints and the (step+1)/loadSteps ramp of the traction, full Newton:
// a fresh solver per iteration, residual forces and the tangent with
// incremental lifting, |F| = solver.norm(), stop when |F| < tol * E,
// cgSolve, addToDoFsFromSolver, |x| = solver.norm(), stop when |x| < tol.
// After the set-up and after every load step it writes base.NNNN.vtk next
// to the mesh (ref06::writeVTKFile, compressible.cpp:60-89): the grid, the
// displacement at the nodes and the Cauchy stress at the element centroids,
// evaluated on the device.
// Usage: compressible_q2 mesh.smf input.dat
// input.dat: E nu maxIter loadSteps pull traction dispControlled tolerance
// Prints the reference's table: step iter |F| |x|
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

#include "insilico/insilico.hpp"
#include "insilico/post.hpp"

namespace {
template <typename MATERIAL>
void writeVTKFile(const std::string& baseName, unsigned step, const insilico::Mesh& mesh,
                  const insilico::Field& disp, const MATERIAL& material)
{
    const std::string vtkFile = baseName + "." + base::io::leadingZeros(step) + ".vtk";
    std::ofstream vtk(vtkFile.c_str());
    if (!vtk) throw insilico::Error(INS_ERR_ARG, "cannot write " + vtkFile);
    base::io::vtk::LegacyWriter vtkWriter(vtk);
    vtkWriter.writeUnstructuredGrid(mesh);
    base::io::vtk::writePointData(vtkWriter, mesh, disp, "disp");
    const std::array<double, 3> xi = base::shapeCentroid();
    base::io::vtk::writeCellData(vtkWriter, solid::cauchy(mesh, disp, material, xi), 9, "sigma");
}
}  // namespace

int main(int argc, char* argv[])
{
    using namespace insilico;
    if (argc != 3) {
        std::printf("Usage:  %s  mesh.smf input.dat \n", argv[0]);
        return 0;
    }
    const std::string meshFile = argv[1], inputFile = argv[2];
    try {
        const auto prop = readProperties(inputFile, {"E", "nu", "pull", "maxIter", "loadSteps",
                                                     "traction", "dispControlled", "tolerance"});
        const double E = std::atof(prop.at("E").c_str()), nu = std::atof(prop.at("nu").c_str());
        const double pull = std::atof(prop.at("pull").c_str());
        const double traction = std::atof(prop.at("traction").c_str());
        const double tolerance = std::atof(prop.at("tolerance").c_str());
        const unsigned maxIter = (unsigned)std::atoi(prop.at("maxIter").c_str());
        const unsigned loadSteps = (unsigned)std::atoi(prop.at("loadSteps").c_str());
        const bool dispControlled = std::atoi(prop.at("dispControlled").c_str()) != 0;
        const std::string baseName = base::io::baseName(meshFile, ".smf");

        const int dim = smfDimension(meshFile);
        Mesh mesh = readSMF(meshFile, dim);
        Field field = generateField(mesh, 2, dim);   // base::dof::generate<FEBasis>
        const double firstPull = pull / static_cast<double>(loadSteps);
        const double coordTol = 1.e-5;
        constrainBoundary(mesh, field, [&](const std::array<double, 3>& x, std::size_t node, Field& f) {
            if (std::abs(x[0] - 0.) < coordTol)
                for (int d = 0; d < dim; d++)
                    if (f.isActive(node, d)) f.constrainValue(node, d, 0.0);
            if (std::abs(x[0] - 1.) < coordTol && dispControlled && f.isActive(node, 0))
                f.constrainValue(node, 0, firstPull);
        });
        // the Neumann function is zero away from x = 1: the faces there
        std::vector<std::vector<int32_t>> rightFaces;
        for (auto& face : boundaryFieldFaces(mesh, field)) {
            bool right = true;
            for (int32_t n : face) right = right && std::abs(field.support[(std::size_t)n * dim] - 1.) < coordTol;
            if (right) rightFaces.push_back(face);
        }
        const std::size_t numDofs = numberDoFsConsecutively(field);
        std::printf("# Number of dofs %zu\n", numDofs);
        std::printf("#  Step  Iter  |F|          |x|  \n");

        typedef mat::hypel::NeoHookeanCompressible Material;
        typedef base::FieldBinder<Mesh, Field>::TupleBinder<1, 1>::Type FTB;
        base::FieldBinder<Mesh, Field> fieldBinder(mesh, field);
        const Material material(mat::Lame::lambda(E, nu), mat::Lame::mu(E, nu));
        solid::HyperElastic<Material, FTB> hyperElastic(material);
        writeVTKFile(baseName, 0, mesh, field, material);

        base::solver::Hip solver(numDofs);
        for (unsigned step = 0; step < loadSteps; step++) {
            const double pullFactor = (step == 0 ? static_cast<double>(step + 1)
                                                 : static_cast<double>(step + 1) / static_cast<double>(step));
            scaleConstraints(field, pullFactor);
            unsigned iter = 0;
            while (iter < maxIter) {
                solver.reset();
                if (!dispControlled) {
                    const double tractionFactor =
                        traction * static_cast<double>(step + 1) / static_cast<double>(loadSteps);
                    base::asmb::neumannForceComputation(
                        solver, fieldBinder, rightFaces,
                        base::asmb::LinearTraction{{tractionFactor, 0., 0.}, 1, 0.5});
                }
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
                    std::printf("%3u %5u %15.6g\n", step, iter, conv1);
                    break;
                }
                // the true residual of these systems stalls near 1e-12 in
                // FP64 (1.3e-12 under traction: small loads against E =
                // 1000), so cgSolve asks for 1e-11
                solver.cgSolve(1e-11);
                base::dof::addToDoFsFromSolver(solver, field);
                const double conv2 = solver.norm();
                std::printf("%3u %5u %15.6g %15.6g\n", step, iter, conv1, conv2);
                iter++;
                if (conv2 < tolerance) break;
            }
            if (iter == maxIter)
                std::printf("# (WW) Step %u has not converged within %u iterations \n", step, maxIter);
            writeVTKFile(baseName, step + 1, mesh, field, material);
        }
    } catch (const Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
        return 1;
    }
    return 0;
}
