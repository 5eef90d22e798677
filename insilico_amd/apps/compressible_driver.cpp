// reference/06-elastic/compressibleWithDriver.cpp on the drop-in layer: a
// compressible neo-Hookean block (0,1)^2 with a degree-2 field, clamped at
// x = 0 and pulled to u_x = pull at x = 1 over loadSteps steps, through
// base::BoundaryValueProblem<Mesh, 2, 2> and solid::CompressibleDriver with
// the caller's Dirichlet function (reference/06-elastic/PulledSheet.hpp).
// Prints "step iterations" per load step as the reference does, writes
// <mesh>.vtk (displacement at the nodes, its gradient at the centroids)
// and |u|_2 over the DoF objects.
// Usage: compressible_driver mesh.smf input.dat
// (input.dat: E nu pull maxIter loadSteps tolerance)
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include "insilico/driver.hpp"

namespace {
const double coordTol = 1.e-6;

template <unsigned DIM>
struct PulledSheet {
    typedef base::VecN<DIM> VecDim;
    // fix x_0 = 0, and optionally pull x_0 = 1 to u_0 = value
    template <typename DOF>
    static void dirichletBC(const VecDim& x, DOF* doFPtr, const bool pullRightSide,
                            const double value)
    {
        const bool onLeftBdr = std::abs(x[0] - 0.) < coordTol;
        const bool onRightBdr = std::abs(x[0] - 1.) < coordTol;
        if (onLeftBdr)
            for (unsigned d = 0; d < DOF::size; d++)
                if (doFPtr->isActive(d)) doFPtr->constrainValue(d, 0.0);
        if (onRightBdr && pullRightSide && doFPtr->isActive(0)) doFPtr->constrainValue(0, value);
    }
};
}  // namespace

int main(int argc, char* argv[])
{
    if (argc != 3) {
        std::printf("Usage:  %s  mesh.smf input.dat \n", argv[0]);
        return 0;
    }
    const unsigned geomDeg = 1;
    const unsigned fieldDeg = 2;
    const base::Shape shape = base::HyperCubeShape<2>::value;
    const std::string meshFile = argv[1];
    const std::string inputFile = argv[2];
    try {
        const auto prop = insilico::readProperties(
            inputFile, {"E", "nu", "pull", "maxIter", "loadSteps", "tolerance"});
        const double E = std::atof(prop.at("E").c_str());
        const double nu = std::atof(prop.at("nu").c_str());
        const double pull = std::atof(prop.at("pull").c_str());
        const unsigned maxIter = (unsigned)std::atoi(prop.at("maxIter").c_str());
        const unsigned loadSteps = (unsigned)std::atoi(prop.at("loadSteps").c_str());
        const double tolerance = std::atof(prop.at("tolerance").c_str());
        const std::string baseName = base::io::baseName(meshFile, ".smf");

        typedef base::Unstructured<shape, geomDeg> Mesh;
        const unsigned dim = Mesh::Node::dim;
        Mesh mesh;
        {
            std::ifstream smf(meshFile.c_str());
            if (!smf) throw insilico::Error(INS_ERR_ARG, "cannot open " + meshFile);
            base::io::smf::readMesh(smf, mesh);
        }
        typedef base::BoundaryValueProblem<Mesh, dim, fieldDeg> BoundaryValueProblem;
        BoundaryValueProblem bvp(mesh);
        typedef mat::hypel::NeoHookeanCompressible Material;
        Material material(mat::Lame::lambda(E, nu), mat::Lame::mu(E, nu));
        solid::CompressibleDriver<BoundaryValueProblem, Material> driver(bvp, material);

        const double firstPull = pull / static_cast<double>(loadSteps);
        using std::placeholders::_1;
        using std::placeholders::_2;
        driver.dirichlet(std::bind(&PulledSheet<dim>::dirichletBC<BoundaryValueProblem::DoF>, _1,
                                   _2, true, firstPull));
        for (unsigned step = 0; step < loadSteps; step++)
            std::printf("%u  %u\n", step, driver.solve(step, maxIter, tolerance));

        bvp.writeVTKFile(baseName);
        const insilico::Field& f = bvp.getField();
        double s = 0.;
        for (double v : f.value) s += v * v;
        std::printf("|u| = %.15e\n", std::sqrt(s));
    } catch (const insilico::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
        return 1;
    }
    return 0;
}
