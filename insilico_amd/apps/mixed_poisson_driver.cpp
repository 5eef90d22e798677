// reference/05-mixedPoisson/mixedPoissonWithDriver.cpp on the drop-in layer:
// the problem of mixed_poisson (u = cos(3x) cos(5y) [cos(4z)] prescribed
// where some x_d = 0, the co-normal derivative grad u . n as Neumann data,
// f = -Lap u), set up through base::BoundaryValueProblem<Mesh, 1, 2> and
// solved by heat::PoissonDriver with the caller's functions -- a Dirichlet
// function on the DoF objects, a body force f(x) and a Neumann function
// ff(x, n) -- evaluated on the host at the library's quadrature points.
// Writes test.vtk in the working directory (u at the nodes, grad u at the
// element centroids) and prints the number of DoFs and |u|_2 over the DoF
// objects.  Usage: mixed_poisson_driver file.smf [dim]
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>

#include "insilico/driver.hpp"

namespace {
const double coordTol = 1.e-5;

// the constructed solution prod_d cos(alpha_d x_d) with its gradient and
// Laplacian (reference/05-mixedPoisson/ReferenceSolution.hpp)
template <unsigned DIM>
struct CosineSolution {
    typedef base::VecN<DIM> VecDim;
    double alpha[3];
    double value(const VecDim& x) const
    {
        double v = 1.;
        for (unsigned d = 0; d < DIM; d++) v *= std::cos(alpha[d] * x[d]);
        return v;
    }
    VecDim gradient(const VecDim& x) const
    {
        VecDim g;
        for (unsigned d = 0; d < DIM; d++) {
            double v = -alpha[d] * std::sin(alpha[d] * x[d]);
            for (unsigned e = 0; e < DIM; e++)
                if (e != d) v *= std::cos(alpha[e] * x[e]);
            g[d] = v;
        }
        return g;
    }
    double laplacian(const VecDim& x) const
    {
        double a2 = 0.;
        for (unsigned d = 0; d < DIM; d++) a2 += alpha[d] * alpha[d];
        return -a2 * value(x);
    }
};

// the problem data (reference/05-mixedPoisson/GivenData.hpp)
template <typename BVP>
struct GivenData {
    typedef typename BVP::VecDim VecDim;
    typedef typename BVP::VecDoF VecDoF;
    CosineSolution<BVP::dim> sol;
    void dirichletBC(const VecDim& x, typename BVP::DoF* doFPtr) const
    {
        bool onDirichletBdr = false;
        for (unsigned d = 0; d < BVP::dim; d++)
            if (std::abs(x[d] - 0.) < coordTol) onDirichletBdr = true;
        if (onDirichletBdr && doFPtr->isActive(0)) doFPtr->constrainValue(0, sol.value(x));
    }
    VecDoF forceFun(const VecDim& x) const
    {
        VecDoF f;
        f[0] = -sol.laplacian(x);
        return f;
    }
    VecDoF neumannBC(const VecDim& x, const VecDim& normal) const
    {
        VecDoF t;
        t[0] = sol.gradient(x).dot(normal);
        return t;
    }
};

template <base::Shape SHAPE>
int run(const std::string& smfFile)
{
    typedef base::Unstructured<SHAPE, 1> Mesh;
    Mesh mesh;
    {
        std::ifstream smf(smfFile.c_str());
        if (!smf) throw insilico::Error(INS_ERR_ARG, "cannot open " + smfFile);
        base::io::smf::readMesh(smf, mesh);
    }
    typedef base::BoundaryValueProblem<Mesh, 1, 2> BoundaryValueProblem;
    BoundaryValueProblem bvp(mesh);
    GivenData<BoundaryValueProblem> givenData{{{3., 5., 4.}}};

    heat::PoissonDriver<BoundaryValueProblem> ppd(bvp, 1.);
    using std::placeholders::_1;
    using std::placeholders::_2;
    ppd.dirichlet(std::bind(&GivenData<BoundaryValueProblem>::dirichletBC, &givenData, _1, _2));
    ppd.neumann(std::bind(&GivenData<BoundaryValueProblem>::neumannBC, &givenData, _1, _2));
    ppd.force(std::bind(&GivenData<BoundaryValueProblem>::forceFun, &givenData, _1));
    ppd.solve();

    bvp.writeVTKFile("test");
    const insilico::Field& f = bvp.getField();
    std::size_t active = 0;
    double s = 0.;
    for (std::size_t k = 0; k < f.value.size(); k++) {
        s += f.value[k] * f.value[k];
        active += f.status[k] == insilico::ACTIVE;
    }
    std::printf("Number of dofs %zu\n", active);
    std::printf("|u| = %.15e\n", std::sqrt(s));
    return 0;
}
}  // namespace

int main(int argc, char* argv[])
{
    if (argc < 2) {
        std::printf("Usage:  %s file.smf [dim]\n\n", argv[0]);
        return -1;
    }
    try {
        const int dim = argc > 2 ? std::atoi(argv[2]) : 2;
        return dim == 3 ? run<base::HEX>(argv[1]) : run<base::QUAD>(argv[1]);
    } catch (const insilico::Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
        return 1;
    }
}
