// Boundary value problems and drivers in the reference's call shape, over
// the device solver base::solver::Hip:
//   base::BoundaryValueProblem<MESH, DOFSIZE, FDEG>  (base/BoundaryValueProblem.hpp:149-371)
//   heat::PoissonDriver<BVP, MATERIAL>               (heat/PoissonDriver.hpp:53-210)
//   solid::CompressibleDriver<BVP, MATERIAL>         (solid/CompressibleDriver.hpp:40-214)
// The caller's Dirichlet, body force and Neumann functions are host
// functions as in the reference: constraints are set on the host field, the
// load functions are evaluated at the library's quadrature points and
// uploaded as per-point values, assembly and solves run on the device.
// This header, with insilico.hpp and post.hpp, takes the place of the
// reference's base/, heat/ and solid/ headers for a driver (the same
// namespaces and names; it is not meant to be included next to them).
#pragma once

#include <fstream>
#include <functional>
#include <istream>
#include <memory>
#include <string>

#include "insilico.hpp"
#include "post.hpp"

namespace base {

// base::HyperCubeShape<DIM> (base/shape.hpp): QUAD in 2-D, HEX in 3-D
template <unsigned DIM>
struct HyperCubeShape;
template <>
struct HyperCubeShape<2> {
    static constexpr Shape value = QUAD;
};
template <>
struct HyperCubeShape<3> {
    static constexpr Shape value = HEX;
};

namespace io {
namespace smf {
// base::io::smf::readMesh (base/io/smf/Reader.hpp:97-312) from a stream
template <Shape S, unsigned G>
void readMesh(std::istream& smf, Unstructured<S, G>& mesh)
{
    static_cast<insilico::Mesh&>(mesh) = insilico::readSMF(smf, (int)ShapeDim<S>::value);
}
}  // namespace smf
}  // namespace io

namespace dof {
// base::dof::scaleConstraints (base/dof/Distribute.hpp:76-80)
inline void scaleConstraints(insilico::Field& f, double factor) { insilico::scaleConstraints(f, factor); }
}  // namespace dof

namespace time {
// no time integration: the field keeps one value per DoF
struct NoTimeIntegration {
    static constexpr unsigned numSteps = 1;
};
}  // namespace time

// base::BoundaryValueProblem (base/BoundaryValueProblem.hpp:149-371): the
// mesh, the field of DOFSIZE components and degree FDEG on it, the boundary
// faces, and the calls that constrain, number, assemble into a solver and
// take the solution back.
template <typename MESH, unsigned DOFSIZE, unsigned FDEG = MESH::Element::GeomFun::degree,
          typename TIMEINT = time::NoTimeIntegration>
class BoundaryValueProblem {
public:
    typedef MESH Mesh;
    typedef TIMEINT TimeIntegration;
    static const unsigned doFSize = DOFSIZE;
    static const unsigned fieldDeg = FDEG;
    static const unsigned dim = Mesh::Node::dim;
    static const Shape shape = Mesh::Element::shape;

    typedef insilico::Field Field;
    // the DegreeOfFreedom a Dirichlet function receives
    // (base/dof/DegreeOfFreedom.hpp:201-260)
    struct DoF {
        static constexpr unsigned size = DOFSIZE;
        Field* field;
        std::size_t node;
        bool isActive(unsigned d) const { return field->isActive(node, (int)d); }
        bool isConstrained(unsigned d) const { return !isActive(d); }
        void constrainValue(unsigned d, double v) { field->constrainValue(node, (int)d, v); }
        double getValue(unsigned d) const { return field->value[node * DOFSIZE + d]; }
        void setValue(unsigned d, double v) { field->value[node * DOFSIZE + d] = v; }
    };
    typedef base::FieldBinder<Mesh, Field> FieldBinder;
    typedef typename FieldBinder::template TupleBinder<1, 1>::Type UU;
    typedef VecN<dim> VecDim;
    typedef VecN<DOFSIZE> VecDoF;
    typedef base::mesh::BoundaryMesh BoundaryMesh;
    typedef base::asmb::SurfaceFieldBinder<BoundaryMesh, Field> SurfaceFieldBinder;
    typedef typename SurfaceFieldBinder::template TupleBinder<1, 1>::Type SUU;
    typedef std::function<void(const VecDim&, DoF*)> DirichletFun;
    typedef std::function<VecDoF(const VecDim&)> BodyForceFun;
    typedef std::function<VecDoF(const VecDim&, const VecDim&)> SurfaceForceFun;

    explicit BoundaryValueProblem(Mesh& mesh, double density = 1.)
        : mesh_(mesh), density_(density)
    {
        if (mesh_.numElements() == 0)
            throw insilico::Error(INS_ERR_ARG, "Mesh is empty - a proper mesh is needed here ");
        field_ = insilico::generateField(mesh_, (int)FDEG, (int)DOFSIZE);
    }

    void generateBoundaryMesh()
    {
        if (haveBoundaryMesh_) return;
        boundaryMesh_.volume = &mesh_;
        boundaryMesh_.faces = insilico::boundaryFieldFaces(mesh_, field_);
        haveBoundaryMesh_ = true;
    }

    // base::dof::constrainBoundary: the function at every DoF object of the
    // boundary faces
    void applyDirichletConstraints(DirichletFun diriFun)
    {
        if (doFsAreNumbered_)
            throw insilico::Error(INS_ERR_STATE,
                                  "Constraints have to be applied before numbering the DoFs");
        insilico::constrainBoundary(mesh_, field_,
                                    [&](const std::array<double, 3>& x, std::size_t node, Field& f) {
                                        VecDim xv;
                                        for (unsigned d = 0; d < dim; d++) xv[d] = x[d];
                                        DoF dof{&f, node};
                                        diriFun(xv, &dof);
                                    });
    }

    std::size_t numberDoFs(const std::size_t init = 0)
    {
        if (init != 0)
            throw insilico::Error(INS_ERR_ARG, "numberDoFs: the solver rows start at 0");
        const std::size_t n = insilico::numberDoFsConsecutively(field_);
        doFsAreNumbered_ = true;
        return n;
    }

    template <typename SOLVER>
    void registerInSolver(SOLVER& solver)
    {
        requireNumbered_();
        solver.bind(mesh_, field_);
    }

    template <typename KERNEL, typename QUAD, typename SOLVER>
    void assembleBilinearForm(const KERNEL& kernel, const QUAD& quadrature, SOLVER& solver,
                              const bool incremental = false)
    {
        requireNumbered_();
        FieldBinder fb(mesh_, field_);
        base::asmb::stiffnessMatrixComputation<UU>(quadrature, solver, fb, kernel, incremental);
        base::asmb::computeResidualForces<UU>(quadrature, solver, fb, kernel);
    }

    template <typename QUAD, typename SOLVER>
    void applyBodyForce(const QUAD& quadrature, SOLVER& solver, BodyForceFun forceFun)
    {
        requireNumbered_();
        FieldBinder fb(mesh_, field_);
        base::asmb::bodyForceComputation<UU>(quadrature, solver, fb, forceFun);
    }

    template <typename SQUAD, typename SOLVER>
    void applyBoundaryForces(const SQUAD& surfaceQuadrature, SOLVER& solver,
                             SurfaceForceFun neumannForceFun)
    {
        requireNumbered_();
        if (!haveBoundaryMesh_)
            throw insilico::Error(INS_ERR_STATE, "Cannot apply Neumann BC without boundary mesh");
        SurfaceFieldBinder sfb(boundaryMesh_, field_);
        base::asmb::neumannForceComputation<SUU>(surfaceQuadrature, solver, sfb, neumannForceFun);
    }

    template <typename SOLVER>
    void setDoFsFromSolver(const SOLVER& solver)
    {
        base::dof::setDoFsFromSolver(solver, field_);
    }
    template <typename SOLVER>
    void addToDoFsFromSolver(const SOLVER& solver)
    {
        base::dof::addToDoFsFromSolver(solver, field_);
    }

    Mesh& getMesh() { return mesh_; }
    Field& getField() { return field_; }
    BoundaryMesh& getBoundaryMesh() { return boundaryMesh_; }
    double density() const { return density_; }

    // the grid, the field at the nodes and its gradient at the element
    // centroids (base/BoundaryValueProblem.hpp:378-398)
    void writeVTKFile(const std::string& baseName, const std::string& fieldName = "u",
                      const std::string& gradName = "gradU") const
    {
        const std::string vtkFile = baseName + ".vtk";
        std::ofstream vtk(vtkFile.c_str());
        if (!vtk) throw insilico::Error(INS_ERR_ARG, "cannot write " + vtkFile);
        base::io::vtk::LegacyWriter vtkWriter(vtk);
        vtkWriter.writeUnstructuredGrid(mesh_);
        base::io::vtk::writePointData(vtkWriter, mesh_, field_, fieldName);
        base::io::vtk::writeCellData(
            vtkWriter,
            base::post::evaluateFieldGradientAtCells(mesh_, field_, base::shapeCentroid()),
            (int)(dim * DOFSIZE), gradName);
    }

private:
    void requireNumbered_() const
    {
        if (!doFsAreNumbered_) throw insilico::Error(INS_ERR_STATE, "DoFs need to be numbered first");
    }

    Mesh& mesh_;
    double density_;
    Field field_;
    BoundaryMesh boundaryMesh_;
    bool haveBoundaryMesh_ = false;
    bool doFsAreNumbered_ = false;
};
}  // namespace base

namespace heat {
// heat::PoissonDriver (heat/PoissonDriver.hpp:53-210): Dirichlet, Neumann and
// body force functions, then one black-box solve -- numbering, the
// quadratures of kernel degree 2 * fieldDeg, device assembly of the Static
// kernel and the loads, the direct solve (choleskySolve / luSolve), the
// field from the solution.
template <typename BVP, typename MATERIAL = mat::thermal::IsotropicConstant>
class PoissonDriver {
public:
    typedef BVP BoundaryValueProblem;
    typedef MATERIAL Material;
    static_assert(BVP::doFSize == 1, "Only scalar DoFs are supported");
    typedef typename BVP::DirichletFun DirichletFun;
    typedef typename BVP::BodyForceFun BodyForceFun;
    typedef typename BVP::SurfaceForceFun SurfaceForceFun;
    typedef heat::Static<Material, typename BVP::UU> Kernel;

    PoissonDriver(BVP& bvp, const Material& material)
        : bvp_(bvp), material_(material), kernel_(material_) {}
    PoissonDriver(BVP& bvp, const double materialParam = 1)
        : bvp_(bvp), material_(materialParam), kernel_(material_) {}

    void dirichlet(DirichletFun dirichletFun) { bvp_.applyDirichletConstraints(dirichletFun); }
    void neumann(SurfaceForceFun surfaceForceFun)
    {
        surfaceForceFun_ = surfaceForceFun;
        bvp_.generateBoundaryMesh();
        withNeumann_ = true;
    }
    void force(BodyForceFun bodyForceFun)
    {
        bodyForceFun_ = bodyForceFun;
        withBodyForce_ = true;
    }

    template <typename SOLVER, typename QUAD, typename SQUAD>
    void assemble(SOLVER& solver, const QUAD& quadrature, const SQUAD& surfaceQuadrature)
    {
        bvp_.assembleBilinearForm(kernel_, quadrature, solver);
        if (withBodyForce_) bvp_.applyBodyForce(quadrature, solver, bodyForceFun_);
        if (withNeumann_) bvp_.applyBoundaryForces(surfaceQuadrature, solver, surfaceForceFun_);
    }

    void solve(const bool symmetric = true)
    {
        const std::size_t numDoFs = bvp_.numberDoFs();
        static constexpr unsigned kernelDegree = 2 * BVP::fieldDeg;
        base::Quadrature<kernelDegree, BVP::shape> quadrature;
        base::SurfaceQuadrature<kernelDegree, BVP::shape> surfaceQuadrature;
        base::solver::Hip solver(numDoFs);
        bvp_.registerInSolver(solver);
        this->assemble(solver, quadrature, surfaceQuadrature);
        solver.finishAssembly();
        if (symmetric) solver.choleskySolve();
        else solver.luSolve();
        bvp_.setDoFsFromSolver(solver);
    }

    Kernel& getKernel() { return kernel_; }

protected:
    BVP& bvp_;

private:
    BodyForceFun bodyForceFun_;
    SurfaceForceFun surfaceForceFun_;
    Material material_;
    Kernel kernel_;
    bool withNeumann_ = false;
    bool withBodyForce_ = false;
};
}  // namespace heat

namespace solid {
// solid::CompressibleDriver (solid/CompressibleDriver.hpp:40-214): the load
// step's Newton iteration -- constraints scaled to the step, assembly of
// tangent and residual (incremental lifting) with the loads, |F| = norm()
// test, direct solve, DoF update, |dx| = norm() test.  One device context
// serves all iterations (the reference builds a new solver each time; its
// pattern does not change).
template <typename BVP, typename MATERIAL>
class CompressibleDriver {
public:
    typedef BVP BoundaryValueProblem;
    typedef MATERIAL Material;
    typedef typename BVP::DirichletFun DirichletFun;
    typedef typename BVP::BodyForceFun BodyForceFun;
    typedef typename BVP::SurfaceForceFun SurfaceForceFun;
    typedef solid::HyperElastic<Material, typename BVP::UU> Kernel;

    CompressibleDriver(BVP& bvp, const Material& material)
        : bvp_(bvp), material_(material), kernel_(material_) {}

    void dirichlet(DirichletFun dirichletFun) { bvp_.applyDirichletConstraints(dirichletFun); }
    void neumann(SurfaceForceFun surfaceForceFun)
    {
        surfaceForceFun_ = surfaceForceFun;
        bvp_.generateBoundaryMesh();
        withNeumann_ = true;
    }
    void force(BodyForceFun bodyForceFun)
    {
        bodyForceFun_ = bodyForceFun;
        withBodyForce_ = true;
    }

    template <typename SOLVER, typename QUAD, typename SQUAD>
    void assemble(SOLVER& solver, const QUAD& quadrature, const SQUAD& surfaceQuadrature)
    {
        bvp_.assembleBilinearForm(kernel_, quadrature, solver, true);
        if (withBodyForce_) bvp_.applyBodyForce(quadrature, solver, bodyForceFun_);
        if (withNeumann_) bvp_.applyBoundaryForces(surfaceQuadrature, solver, surfaceForceFun_);
    }

    unsigned solve(const unsigned stepNum, const unsigned maxIter, const double tolerance)
    {
        if (numDoFs_ == 0) numDoFs_ = bvp_.numberDoFs();
        // rescale the constraints in every load step: new value / old value
        const double constraintFactor =
            stepNum == 0 ? (double)(stepNum + 1) : (double)(stepNum + 1) / (double)stepNum;
        base::dof::scaleConstraints(bvp_.getField(), constraintFactor);
        static constexpr unsigned kernelDegree = 2 * BVP::fieldDeg;
        base::Quadrature<kernelDegree, BVP::shape> quadrature;
        base::SurfaceQuadrature<kernelDegree, BVP::shape> surfaceQuadrature;
        if (!solver_) solver_.reset(new base::solver::Hip(numDoFs_));
        unsigned iter = 0;
        while (iter < maxIter) {
            solver_->reset();
            bvp_.registerInSolver(*solver_);
            this->assemble(*solver_, quadrature, surfaceQuadrature);
            solver_->finishAssembly();
            const double normR = solver_->norm();
            if (normR < tolerance) break;
            solver_->choleskySolve();
            bvp_.addToDoFsFromSolver(*solver_);
            const double normDU = solver_->norm();
            iter++;
            if (normDU < tolerance) break;
        }
        return iter;
    }

    Kernel& getKernel() { return kernel_; }

protected:
    BVP& bvp_;
    std::size_t numDoFs_ = 0;

private:
    BodyForceFun bodyForceFun_;
    SurfaceForceFun surfaceForceFun_;
    Material material_;
    Kernel kernel_;
    std::unique_ptr<base::solver::Hip> solver_;
    bool withNeumann_ = false;
    bool withBodyForce_ = false;
};
}  // namespace solid
