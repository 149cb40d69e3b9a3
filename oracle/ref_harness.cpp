// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference's own headers (elicase4/residuum-solver, compiled
// from /root/reference/include by oracle/Makefile into oracle/_ref/libref.so)
// to produce golden vectors for the oracle and the HIP path.  Nothing here
// ships: the product never links it.
//
// Only reference components that build without exprtk / yaml-cpp are used:
// BlockMesh2D/3D, TopologicalDOF, Assembler<CPU>, BoundaryApplicator<CPU>,
// the heat-equation forms, EvalElement / EvalQuadraturePointVolume /
// HeatStateVector, CSROperator, cg::Solver, gmres::Solver, Jacobi, Identity.
// The harness supplies what the reference's own fem-tier tests supply
// (tests/integration/fem/heateq/cpu/HeatEquationMinimal2D.cpp:83-87): plain
// callables for boundary values and sources, and a conductivity model with
// ConductivityModel's eval/evalGradient semantics (ConductivityModel.hpp:60-96)
// in place of the exprtk-backed one.  A no-op form hooks gatherElementData
// (fem/form/FormRegistry.hpp:32-40) to expose a per-element conductivity
// factor, which is this build's extension (no reference counterpart).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <type_traits>
#include <vector>

#include "core/Types.hpp"
#include "mesh/Mesh.hpp"
#include "mesh/generator/BlockMesh2D.hpp"
#include "mesh/generator/BlockMesh3D.hpp"
#include "topology/TopologicalDOF.hpp"
#include "fem/assembly/Assembler.hpp"
#include "fem/boundary/BoundaryApplicator.hpp"
#include "fem/boundary/EssentialBoundaryRegistry.hpp"
#include "fem/boundary/NaturalBoundaryRegistry.hpp"
#include "equation/heateq/boundary/BoundaryFluxFunction.hpp"
#include "equation/heateq/form/FluxBoundaryForm.hpp"
#include "equation/heateq/form/NodalFluxForm.hpp"
#include "fem/basis/LagrangeQuad.hpp"
#include "fem/basis/LagrangeHex.hpp"
#include "fem/quadrature/GaussQuadratureQuad.hpp"
#include "fem/quadrature/GaussQuadratureHex.hpp"
#include "fem/geometry/JacobianTransform.hpp"
#include "fem/form/FormRegistry.hpp"
#include "fem/form/ScaledForm.hpp"
#include "fem/evaluator/ModelRegistry.hpp"
#include "equation/heateq/evaluator/EvalElement.hpp"
#include "equation/heateq/evaluator/EvalQuadraturePointVolume.hpp"
#include "equation/heateq/evaluator/HeatStateVector.hpp"
#include "equation/heateq/evaluator/DensityModel.hpp"
#include "equation/heateq/evaluator/SourceFunction.hpp"
#include "equation/heateq/form/DiffusionForm.hpp"
#include "equation/heateq/form/MassForm.hpp"
#include "equation/heateq/form/TangentDiffusionForm.hpp"
#include "equation/heateq/form/SourceForm.hpp"
#include "equation/heateq/form/NodalSourceForm.hpp"
#include "equation/heateq/boundary/BoundaryValueFunction.hpp"
#include "linalg/types/Vector.hpp"
#include "linalg/types/CSRMatrix.hpp"
#include "linalg/operator/CSROperator.hpp"
#include "linalg/operator/FEMOperator.hpp"
#include "fem/quadrature/GaussQuadrature1D.hpp"
#include "fem/quantity/BoundaryQuantityRegistry.hpp"
#include "fem/quantity/QuantityEvaluator.hpp"
#include "equation/heateq/evaluator/EvalQuadraturePointBoundary.hpp"
#include "equation/heateq/evaluator/HeatBoundaryStateVector.hpp"
#include "equation/heateq/quantity/HeatFluxIntegrand.hpp"
#include "fem/quantity/QuantityForms.hpp"
#include "linalg/solver/iterative/cg/Solver.hpp"
#include "linalg/solver/iterative/gmres/Solver.hpp"
#include "linalg/solver/preconditioner/Jacobi.hpp"
#include "linalg/solver/preconditioner/Identity.hpp"
#include "utils/logging/core/NullLogger.hpp"

using namespace residuum;
using Backend = linalg::types::backend::CPU;
using VecT = linalg::types::Vector<Real, Backend>;
using MatT = linalg::types::CSRMatrix<Real, Backend>;

namespace {

// function kinds shared with the tests (tests/golden/make_golden.py):
// 0 "0", 1 "x", 2 "(1-x)*(1-y)", 3 "10*sin(pi*x)*sin(pi*y)", 4 "x*y",
// 5 "x*y*z", 6 "1", 7 "x+2*y+3*z"
Real eval_kind(int kind, int sd, const Real* x) {
    const Real X = x[0], Y = x[1], Z = sd == 3 ? x[2] : 0.0;
    switch (kind) {
    case 1: return X;
    case 2: return (1.0 - X) * (1.0 - Y);
    case 3: return 10.0 * std::sin(M_PI * X) * std::sin(M_PI * Y);
    case 4: return X * Y;
    case 5: return X * Y * Z;
    case 6: return 1.0;
    case 7: return X + 2.0 * Y + 3.0 * Z;
    default: return 0.0;
    }
}

template<int SD>
struct KindFn {
    int kind = 0;
    void operator()(Real, const Real* x, Real* out) const { out[0] = eval_kind(kind, SD, x); }
};

// per-element conductivity factor, set per element by KappaHook
thread_local const Real* g_kappa = nullptr;
thread_local const int64_t* g_node0_to_elem = nullptr;
thread_local Real g_cur_kappa = 1.0;

struct KappaHook {
    void gatherElementData(const Index* nodeIDs, Index) const {
        g_cur_kappa = g_kappa ? g_kappa[g_node0_to_elem[nodeIDs[0]]] : 1.0;
    }
    template<typename QP> void computeElementLevelMatrix(const QP&, Real*) const {}
    template<typename QP> void computeElementLevelVector(const QP&, const Real*, Real*) const {}
};

// ConductivityModel semantics: K = k * shape, dKdT = k' * shape
template<typename QP>
struct HarnessConductivity {
    static constexpr Index SD = QP::SpatialDim;
    bool tdep = false;
    Real ka = 1.0, kb = 0.0;
    bool aniso = false;
    Real tensor[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    void eval(QP& qp) const {
        Real k = tdep ? (ka + kb * qp.T.value()) : ka;
        if (g_kappa) k = g_cur_kappa * k;
        for (Index i = 0; i < SD; ++i)
            for (Index j = 0; j < SD; ++j) {
                const Real shape = aniso ? tensor[i * SD + j] : ((i == j) ? Real(1) : Real(0));
                qp.K[i * SD + j] = k * shape;
            }
    }
    void evalGradient(QP& qp) const {
        Real dk = tdep ? kb : Real(0);
        if (g_kappa) dk = g_cur_kappa * dk;
        for (Index i = 0; i < SD; ++i)
            for (Index j = 0; j < SD; ++j) {
                const Real shape = aniso ? tensor[i * SD + j] : ((i == j) ? Real(1) : Real(0));
                qp.dKdT[i * SD + j] = dk * shape;
            }
    }
};

template<typename QP>
struct HarnessCp {
    Real value = 1.0;
    void eval(QP& qp) const { qp.cp = value; }
    void evalGradient(QP& qp) const { qp.dcpdT = 0.0; }
};

// flux vector kinds shared with the tests (tests/oracle_py.py FLUX_KINDS):
// 0 (1.5, -0.5, 0.25), 1 (x, y*y, x*z), 2 (sin(pi*x) + t, cos(y), x*y)
template<int SD>
struct FluxKindFn {
    int kind = 0;
    void operator()(Real t, const Real* x, Real* out) const {
        const Real X = x[0], Y = x[1], Z = SD == 3 ? x[2] : 0.0;
        Real v[3];
        if (kind == 0) { v[0] = 1.5; v[1] = -0.5; v[2] = 0.25; }
        else if (kind == 1) { v[0] = X; v[1] = Y * Y; v[2] = X * Z; }
        else { v[0] = std::sin(M_PI * X) + t; v[1] = std::cos(Y); v[2] = X * Y; }
        for (int i = 0; i < SD; ++i) out[i] = v[i];
    }
};

// nodal flux vectors [n_nodes * SD]
template<int SD>
struct NodalVec {
    static constexpr Index NumComponents = SD;
    const Real* data = nullptr;
    void eval(Index node, Real* out) const {
        for (int i = 0; i < SD; ++i) out[i] = data[node * SD + i];
    }
};

struct NodalArray {
    static constexpr Index NumComponents = 1;
    const Real* data = nullptr;
    void eval(Index node, Real* out) const { out[0] = data[node]; }
};

struct Base {
    virtual ~Base() = default;
    int dim = 0;
    mesh::Mesh mesh;
    std::vector<int64_t> node0;
    std::unique_ptr<MatT> K;
    virtual void sizes(int64_t* out) = 0;
    virtual void dofs(int32_t* t2a, int32_t* ctag) = 0;
    virtual void map(int64_t* m) = 0;
    virtual void assemble_matrix(int set, Real coef, const Real* U, Real* val) = 0;
    virtual void assemble_vector(int set, Real coef, const Real* U, Real* F) = 0;
    virtual void load(int kind, const Real* nodal, Real* F) = 0;
    virtual void lift(int set, Real coef, Real* F) = 0;
    virtual void assemble_diagonal(int set, Real coef, const Real* U, Real* d) = 0;
    virtual void boundary_reaction(Real time, int src_kind, const Real* U, const Real* Udot, int n_tags,
                                   const int32_t* tags, Real* result, Real* measure) = 0;
    virtual void boundary_flux(Real time, const Real* U, int n_tags, const int32_t* tags, Real* result,
                               Real* measure) = 0;
    virtual void natural(Real time, int n_entries, const int32_t* tags, const int32_t* kinds,
                         const Real* const* nodal, Real* F) = 0;
    virtual int cg_fem(int set, Real coef, const Real* U, int pc, const Real* rhs, Real* x, Real tol, int64_t maxit,
                       int64_t* iters) = 0;
    virtual int be_run(Real t0, Real tf, Real dt, Real tol, int64_t maxit, Real* U, int64_t* iters) = 0;
    virtual int newton(Real abs_tol, Real rel_tol, int64_t max_newton, int src, Real lin_tol,
                       int64_t lin_max, int64_t krylov, Real* U, int64_t* lin_iters) = 0;
    // model
    bool tdep = false, aniso = false;
    Real ka = 1.0, kb = 0.0, tensor[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, rho = 1.0, cp = 1.0;
};

template<int D>
struct Ctx : Base {
    using Basis = std::conditional_t<D == 2, fem::basis::LagrangeQuad, fem::basis::LagrangeHex>;
    using Quad = std::conditional_t<D == 2, fem::quadrature::GaussQuadratureQuad, fem::quadrature::GaussQuadratureHex>;
    using EvalEle = equation::heateq::evaluator::EvalElement<Basis, D>;
    using QP = equation::heateq::evaluator::EvalQuadraturePointVolume<EvalEle, Basis, fem::geometry::JacobianTransform<D, D>>;
    using State = equation::heateq::evaluator::HeatStateVector<1, D, D, VecT>;
    using Cond = HarnessConductivity<QP>;
    using Density = equation::heateq::evaluator::DensityModel<QP>;
    using MassModel = fem::evaluator::ModelRegistry<Density, HarnessCp<QP>>;
    using OpModel = fem::evaluator::ModelRegistry<Cond, MassModel>;
    using Diff = equation::heateq::DiffusionForm<QP>;
    using Mass = equation::heateq::MassForm<QP>;
    using TDiff = equation::heateq::TangentDiffusionForm<QP>;
    using StiffForms = fem::form::FormRegistry<KappaHook, Diff>;
    using MassForms = fem::form::FormRegistry<Mass>;
    using BEForms = fem::form::FormRegistry<StiffForms, fem::form::ScaledForm<1, MassForms>>;
    using JacForms = fem::form::FormRegistry<StiffForms, fem::form::FormRegistry<TDiff>>;
    using DirBC = equation::heateq::BoundaryValueFunction<D, 1, KindFn<D>>;

    Basis basis;
    Quad quad;
    EvalEle ele;
    std::unique_ptr<topology::TopologicalDOF<1>> topo;
    fem::boundary::EssentialBoundaryRegistry bcs;
    std::vector<std::shared_ptr<fem::boundary::BoundaryCondition<DirBC>>> bc_keep;
    fem::assembly::Assembler<Backend> asmb;
    fem::boundary::BoundaryApplicator<Backend> app;

    Ctx(Basis b, Quad q) : basis(b), quad(q), ele(b) { dim = D; }

    Cond cond() const {
        Cond c; c.tdep = tdep; c.ka = ka; c.kb = kb; c.aniso = aniso;
        for (int i = 0; i < 9; ++i) c.tensor[i] = tensor[i];
        return c;
    }
    MassModel mass_model() const {
        Density d; d.value = rho;
        HarnessCp<QP> c; c.value = cp;
        return MassModel(d, c);
    }

    void sizes(int64_t* out) override {
        out[0] = mesh.data.numNodes; out[1] = mesh.data.numElements;
        out[2] = mesh.data.nodesPerElement; out[3] = topo->numFreeDOFs();
        out[4] = K ? (int64_t)K->rowPtr()[K->nRows()] : 0;
    }
    void dofs(int32_t* t2a, int32_t* ctag) override {
        for (Index i = 0; i < mesh.data.numNodes; ++i) {
            t2a[i] = topo->isConstrained(i) ? -1 : (int32_t)topo->toAlgebraic(i);
            ctag[i] = topo->getConstraintTag(i);
        }
    }
    void map(int64_t* m) override {
        const Index npe = mesh.data.nodesPerElement;
        for (Index e = 0; e < mesh.data.numElements; ++e) {
            const Index* nd = mesh.getElementNodes(e);
            for (Index a = 0; a < npe; ++a)
                for (Index b = 0; b < npe; ++b) {
                    int64_t v = -1;
                    if (!topo->isConstrained(nd[a]) && !topo->isConstrained(nd[b]))
                        v = (int64_t)K->getDataIndex(topo->toAlgebraic(nd[a]), topo->toAlgebraic(nd[b]));
                    m[(e * npe + a) * npe + b] = v;
                }
        }
    }
    VecT vec_from(const Real* U) {
        VecT v(topo->numFreeDOFs());
        for (Index i = 0; i < v.size(); ++i) v.data()[i] = U ? U[i] : 0.0;
        return v;
    }
    // set: 0 diffusion, 1 mass, 2 BE operator (diffusion + coef*mass), 3 Jacobian
    void assemble_matrix(int set, Real coef, const Real* U, Real* val) override {
        VecT u = vec_from(U);
        State st; st.valueSource = &u;
        const Index nnz = K->rowPtr()[K->nRows()];
        Cond c = cond();
        if (set == 0) asmb.template assembleMatrix<1, EvalEle, QP, Cond, StiffForms, Quad, State>(mesh, *topo, 0.0, c, StiffForms{}, ele, quad, st, *K, &bcs);
        else if (set == 1) asmb.template assembleMatrix<1, EvalEle, QP, MassModel, MassForms, Quad, State>(mesh, *topo, 0.0, mass_model(), MassForms{}, ele, quad, st, *K, &bcs);
        else if (set == 2) {
            BEForms f(StiffForms{}, fem::form::ScaledForm<1, MassForms>(coef));
            OpModel om(c, mass_model());
            asmb.template assembleMatrix<1, EvalEle, QP, OpModel, BEForms, Quad, State>(mesh, *topo, 0.0, om, f, ele, quad, st, *K, &bcs);
        } else {
            asmb.template assembleMatrix<1, EvalEle, QP, Cond, JacForms, Quad, State>(mesh, *topo, 0.0, c, JacForms{}, ele, quad, st, *K, &bcs);
        }
        for (Index i = 0; i < nnz; ++i) val[i] = K->data()[i];
    }
    // set: 0 K(U)U (diffusion, Full gather), 1 M U (mass)
    void assemble_vector(int set, Real, const Real* U, Real* F) override {
        VecT u = vec_from(U);
        VecT out(topo->numFreeDOFs());
        State st; st.valueSource = &u;
        auto sel = [](State& s) -> Real* { return s.valueBuffer(); };
        Cond c = cond();
        if (set == 0)
            asmb.template assembleVector<1, EvalEle, QP, Cond, Quad, State, fem::assembly::FormsWithOperand<StiffForms, decltype(sel)>>(mesh, *topo, 0.0, c, ele, quad, st, out, &bcs, fem::assembly::FormsWithOperand<StiffForms, decltype(sel)>{StiffForms{}, sel});
        else {
            MassForms mf;
            asmb.template assembleVector<1, EvalEle, QP, MassModel, Quad, State, fem::assembly::FormsWithOperand<MassForms, decltype(sel)>>(mesh, *topo, 0.0, mass_model(), ele, quad, st, out, &bcs, fem::assembly::FormsWithOperand<MassForms, decltype(sel)>{mf, sel});
        }
        for (Index i = 0; i < out.size(); ++i) F[i] = out.data()[i];
    }
    void load(int kind, const Real* nodal, Real* F) override {
        VecT u = vec_from(nullptr);
        VecT out(topo->numFreeDOFs());
        State st; st.valueSource = &u;
        auto sel = [](State& s) -> Real* { return s.valueBuffer(); };
        struct Default { void eval(QP&) const {} void evalGradient(QP&) const {} } dm;
        if (nodal) {
            using NS = equation::heateq::NodalSourceForm<QP, NodalArray, 1>;
            using NSF = fem::form::FormRegistry<NS>;
            NodalArray na; na.data = nodal;
            NSF f{NS{na}};
            asmb.template assembleVector<1, EvalEle, QP, Default, Quad, State, fem::assembly::FormsWithOperand<NSF, decltype(sel)>>(mesh, *topo, 0.0, dm, ele, quad, st, out, &bcs, fem::assembly::FormsWithOperand<NSF, decltype(sel)>{f, sel});
        } else {
            using SF = equation::heateq::SourceForm<QP, equation::heateq::evaluator::SourceFunction<D, 1, KindFn<D>>>;
            using SFR = fem::form::FormRegistry<SF>;
            KindFn<D> fn; fn.kind = kind;
            SFR f{SF{fn}};
            asmb.template assembleVector<1, EvalEle, QP, Default, Quad, State, fem::assembly::FormsWithOperand<SFR, decltype(sel)>>(mesh, *topo, 0.0, dm, ele, quad, st, out, &bcs, fem::assembly::FormsWithOperand<SFR, decltype(sel)>{f, sel});
        }
        for (Index i = 0; i < out.size(); ++i) F[i] = out.data()[i];
    }
    void lift(int set, Real coef, Real* F) override {
        VecT out = vec_from(F);
        Cond c = cond();
        if (set == 0) app.template applyEssentialBCs<1, EvalEle, QP, Cond, StiffForms, Quad>(mesh, *topo, bcs, 0.0, c, StiffForms{}, ele, quad, out);
        else {
            BEForms f(StiffForms{}, fem::form::ScaledForm<1, MassForms>(coef));
            OpModel om(c, mass_model());
            app.template applyEssentialBCs<1, EvalEle, QP, OpModel, BEForms, Quad>(mesh, *topo, bcs, 0.0, om, f, ele, quad, out);
        }
        for (Index i = 0; i < out.size(); ++i) F[i] = out.data()[i];
    }

    // Assembler::assembleDiagonal (Assembler.tpp:142-197); set as assemble_matrix (0 diffusion, 1 mass, 2 BE)
    void assemble_diagonal(int set, Real coef, const Real* U, Real* d) override {
        VecT u = vec_from(U);
        VecT out(topo->numFreeDOFs());
        State st; st.valueSource = &u;
        Cond c = cond();
        if (set == 0) asmb.template assembleDiagonal<1, EvalEle, QP, Cond, StiffForms, Quad, State>(mesh, *topo, 0.0, c, StiffForms{}, ele, quad, st, out, &bcs);
        else if (set == 1) asmb.template assembleDiagonal<1, EvalEle, QP, MassModel, MassForms, Quad, State>(mesh, *topo, 0.0, mass_model(), MassForms{}, ele, quad, st, out, &bcs);
        else {
            BEForms f(StiffForms{}, fem::form::ScaledForm<1, MassForms>(coef));
            OpModel om(c, mass_model());
            asmb.template assembleDiagonal<1, EvalEle, QP, OpModel, BEForms, Quad, State>(mesh, *topo, 0.0, om, f, ele, quad, st, out, &bcs);
        }
        for (Index i = 0; i < out.size(); ++i) d[i] = out.data()[i];
    }

    // QuantityEvaluator<CPU>::evaluateBoundaryReaction (QuantityEvaluator.tpp:95-199) with the
    // weak heat-flux pairs of HeatProblem.tpp:385-410 (source, -stiffness, -mass on the rate)
    void boundary_reaction(Real time, int src_kind, const Real* U, const Real* Udot, int n_tags, const int32_t* tags,
                           Real* result, Real* measure) override {
        using QuadBdy = std::conditional_t<D == 2, fem::quadrature::GaussQuadrature1D, fem::quadrature::GaussQuadratureQuad>;
        using QPB = equation::heateq::evaluator::EvalQuadraturePointBoundary<EvalEle, Basis, fem::geometry::JacobianTransform<D, D>>;
        using SF = equation::heateq::SourceForm<QP, equation::heateq::evaluator::SourceFunction<D, 1, KindFn<D>>>;
        using SFR = fem::form::FormRegistry<SF>;
        using NegStiff = fem::form::ScaledForm<1, StiffForms>;
        using NegMass = fem::form::ScaledForm<1, MassForms>;
        VecT u = vec_from(U), ud = vec_from(Udot);
        State st; st.valueSource = &u; st.rateSource = &ud;
        auto selV = [](State& s) -> Real* { return s.valueBuffer(); };
        auto selR = [](State& s) -> Real* { return s.rateBuffer(); };
        using SP = fem::assembly::FormsWithOperand<SFR, decltype(selV)>;
        using KP = fem::assembly::FormsWithOperand<NegStiff, decltype(selV)>;
        using MP = fem::assembly::FormsWithOperand<NegMass, decltype(selR)>;
        KindFn<D> fn; fn.kind = src_kind;
        SFR sf{SF{fn}};
        Cond c = cond();
        OpModel om(c, mass_model());
        const Index nb = quad.numPointsTotal() == 0 ? 1 : (Index)std::llround(std::pow((double)quad.numPointsTotal(), 1.0 / D));
        QuadBdy qb = [&] { if constexpr (D == 2) return QuadBdy(nb); else return QuadBdy(nb, nb); }();
        fem::quantity::BoundaryQuantityRegistry<1> reg;
        for (int k = 0; k < n_tags; ++k) reg.registerTag(tags[k]);
        if (Udot)
            fem::quantity::QuantityEvaluator<Backend>::template evaluateBoundaryReaction<1, EvalEle, QP, QPB, OpModel, Quad, QuadBdy, State, SP, KP, MP>(
                mesh, *topo, bcs, time, om, ele, quad, qb, st, reg, SP{sf, selV}, KP{NegStiff(Real(-1))}, MP{NegMass(Real(-1))});
        else
            fem::quantity::QuantityEvaluator<Backend>::template evaluateBoundaryReaction<1, EvalEle, QP, QPB, OpModel, Quad, QuadBdy, State, SP, KP>(
                mesh, *topo, bcs, time, om, ele, quad, qb, st, reg, SP{sf, selV}, KP{NegStiff(Real(-1))});
        for (int k = 0; k < n_tags; ++k) {
            result[k] = reg.result(tags[k])[0];
            measure[k] = reg.entries().at(tags[k]).measureSum;
        }
    }

    // QuantityEvaluator<CPU>::evaluateBoundaryRegistry (QuantityEvaluator.tpp:14-93) with
    // HeatFluxIntegrand, Integral reduction, on the boundary state HeatBoundaryStateVector
    void boundary_flux(Real time, const Real* U, int n_tags, const int32_t* tags, Real* result,
                       Real* measure) override {
        using QuadBdy = std::conditional_t<D == 2, fem::quadrature::GaussQuadrature1D, fem::quadrature::GaussQuadratureQuad>;
        using QPB = equation::heateq::evaluator::EvalQuadraturePointBoundary<EvalEle, Basis, fem::geometry::JacobianTransform<D, D>>;
        using BState = equation::heateq::evaluator::HeatBoundaryStateVector<1, D, D, VecT>;
        using CondB = HarnessConductivity<QPB>;
        using Flux = equation::heateq::quantity::HeatFluxIntegrand<QPB>;
        using Qs = fem::quantity::QuantityForms<fem::quantity::ReducedQuantity<Flux, fem::quantity::Reduction::Integral>>;
        VecT u = vec_from(U);
        BState st; st.valueSource = &u;
        CondB c; c.tdep = tdep; c.ka = ka; c.kb = kb; c.aniso = aniso;
        for (int i = 0; i < 9; ++i) c.tensor[i] = tensor[i];
        Qs qf{Flux{}};
        const Index nb = quad.numPointsTotal() == 0 ? 1 : (Index)std::llround(std::pow((double)quad.numPointsTotal(), 1.0 / D));
        QuadBdy qb = [&] { if constexpr (D == 2) return QuadBdy(nb); else return QuadBdy(nb, nb); }();
        fem::quantity::BoundaryQuantityRegistry<1> reg;
        for (int k = 0; k < n_tags; ++k) reg.registerTag(tags[k]);
        fem::quantity::QuantityEvaluator<Backend>::template evaluateBoundaryRegistry<1, EvalEle, QPB, CondB, Qs, QuadBdy, BState>(
            mesh, *topo, bcs, time, c, qf, ele, qb, st, reg);
        for (int k = 0; k < n_tags; ++k) {
            result[k] = reg.result(tags[k])[0];
            measure[k] = reg.entries().at(tags[k]).measureSum;
        }
    }

    // BoundaryApplicator<CPU>::applyNaturalBCs (BoundaryApplicator.tpp:66-140): a
    // NaturalBoundaryRegistry with one FluxBoundaryForm (flux kind) or
    // NodalFluxForm (nodal vectors) entry per (tag, kind) pair, in order
    void natural(Real time, int n_entries, const int32_t* tags, const int32_t* kinds, const Real* const* nodal,
                 Real* F) override {
        using QuadBdy = std::conditional_t<D == 2, fem::quadrature::GaussQuadrature1D, fem::quadrature::GaussQuadratureQuad>;
        using QPB = equation::heateq::evaluator::EvalQuadraturePointBoundary<EvalEle, Basis, fem::geometry::JacobianTransform<D, D>>;
        using FluxFn = equation::heateq::BoundaryFluxFunction<D, 1, FluxKindFn<D>>;
        using FluxForms = fem::form::FormRegistry<equation::heateq::FluxBoundaryForm<QPB, FluxFn>>;
        using NodalForms = fem::form::FormRegistry<equation::heateq::NodalFluxForm<QPB, NodalVec<D>, 1, D>>;
        struct DefaultB { void eval(QPB&) const {} void evalGradient(QPB&) const {} };
        const Index nb = (Index)std::llround(std::pow((double)quad.numPointsTotal(), 1.0 / D));
        QuadBdy qb = [&] { if constexpr (D == 2) return QuadBdy(nb); else return QuadBdy(nb, nb); }();
        fem::boundary::NaturalBoundaryRegistry<QPB> reg;
        DefaultB model;
        std::vector<std::unique_ptr<FluxForms>> ff;
        std::vector<std::unique_ptr<NodalForms>> nf;
        for (int k = 0; k < n_entries; ++k) {
            FluxKindFn<D> fn; fn.kind = kinds[k];
            auto bc = std::shared_ptr<fem::boundary::BoundaryCondition<FluxFn>>(new fem::boundary::BoundaryCondition<FluxFn>{
                tags[k], {fem::boundary::BCCategory::Natural}, FluxFn{fn}});
            if (nodal && nodal[k]) {
                NodalVec<D> nv; nv.data = nodal[k];
                nf.push_back(std::make_unique<NodalForms>(equation::heateq::NodalFluxForm<QPB, NodalVec<D>, 1, D>{nv}));
                reg.template registerBC<FluxFn>(bc, *nf.back(), model);
            } else {
                ff.push_back(std::make_unique<FluxForms>(equation::heateq::FluxBoundaryForm<QPB, FluxFn>{fn}));
                reg.template registerBC<FluxFn>(bc, *ff.back(), model);
            }
        }
        VecT out = vec_from(F);
        app.template applyNaturalBCs<1, EvalEle, QPB, QuadBdy>(mesh, *topo, reg, time, ele, qb, out);
        for (Index i = 0; i < out.size(); ++i) F[i] = out.data()[i];
    }

    // cg::Solver over the matrix-free FEMOperator (FEMOperator.hpp:27-71); set 0 diffusion, 2 BE
    template<typename ModelT, typename FormsT>
    int cg_fem_t(const ModelT& model, const FormsT& forms, const Real* U, int pc, const Real* rhs, Real* x, Real tol,
                 int64_t maxit, int64_t* iters) {
        const Index n = topo->numFreeDOFs();
        VecT u = vec_from(U), B = vec_from(rhs), X = vec_from(x);
        State st; st.valueSource = &u;
        Real time = 0.0;
        using Op = linalg::op::FEMOperator<fem::assembly::Assembler<Backend>, topology::TopologicalDOF<1>, EvalEle, QP,
                                           ModelT, FormsT, Quad, State, VecT>;
        Op op(asmb, mesh, *topo, &time, model, forms, ele, quad, &bcs, st);
        utils::logging::NullLogger lg;
        linalg::solver::SolverReport<VecT> rep;
        auto run = [&](auto pcT) {
            using PC = decltype(pcT);
            PC M;
            using S = linalg::solver::iterative::cg::Solver<Op, VecT, PC, utils::logging::NullLogger>;
            typename S::Config cfg; cfg.tol = tol; cfg.maxIters = (Index)maxit;
            S s(cfg); typename S::Workspace W(n);
            return s.solve(rep, lg, W, M, op, B, X);
        };
        const bool ok = pc == 1 ? run(linalg::solver::preconditioner::Jacobi<VecT>{})
                                : run(linalg::solver::preconditioner::Identity<VecT>{});
        for (Index i = 0; i < n; ++i) x[i] = X.data()[i];
        if (iters) *iters = (int64_t)rep.iterations;
        return ok ? 1 : 0;
    }
    int cg_fem(int set, Real coef, const Real* U, int pc, const Real* rhs, Real* x, Real tol, int64_t maxit,
               int64_t* iters) override {
        Cond c = cond();
        if (set == 0) return cg_fem_t(c, StiffForms{}, U, pc, rhs, x, tol, maxit, iters);
        BEForms f(StiffForms{}, fem::form::ScaledForm<1, MassForms>(coef));
        OpModel om(c, mass_model());
        return cg_fem_t(om, f, U, pc, rhs, x, tol, maxit, iters);
    }

    template<typename PC>
    bool cg(const MatT& A, const VecT& b, VecT& x, Real tol, int64_t maxit, int64_t* iters) {
        using Op = linalg::op::CSROperator<MatT>;
        using S = linalg::solver::iterative::cg::Solver<Op, VecT, PC, utils::logging::NullLogger>;
        typename S::Config cfg; cfg.tol = tol; cfg.maxIters = (Index)maxit;
        S solver(cfg);
        typename S::Workspace W(b.size());
        PC M;
        utils::logging::NullLogger lg;
        linalg::solver::SolverReport<VecT> rep;
        Op op(A);
        bool ok = solver.solve(rep, lg, W, M, op, b, x);
        if (iters) *iters = (int64_t)rep.iterations;
        return ok;
    }

    // BackwardEulerStage::assemble/solve/advance + BackwardEulerRunner::step
    // (BackwardEulerStage.hpp:221-258, BackwardEuler.hpp:132-172), linear, CSR, CG+Jacobi
    int be_run(Real t0, Real tf, Real dtp, Real tol, int64_t maxit, Real* Uio, int64_t* iters) override {
        const Index n = topo->numFreeDOFs();
        VecT U = vec_from(Uio), Up = vec_from(Uio), Ud(n), F(n), ms(n);
        Ud.zero();
        State st; st.valueSource = &U; st.rateSource = &Ud;
        Real time = t0;
        int64_t step = 0;
        Cond c = cond();
        while (!(time >= tf - Real(1e-9))) {
            const Real dt = std::min(dtp, tf - time);
            const Real ct = time + dt;
            BEForms f(StiffForms{}, fem::form::ScaledForm<1, MassForms>(Real(1) / dt));
            OpModel om(c, mass_model());
            linalg::operations::copy(U, Ud);
            linalg::operations::axpby(-Real(1) / dt, Real(1) / dt, Up, Ud);
            asmb.template assembleMatrix<1, EvalEle, QP, OpModel, BEForms, Quad, State>(mesh, *topo, ct, om, f, ele, quad, st, *K, &bcs);
            F.zero();
            State ps; ps.valueSource = &Up;
            auto sel = [](State& s) -> Real* { return s.valueBuffer(); };
            MassForms mf;
            asmb.template assembleVector<1, EvalEle, QP, MassModel, Quad, State, fem::assembly::FormsWithOperand<MassForms, decltype(sel)>>(mesh, *topo, ct - dt, mass_model(), ele, quad, ps, ms, &bcs, fem::assembly::FormsWithOperand<MassForms, decltype(sel)>{mf, sel});
            linalg::operations::axpy(Real(1) / dt, ms, F);
            app.template applyEssentialBCs<1, EvalEle, QP, OpModel, BEForms, Quad>(mesh, *topo, bcs, ct, om, f, ele, quad, F);
            int64_t it = 0;
            bool ok = cg<linalg::solver::preconditioner::Jacobi<VecT>>(*K, F, U, tol, maxit, &it);
            if (iters) iters[step] = it;
            if (!ok) return -1;
            linalg::operations::copy(U, Up);
            time += dt;
            ++step;
        }
        for (Index i = 0; i < n; ++i) Uio[i] = U.data()[i];
        return (int)step;
    }

    // NewtonRunner::solve over SteadyStage (Newton.hpp:32-83, SteadyStage.hpp:60-116)
    int newton(Real abs_tol, Real rel_tol, int64_t max_newton, int src, Real lin_tol, int64_t lin_max,
               int64_t krylov, Real* Uio, int64_t* lin_iters) override {
        const Index n = topo->numFreeDOFs();
        VecT U = vec_from(Uio), F(n), R(n), dU(n);
        State st; st.valueSource = &U;
        Cond c = cond();
        Real res0 = 0;
        int result = -1;
        for (int64_t it = 0; it < max_newton; ++it) {
            std::vector<Real> f(n);
            load(src, nullptr, f.data());
            for (Index i = 0; i < n; ++i) F.data()[i] = f[i];
            auto sel = [](State& s) -> Real* { return s.valueBuffer(); };
            asmb.template assembleVector<1, EvalEle, QP, Cond, Quad, State, fem::assembly::FormsWithOperand<StiffForms, decltype(sel)>>(mesh, *topo, 0.0, c, ele, quad, st, R, &bcs, fem::assembly::FormsWithOperand<StiffForms, decltype(sel)>{StiffForms{}, sel});
            linalg::operations::axpby(Real(1), Real(-1), F, R);
            const Real rn = linalg::operations::norm(R);
            if (it == 0) res0 = rn;
            const Real rr = res0 > 0 ? rn / res0 : 0;
            if (rn < abs_tol || rr < rel_tol) { result = (int)it; break; }
            asmb.template assembleMatrix<1, EvalEle, QP, Cond, JacForms, Quad, State>(mesh, *topo, 0.0, c, JacForms{}, ele, quad, st, *K, &bcs);
            dU.zero();
            using Op = linalg::op::CSROperator<MatT>;
            using PC = linalg::solver::preconditioner::Jacobi<VecT>;
            using S = linalg::solver::iterative::gmres::Solver<Op, VecT, PC, utils::logging::NullLogger>;
            typename S::Config cfg; cfg.tol = lin_tol; cfg.maxIters = (Index)lin_max; cfg.krylovDim = (Index)krylov;
            S solver(cfg);
            typename S::Workspace W(n, (Index)krylov);
            PC M; utils::logging::NullLogger lg; linalg::solver::SolverReport<VecT> rep; Op op(*K);
            bool ok = solver.solve(rep, lg, W, M, op, R, dU);
            if (lin_iters) lin_iters[it] = (int64_t)rep.iterations;
            if (!ok) { result = -2; break; }
            linalg::operations::axpy(Real(1), dU, U);
        }
        for (Index i = 0; i < n; ++i) Uio[i] = U.data()[i];
        return result;
    }
};

} // namespace

extern "C" {

void* ref_create(int dim, int64_t nx, int64_t ny, int64_t nz, const double* ext, int64_t px,
                 int64_t py, int64_t pz, int nq, int ntags, const int32_t* tags, const int32_t* kinds) {
    Base* b;
    if (dim == 2) {
        auto* c = new Ctx<2>(fem::basis::LagrangeQuad(px, py), fem::quadrature::GaussQuadratureQuad(nq, nq));
        c->mesh = mesh::generator::BlockMesh2D(nx, ny, ext[0], ext[1], ext[2], ext[3], px, py).generate();
        for (int k = 0; k < ntags; ++k) {
            using BC = fem::boundary::BoundaryCondition<Ctx<2>::DirBC>;
            KindFn<2> fn; fn.kind = kinds[k];
            auto bc = std::make_shared<BC>(BC{tags[k], {fem::boundary::BCCategory::Essential}, Ctx<2>::DirBC{fn}});
            c->bc_keep.push_back(bc);
            c->bcs.registerBC<Ctx<2>::DirBC>(bc);
        }
        c->topo = std::make_unique<topology::TopologicalDOF<1>>(c->mesh, fem::dof::DOFOrdering::Interleaved);
        c->topo->buildConstraints(c->basis, c->bcs);
        c->K = std::make_unique<MatT>(fem::assembly::Assembler<Backend>::createMatrix<1>(c->mesh, *c->topo));
        b = c;
    } else {
        auto* c = new Ctx<3>(fem::basis::LagrangeHex(px, py, pz), fem::quadrature::GaussQuadratureHex(nq, nq, nq));
        c->mesh = mesh::generator::BlockMesh3D(nx, ny, nz, ext[0], ext[1], ext[2], ext[3], ext[4], ext[5], px, py, pz).generate();
        for (int k = 0; k < ntags; ++k) {
            using BC = fem::boundary::BoundaryCondition<Ctx<3>::DirBC>;
            KindFn<3> fn; fn.kind = kinds[k];
            auto bc = std::make_shared<BC>(BC{tags[k], {fem::boundary::BCCategory::Essential}, Ctx<3>::DirBC{fn}});
            c->bc_keep.push_back(bc);
            c->bcs.registerBC<Ctx<3>::DirBC>(bc);
        }
        c->topo = std::make_unique<topology::TopologicalDOF<1>>(c->mesh, fem::dof::DOFOrdering::Interleaved);
        c->topo->buildConstraints(c->basis, c->bcs);
        c->K = std::make_unique<MatT>(fem::assembly::Assembler<Backend>::createMatrix<1>(c->mesh, *c->topo));
        b = c;
    }
    b->node0.assign(b->mesh.data.numNodes, -1);
    for (Index e = 0; e < b->mesh.data.numElements; ++e) b->node0[b->mesh.getElementNodes(e)[0]] = (int64_t)e;
    return b;
}

void ref_destroy(void* h) { delete static_cast<Base*>(h); }

void ref_set_model(void* h, int tdep, double ka, double kb, int aniso, const double* tensor, double rho, double cp) {
    Base* b = static_cast<Base*>(h);
    b->tdep = tdep; b->ka = ka; b->kb = kb; b->aniso = aniso; b->rho = rho; b->cp = cp;
    if (tensor) for (int i = 0; i < 9; ++i) b->tensor[i] = tensor[i];
}

void ref_sizes(void* h, int64_t* out) { static_cast<Base*>(h)->sizes(out); }

void ref_mesh(void* h, double* xyz, int64_t* ien, int32_t* rng) {
    Base* b = static_cast<Base*>(h);
    const auto& d = b->mesh.data;
    for (Index i = 0; i < d.xyz.size(); ++i) xyz[i] = d.xyz[i];
    for (Index i = 0; i < d.ien.size(); ++i) ien[i] = (int64_t)d.ien[i];
    for (Index i = 0; i < d.rng.size(); ++i) rng[i] = d.rng[i];
}

void ref_dofs(void* h, int32_t* t2a, int32_t* ctag) { static_cast<Base*>(h)->dofs(t2a, ctag); }

void ref_pattern(void* h, int64_t* rp, int64_t* ci) {
    Base* b = static_cast<Base*>(h);
    const Index n = b->K->nRows();
    for (Index i = 0; i <= n; ++i) rp[i] = (int64_t)b->K->rowPtr()[i];
    for (Index i = 0; i < b->K->rowPtr()[n]; ++i) ci[i] = (int64_t)b->K->colIdx()[i];
}

void ref_map(void* h, int64_t* m) { static_cast<Base*>(h)->map(m); }

static void with_kappa(Base* b, const double* kappa) {
    g_kappa = kappa;
    g_node0_to_elem = b->node0.data();
}

void ref_assemble_matrix(void* h, int set, double coef, const double* U, const double* kappa, double* val) {
    Base* b = static_cast<Base*>(h);
    with_kappa(b, kappa);
    b->assemble_matrix(set, coef, U, val);
    g_kappa = nullptr;
}

void ref_assemble_vector(void* h, int set, const double* U, const double* kappa, double* F) {
    Base* b = static_cast<Base*>(h);
    with_kappa(b, kappa);
    b->assemble_vector(set, 0.0, U, F);
    g_kappa = nullptr;
}

void ref_load(void* h, int kind, const double* nodal, double* F) { static_cast<Base*>(h)->load(kind, nodal, F); }

void ref_lift(void* h, int set, double coef, const double* kappa, double* F) {
    Base* b = static_cast<Base*>(h);
    with_kappa(b, kappa);
    b->lift(set, coef, F);
    g_kappa = nullptr;
}

// CG / GMRES on the handle's pattern with caller-supplied values
int ref_solve(void* h, int solver, int pc, const double* val, const double* rhs, double* x, double tol,
              int64_t maxit, int64_t krylov, int64_t* iters) {
    Base* b = static_cast<Base*>(h);
    const Index n = b->K->nRows(), nnz = b->K->rowPtr()[n];
    for (Index i = 0; i < nnz; ++i) b->K->data()[i] = val[i];
    VecT B(n), X(n);
    for (Index i = 0; i < n; ++i) { B.data()[i] = rhs[i]; X.data()[i] = x[i]; }
    using Op = linalg::op::CSROperator<MatT>;
    Op op(*b->K);
    utils::logging::NullLogger lg;
    linalg::solver::SolverReport<VecT> rep;
    bool ok;
    auto run = [&](auto pcT) {
        using PC = decltype(pcT);
        PC M;
        if (solver == 0) {
            using S = linalg::solver::iterative::cg::Solver<Op, VecT, PC, utils::logging::NullLogger>;
            typename S::Config cfg; cfg.tol = tol; cfg.maxIters = (Index)maxit;
            S s(cfg); typename S::Workspace W(n);
            return s.solve(rep, lg, W, M, op, B, X);
        }
        using S = linalg::solver::iterative::gmres::Solver<Op, VecT, PC, utils::logging::NullLogger>;
        typename S::Config cfg; cfg.tol = tol; cfg.maxIters = (Index)maxit; cfg.krylovDim = (Index)krylov;
        S s(cfg); typename S::Workspace W(n, (Index)krylov);
        return s.solve(rep, lg, W, M, op, B, X);
    };
    if (pc == 1) ok = run(linalg::solver::preconditioner::Jacobi<VecT>{});
    else ok = run(linalg::solver::preconditioner::Identity<VecT>{});
    for (Index i = 0; i < n; ++i) x[i] = X.data()[i];
    if (iters) *iters = (int64_t)rep.iterations;
    return ok ? 1 : 0;
}

void ref_assemble_diagonal(void* h, int set, double coef, const double* U, const double* kappa, double* d) {
    Base* b = static_cast<Base*>(h);
    with_kappa(b, kappa);
    b->assemble_diagonal(set, coef, U, d);
    g_kappa = nullptr;
}

int ref_cg_fem(void* h, int set, double coef, const double* U, const double* kappa, int pc, const double* rhs,
               double* x, double tol, int64_t maxit, int64_t* iters) {
    Base* b = static_cast<Base*>(h);
    with_kappa(b, kappa);
    const int r = b->cg_fem(set, coef, U, pc, rhs, x, tol, maxit, iters);
    g_kappa = nullptr;
    return r;
}

void ref_boundary_reaction(void* h, double time, int src_kind, const double* U, const double* Udot,
                           const double* kappa, int n_tags, const int32_t* tags, double* result, double* measure) {
    Base* b = static_cast<Base*>(h);
    with_kappa(b, kappa);
    b->boundary_reaction(time, src_kind, U, Udot, n_tags, tags, result, measure);
    g_kappa = nullptr;
}

void ref_boundary_flux(void* h, double time, const double* U, int n_tags, const int32_t* tags, double* result,
                       double* measure) {
    Base* b = static_cast<Base*>(h);
    with_kappa(b, nullptr);
    b->boundary_flux(time, U, n_tags, tags, result, measure);
}

void ref_natural(void* h, double time, int n_entries, const int32_t* tags, const int32_t* kinds,
                 const double* const* nodal, double* F) {
    Base* b = static_cast<Base*>(h);
    with_kappa(b, nullptr);
    b->natural(time, n_entries, tags, kinds, nodal, F);
}

int ref_be_run(void* h, double t0, double tf, double dt, double tol, int64_t maxit, const double* kappa,
               double* U, int64_t* iters) {
    Base* b = static_cast<Base*>(h);
    with_kappa(b, kappa);
    int r = b->be_run(t0, tf, dt, tol, maxit, U, iters);
    g_kappa = nullptr;
    return r;
}

int ref_newton(void* h, double abs_tol, double rel_tol, int64_t max_newton, int src, double lin_tol,
               int64_t lin_max, int64_t krylov, double* U, int64_t* lin_iters) {
    return static_cast<Base*>(h)->newton(abs_tol, rel_tol, max_newton, src, lin_tol, lin_max, krylov, U, lin_iters);
}

} // extern "C"
