// residuum/hip/HeatProblem.hpp -- the heat problem on either backend tag.
//
// hip::HeatProblem<BackendT, BundleT> is the reference's HeatProblem
// (application/heateq/problem/HeatProblem.hpp:35-205, HeatProblem.tpp:3-289)
// with the same constructor (a HeatConfig, the mesh, basis and the two Gauss
// rules) and the members the solver stages use: it satisfies
// solver::problem::Problem (solver/problem/Problem.hpp:16-50) and
// TransientCapableProblem (TransientCapableProblem.hpp:13-25), so the
// reference's SteadyStage, BackwardEulerStage, NewtonRunner and
// BackwardEulerRunner drive it unchanged.  Everything goes through
// Assembler<BackendT>, BoundaryApplicator<BackendT> and the linear solver
// factory (solver/linear/LinearSolverFactory.hpp:38-118), so on the CPU tag
// it runs the reference's own CPU code and on the Hip tag the device.
//
// hip::HeatEquation is the reference's bundle (equation/heateq/
// HeatEquation.hpp:45-150) with device-evaluable functions and models in
// place of the exprtk-backed ones: hip::Expression for sources and boundary
// values, hip::ConductivityModel / hip::SpecificHeatModel, and a
// hip::NaturalBoundaryRegistry for flux conditions.  The forms are the
// reference's.  Monitors and field output are not part of this problem: a
// configuration that asks for them is refused.
#pragma once

#include <array>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

// the Hip overloads of operations::* before any reference template that calls them
#include "residuum/hip/Solvers.hpp"

#include "application/heateq/config/HeatConfig.hpp"
#include "equation/heateq/boundary/BoundaryValueFunction.hpp"
#include "equation/heateq/evaluator/DefaultModel.hpp"
#include "equation/heateq/evaluator/DensityModel.hpp"
#include "equation/heateq/evaluator/EvalElement.hpp"
#include "equation/heateq/evaluator/EvalQuadraturePointBoundary.hpp"
#include "equation/heateq/evaluator/EvalQuadraturePointVolume.hpp"
#include "equation/heateq/evaluator/HeatStateVector.hpp"
#include "equation/heateq/evaluator/SourceFunction.hpp"
#include "fem/boundary/EssentialBoundaryRegistry.hpp"
#include "fem/dispatch/DiscretizationDispatch.hpp"
#include "fem/evaluator/ModelRegistry.hpp"
#include "fem/geometry/JacobianTransform.hpp"
#include "io/fieldio/NodalFileValueSource.hpp"
#include "io/fieldio/NodalValueSourceAdapter.hpp"
#include "linalg/operator/CSROperator.hpp"
#include "linalg/operator/FEMOperator.hpp"
#include "mesh/ElementFamily.hpp"
#include "mesh/Mesh.hpp"
#include "solver/SolverInstance.hpp"
#include "solver/linear/LinearSolverFactory.hpp"
#include "topology/TopologicalDOF.hpp"

#include "residuum/hip/BoundaryApplicator.hpp"
#include "residuum/hip/Solvers.hpp"

namespace residuum {
namespace hip {

template<Index NSD, Index NPD, mesh::ElementFamily Family>
struct HeatEquation {
    static constexpr Index NumDOFs = 1;
    static constexpr Index SpatialDim = NSD;

    using ElementTraits = fem::dispatch::ElementTypeTraits<Family>;
    using Basis = typename ElementTraits::BasisType;
    using QuadratureVolumeType = typename ElementTraits::QuadratureVolumeType;
    using QuadratureBoundaryType = typename ElementTraits::QuadratureBoundaryType;
    using Transform = fem::geometry::JacobianTransform<NSD, NPD>;
    using EvalEle = equation::heateq::evaluator::EvalElement<Basis, NSD>;
    using EvalQPVol = equation::heateq::evaluator::EvalQuadraturePointVolume<EvalEle, Basis, Transform>;
    using EvalQPBdy = equation::heateq::evaluator::EvalQuadraturePointBoundary<EvalEle, Basis, Transform>;

    using DefaultModel = equation::heateq::evaluator::DefaultModel<EvalQPVol>;
    using ConductivityModel = hip::ConductivityModel<EvalQPVol>;
    using DensityModel = equation::heateq::evaluator::DensityModel<EvalQPVol>;
    using SpecificHeatModel = hip::SpecificHeatModel<EvalQPVol>;
    using MassModel = fem::evaluator::ModelRegistry<DensityModel, SpecificHeatModel>;

    using DiffusionForm = equation::heateq::DiffusionForm<EvalQPVol>;
    using StiffnessForms = fem::form::FormRegistry<DiffusionForm>;
    using MassForm = equation::heateq::MassForm<EvalQPVol>;
    using MassForms = fem::form::FormRegistry<MassForm>;
    using TangentDiffusionForm = equation::heateq::TangentDiffusionForm<EvalQPVol>;
    using TangentDiffusionForms = fem::form::FormRegistry<TangentDiffusionForm>;
    using TangentMassForm = equation::heateq::TangentMassForm<EvalQPVol>;
    using TangentMassForms = fem::form::FormRegistry<TangentMassForm>;

    using ScalarFunction = hip::Expression<NSD, 1>;
    using SourceForm = equation::heateq::SourceForm<EvalQPVol, equation::heateq::evaluator::SourceFunction<NSD, 1, ScalarFunction>>;
    using ExpressionSourceForms = fem::form::FormRegistry<SourceForm>;
    using NodalScalarSource = io::fieldio::NodalFileValueSource<NumDOFs>;
    using NodalSourceForms = fem::form::FormRegistry<equation::heateq::NodalSourceForm<EvalQPVol, NodalScalarSource, NumDOFs>>;
    using DirichletExpression = equation::heateq::BoundaryValueFunction<NSD, NumDOFs, ScalarFunction>;
    using DirichletNodal = equation::heateq::BoundaryValueFunction<NSD, NumDOFs, io::fieldio::NodalValueSourceAdapter<NodalScalarSource>>;
    using NaturalRegistry = hip::NaturalBoundaryRegistry<EvalQPBdy>;
};

template<typename BackendT, typename B>
class HeatProblem {
public:
    using VectorT = linalg::types::Vector<Real, BackendT>;
    using MatrixT = linalg::types::CSRMatrix<Real, BackendT>;
    static constexpr Index NumDOFs = B::NumDOFs;
    using StateVectorT = equation::heateq::evaluator::HeatStateVector<NumDOFs, B::SpatialDim, B::EvalEle::ParametricDim, VectorT>;

    // HeatProblem.tpp:3-176
    HeatProblem(const application::heateq::config::HeatConfig& config, mesh::Mesh mesh, typename B::Basis basis,
                typename B::QuadratureVolumeType quadratureVolume, typename B::QuadratureBoundaryType quadratureBoundary)
        : config_(config), solverInstance_(solver::resolveSolverInstance(config_.solver)), mesh_(std::move(mesh)),
          topoDOF_(mesh_, config_.discretization.dofOrdering), evalEle_(std::move(basis)),
          quadVol_(std::move(quadratureVolume)), quadBdy_(std::move(quadratureBoundary)) {
        namespace hc = application::heateq::config;
        using Mode = solver::config::NodalFieldReadConfig::Mode;
        if (!config_.monitors.empty() || config_.output.has_value())
            throw std::runtime_error("hip::HeatProblem: monitors and field output are not part of this problem");
        switch (config_.conductivity.type) {
        case hc::ConductivityConfig::Type::Constant: conductivity_.setConstant(config_.conductivity.value); break;
        case hc::ConductivityConfig::Type::Anisotropic: conductivity_.setAnisotropic(config_.conductivity.tensor); break;
        case hc::ConductivityConfig::Type::TemperatureDependentIsotropic:
            conductivity_.setTemperatureDependentIsotropic(config_.conductivity.valueExpression,
                                                           config_.conductivity.gradientExpression);
            break;
        case hc::ConductivityConfig::Type::TemperatureDependentAnisotropic:
            conductivity_.setTemperatureDependentAnisotropic(config_.conductivity.tensor, config_.conductivity.valueExpression,
                                                             config_.conductivity.gradientExpression);
            break;
        default: throw std::runtime_error("HeatProblem: unsupported conductivity type");
        }
        const bool transient = solver::isTransient(solverInstance_.mode);
        if (transient) {
            if (!config_.density.has_value() || !config_.specificHeat.has_value())
                throw std::runtime_error("HeatProblem: 'materials.density' and 'materials.specific_heat' are required for a transient driver");
            if (!solverInstance_.timestepper)
                throw std::runtime_error("HeatProblem: solver.timestepper config is required for a transient driver");
            density_.value = config_.density->value;
            if (config_.specificHeat->type == hc::SpecificHeatConfig::Type::Constant)
                specificHeat_.setConstant(config_.specificHeat->value);
            else if (config_.specificHeat->type == hc::SpecificHeatConfig::Type::TemperatureDependent)
                specificHeat_.setTemperatureDependent(config_.specificHeat->valueExpression, config_.specificHeat->gradientExpression);
            else
                throw std::runtime_error("HeatProblem: unsupported specific heat type");
        }
        massModel_ = typename B::MassModel(density_, specificHeat_);
        if (config_.source.read.mode == Mode::Expression)
            sourceForms_.emplace(typename B::SourceForm(typename B::ScalarFunction(config_.source.read.expression)));
        else
            nodalSourceForms_.emplace(typename B::NodalScalarSource(mesh_, config_.source.read.file));
        for (const auto& bc : config_.boundaryConditions) {
            for (auto form : bc.forms) {
                if (form == hc::BoundaryConditionConfig::Form::ValueBC) {
                    if (bc.mode == Mode::File) {
                        using T = typename B::DirichletNodal;
                        auto p = std::shared_ptr<fem::boundary::BoundaryCondition<T>>(new fem::boundary::BoundaryCondition<T>{
                            bc.boundaryID, {fem::boundary::BCCategory::Essential}, T{mesh_, bc.file}});
                        essentialBCs_.template registerBC<T>(p);
                    } else {
                        using T = typename B::DirichletExpression;
                        auto p = std::shared_ptr<fem::boundary::BoundaryCondition<T>>(new fem::boundary::BoundaryCondition<T>{
                            bc.boundaryID, {fem::boundary::BCCategory::Essential}, T{typename B::ScalarFunction(bc.expression)}});
                        essentialBCs_.template registerBC<T>(p);
                    }
                } else if (form == hc::BoundaryConditionConfig::Form::FluxBC) {
                    if (bc.mode == Mode::File) {
                        io::fieldio::NodalFileValueSource<NumDOFs * B::SpatialDim> src(mesh_, bc.file);
                        std::vector<Real> v(mesh_.data.numNodes * B::SpatialDim);
                        for (Index n = 0; n < mesh_.data.numNodes; ++n) src.eval(n, &v[n * B::SpatialDim]);
                        naturalBCs_.registerNodalFlux(bc.boundaryID, std::move(v));
                    } else {
                        if (bc.fluxExpression.size() != B::SpatialDim)
                            throw std::runtime_error("HeatProblem: flux boundary 'expression' must have SpatialDim components");
                        std::array<std::string, B::SpatialDim> q;
                        for (Index i = 0; i < B::SpatialDim; ++i) q[i] = bc.fluxExpression[i];
                        naturalBCs_.registerFlux(bc.boundaryID, Expression<B::SpatialDim, B::SpatialDim>(q));
                    }
                }
            }
        }
        topoDOF_.buildConstraints(evalEle_.basis(), essentialBCs_);
        F_ = std::make_unique<VectorT>(createVector());
        U_ = std::make_unique<VectorT>(createVector());
        U_->zero();
        U_prev_ = std::make_unique<VectorT>(createVector());
        Udot_ = std::make_unique<VectorT>(createVector());
        loadNodalField(config_.initialCondition.read, *U_);
        if (transient) linalg::operations::copy(*U_, *U_prev_);
    }

    // the device problem built for this mesh goes with it, so that a later
    // mesh at the same address cannot match it; matrices still made on it
    // (a solver's copy) are orphaned by the library and refuse further use
    ~HeatProblem() {
        if constexpr (std::is_same_v<BackendT, linalg::types::backend::Hip>) ProblemCache::release(mesh_);
    }
    HeatProblem(const HeatProblem&) = delete;
    HeatProblem& operator=(const HeatProblem&) = delete;


    typename B::StiffnessForms stiffnessForms() const { return {}; }
    typename B::MassForms massForms() const { return {}; }
    typename B::TangentDiffusionForms tangentStiffnessForms() const { return {}; }
    typename B::TangentMassForms tangentMassForms() const { return {}; }
    const typename B::ConductivityModel& stiffnessModel() const { return conductivity_; }
    const typename B::MassModel& massModel() const { return massModel_; }

    MatrixT createMatrix() const { return fem::assembly::Assembler<BackendT>::template createMatrix<NumDOFs>(mesh_, topoDOF_); }
    VectorT createVector() const { return fem::assembly::Assembler<BackendT>::template createVector<NumDOFs>(mesh_, topoDOF_); }

    // HeatProblem.tpp:216-231
    void assembleLoad(Real time) {
        F_->zero();
        StateVectorT state;
        state.valueSource = U_.get();
        auto sel = [](StateVectorT& s) -> Real* { return s.valueBuffer(); };
        typename B::DefaultModel dm;
        if (sourceForms_) {
            using FW = fem::assembly::FormsWithOperand<typename B::ExpressionSourceForms, decltype(sel)>;
            fem::assembly::Assembler<BackendT>::template assembleVector<NumDOFs, typename B::EvalEle, typename B::EvalQPVol, typename B::DefaultModel, typename B::QuadratureVolumeType, StateVectorT, FW>(
                mesh_, topoDOF_, time, dm, evalEle_, quadVol_, state, *F_, &essentialBCs_, FW{*sourceForms_, sel});
        } else {
            using FW = fem::assembly::FormsWithOperand<typename B::NodalSourceForms, decltype(sel)>;
            fem::assembly::Assembler<BackendT>::template assembleVector<NumDOFs, typename B::EvalEle, typename B::EvalQPVol, typename B::DefaultModel, typename B::QuadratureVolumeType, StateVectorT, FW>(
                mesh_, topoDOF_, time, dm, evalEle_, quadVol_, state, *F_, &essentialBCs_, FW{*nodalSourceForms_, sel});
        }
    }

    // HeatProblem.tpp:233-238
    void applyNatural(Real time) {
        fem::boundary::BoundaryApplicator<BackendT>::template applyNaturalBCs<NumDOFs, typename B::EvalEle, typename B::EvalQPBdy, typename B::QuadratureBoundaryType>(
            mesh_, topoDOF_, naturalBCs_, time, evalEle_, quadBdy_, *F_);
    }

    template<typename FormsT, typename ModelT>
    void assembleMatrix(Real time, const FormsT& forms, const ModelT& model, StateVectorT& state, MatrixT& K) {
        fem::assembly::Assembler<BackendT>::template assembleMatrix<NumDOFs, typename B::EvalEle, typename B::EvalQPVol, ModelT, FormsT, typename B::QuadratureVolumeType, StateVectorT>(
            mesh_, topoDOF_, time, model, forms, evalEle_, quadVol_, state, K, &essentialBCs_);
    }

    template<typename FormsT, typename ModelT, typename SelectorT>
    void assembleVector(Real time, const FormsT& forms, const ModelT& model, StateVectorT& state, SelectorT selector, VectorT& V) {
        using FW = fem::assembly::FormsWithOperand<FormsT, SelectorT>;
        fem::assembly::Assembler<BackendT>::template assembleVector<NumDOFs, typename B::EvalEle, typename B::EvalQPVol, ModelT, typename B::QuadratureVolumeType, StateVectorT, FW>(
            mesh_, topoDOF_, time, model, evalEle_, quadVol_, state, V, &essentialBCs_, FW{forms, selector});
    }

    template<typename FormsT, typename ModelT>
    void assembleResidual(Real time, const FormsT& forms, const ModelT& model, StateVectorT& state, VectorT& R) {
        assembleVector(time, forms, model, state, [](StateVectorT& s) -> Real* { return s.valueBuffer(); }, R);
    }

    template<typename FormsT, typename ModelT>
    void applyEssential(Real time, const FormsT& forms, const ModelT& model, VectorT& F) {
        fem::boundary::BoundaryApplicator<BackendT>::template applyEssentialBCs<NumDOFs, typename B::EvalEle, typename B::EvalQPVol, ModelT, FormsT, typename B::QuadratureVolumeType>(
            mesh_, topoDOF_, essentialBCs_, time, model, forms, evalEle_, quadVol_, F);
    }

    // HeatProblem.tpp:272-289
    template<typename FormsT, typename ModelT>
    std::unique_ptr<linalg::solver::LinearSolverRunner<VectorT>> makeLinearRunner(const Real* time, const FormsT& forms, const ModelT& model, StateVectorT& state, MatrixT& K) {
        using CSROperatorT = linalg::op::CSROperator<MatrixT>;
        using FEMOperatorT = linalg::op::FEMOperator<fem::assembly::Assembler<BackendT>, topology::TopologicalDOF<NumDOFs>, typename B::EvalEle, typename B::EvalQPVol, ModelT, FormsT, typename B::QuadratureVolumeType, StateVectorT, VectorT>;
        using OpType = solver::config::LinearSolverConfig::OperatorType;
        const OpType op = solverInstance_.linear->operatorType;
        if (op == OpType::CSR)
            return solver::linear::makeLinearSolverRunner<CSROperatorT, VectorT>(CSROperatorT(K), topoDOF_.numFreeDOFs(), *solverInstance_.linear, config_.logging.linear, equationLabel(), dofNames(), freeDOFsPerField(), dofOrdering());
        if (op == OpType::FEM)
            return solver::linear::makeLinearSolverRunner<FEMOperatorT, VectorT>(FEMOperatorT(assembler_, mesh_, topoDOF_, time, model, forms, evalEle_, quadVol_, &essentialBCs_, state), topoDOF_.numFreeDOFs(), *solverInstance_.linear, config_.logging.linear, equationLabel(), dofNames(), freeDOFsPerField(), dofOrdering());
        throw std::runtime_error("HeatProblem: unsupported operator");
    }

    Index numFreeDOFs() const { return topoDOF_.numFreeDOFs(); }
    std::vector<std::string> dofNames() const { return {"T"}; }
    Index freeDOFsPerField() const { return topoDOF_.numFreeDOFs() / static_cast<Index>(dofNames().size()); }
    fem::dof::DOFOrdering dofOrdering() const { return topoDOF_.ordering(); }
    const solver::SolverInstance& solverInstance() const { return solverInstance_; }
    std::string equationLabel() const { return "Heat Equation"; }
    const solver::config::LoggingConfig& loggingConfig() const { return config_.logging; }

    VectorT& U() { return *U_; }
    VectorT& F() { return *F_; }
    VectorT& U_prev() { return *U_prev_; }
    VectorT& Udot() { return *Udot_; }
    const VectorT& solution() const { return *U_; }
    void advance() { linalg::operations::copy(*U_, *U_prev_); }

    // no monitors or field output are configured (refused at construction)
    void writeOutput(Index, Real) const {}
    void evaluateMonitors(Index, Real) const {}

    const mesh::Mesh& mesh() const { return mesh_; }
    const topology::TopologicalDOF<NumDOFs>& topology() const { return topoDOF_; }
    const fem::boundary::EssentialBoundaryRegistry& essentialRegistry() const { return essentialBCs_; }

private:
    // HeatProblem.tpp:178-214: a nodal field into the free DOFs
    void loadNodalField(const solver::config::NodalFieldReadConfig& cfg, VectorT& target) const {
        std::vector<Real> h(topoDOF_.numFreeDOFs(), Real(0));
        auto scatter = [&](Index node, const Real* val) {
            const Index tdof = topoDOF_.getNodeDOF(node, 0);
            if (!topoDOF_.isConstrained(tdof)) h[topoDOF_.toAlgebraic(tdof)] = val[0];
        };
        if (cfg.mode == solver::config::NodalFieldReadConfig::Mode::Expression) {
            typename B::ScalarFunction f(cfg.expression);
            for (Index n = 0; n < mesh_.data.numNodes; ++n) {
                Real x[3] = {0, 0, 0}, v[1];
                const Real* p = mesh_.getNodeCoord(n);
                for (Index d = 0; d < B::SpatialDim; ++d) x[d] = p[d];
                f(Real(0), x, v);
                scatter(n, v);
            }
        } else {
            typename B::NodalScalarSource src(mesh_, cfg.file);
            for (Index n = 0; n < mesh_.data.numNodes; ++n) {
                Real v[1];
                src.eval(n, v);
                scatter(n, v);
            }
        }
        BackendT::template copy<Real>(target.data(), h.data(), h.size(), linalg::types::backend::CopyKind::HostToDevice);
    }

    application::heateq::config::HeatConfig config_;
    solver::SolverInstance solverInstance_;
    mesh::Mesh mesh_;
    topology::TopologicalDOF<NumDOFs> topoDOF_;
    typename B::EvalEle evalEle_;
    typename B::QuadratureVolumeType quadVol_;
    typename B::QuadratureBoundaryType quadBdy_;
    fem::assembly::Assembler<BackendT> assembler_;
    fem::boundary::EssentialBoundaryRegistry essentialBCs_;
    typename B::NaturalRegistry naturalBCs_;
    typename B::ConductivityModel conductivity_;
    typename B::DensityModel density_;
    typename B::SpecificHeatModel specificHeat_;
    typename B::MassModel massModel_;
    std::optional<typename B::ExpressionSourceForms> sourceForms_;
    std::optional<typename B::NodalSourceForms> nodalSourceForms_;
    std::unique_ptr<VectorT> F_, U_, U_prev_, Udot_;
};

}  // namespace hip
}  // namespace residuum
