// residuum/hip/Assembler.hpp -- fem::assembly::Assembler<Hip>.
//
// The reference's assembly seam is a class template specialised per backend
// (fem/assembly/Assembler.hpp:44-68; the CPU specialisation in
// fem/assembly/backend/cpu/Assembler.tpp).  This header adds
// Assembler<linalg::types::backend::Hip> with the same static members and
// signatures, so code written against the seam -- the reference's
// HeatProblem (application/heateq/problem/HeatProblem.tpp:217-270) and
// FEMOperator (linalg/operator/FEMOperator.hpp:27-71) -- runs its assembly
// on the device:
//
//   createMatrix / createVector   a CSRMatrix<Real, Hip> with the pattern
//                                 of Assembler.tpp:7-76 (bitwise), a
//                                 Vector<Real, Hip> of the free DOFs;
//   assembleMatrix                K.zero(), the element loop, the scatter
//                                 (Assembler.tpp:86-140) as the device's
//                                 coloured element kernels; values in CSR
//                                 order in K's device memory;
//   assembleVector                F.zero() and the element loop with one
//                                 (forms, selector) pair (Assembler.tpp:
//                                 199-255); source forms go to the device's
//                                 load assembly;
//   assembleDiagonal              Assembler.tpp:142-197.
//
// The reference's form types map to the device's form kinds at compile time
// (DiffusionForm, MassForm, TangentDiffusionForm, TangentMassForm, nested
// FormRegistry, ScaledForm of one form); SourceForm with a hip::Expression
// callable and NodalSourceForm are loads.  A form the device does not know
// fails to compile.  Models are read through probe_model (Models.hpp).
// Boundary values are host callables of the EssentialBoundaryRegistry: they
// are evaluated once per time at the constrained nodes, exactly as the
// reference's gather evaluates them, and handed to the device.
//
// Each (mesh, TopologicalDOF) pair gets one device problem (mesh, DOF maps,
// pattern, element->slot map, colouring), built on first use and kept for
// the thread's lifetime or until release(); its DOF maps are checked against
// the TopologicalDOF once, at creation.
#pragma once

#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

// the Hip overloads of operations::* before any reference template that calls them
#include "residuum/hip/Solvers.hpp"

#include "fem/assembly/Assembler.hpp"
#include "fem/boundary/EssentialBoundaryRegistry.hpp"
#include "fem/form/FormRegistry.hpp"
#include "fem/form/ScaledForm.hpp"
#include "equation/heateq/form/DiffusionForm.hpp"
#include "equation/heateq/form/MassForm.hpp"
#include "equation/heateq/form/NodalSourceForm.hpp"
#include "equation/heateq/form/SourceForm.hpp"
#include "equation/heateq/form/TangentDiffusionForm.hpp"
#include "equation/heateq/form/TangentMassForm.hpp"
#include "mesh/Mesh.hpp"
#include "topology/TopologicalDOF.hpp"

#include "residuum/hip/Expression.hpp"
#include "residuum/hip/Models.hpp"
#include "residuum/hip/Operations.hpp"

namespace residuum {
namespace hip {

// ---------------------------------------------------------------------------
// device problem per (mesh, TopologicalDOF)
// ---------------------------------------------------------------------------
class DeviceProblem {
public:
    // nq: Gauss points per direction (0: p + 1 until an assembly names its rule)
    template<Index numDOFs>
    DeviceProblem(const mesh::Mesh& mesh, const topology::TopologicalDOF<numDOFs>& topo, Index nq)
        : mesh_(&mesh), topo_(&topo) {
        static_assert(numDOFs == 1, "the device heat problem has one DOF per node");
        const auto& d = mesh.data;
        if (d.parametricDim != d.spatialDim || (d.spatialDim != 2 && d.spatialDim != 3))
            throw std::invalid_argument("Assembler<Hip>: quadrilateral meshes in 2D and hexahedral meshes in 3D only");
        dim_ = (int32_t)d.spatialDim;
        int32_t p[3] = {1, 1, 1};
        for (Index i = 0; i < (Index)dim_; ++i) p[i] = (int32_t)(i < d.basisOrder.size() ? d.basisOrder[i] : 1);
        // the Dirichlet tags are the constraint tags the TopologicalDOF holds:
        // a tag that never wins a node (TopologicalDOF.tpp:60-68 keeps the last)
        // constrains no node another tag does not, so the device, given these,
        // marks the same nodes with the same tags (checked below)
        std::set<int32_t> tagset;
        for (Index i = 0; i < d.numNodes; ++i)
            if (topo.isConstrained(i)) {
                cnodes_.push_back(i);
                tagset.insert((int32_t)topo.getConstraintTag(i));
            }
        std::vector<int32_t> tags(tagset.begin(), tagset.end());
        std::vector<const char*> ex(tags.size(), "0");
        nq_ = nq ? (int32_t)nq : p[0] + 1;
        check(rsd_problem_create(Context::get(), dim_, p, nq_, (int64_t)d.numNodes, (int64_t)d.numElements,
                                 d.xyz.data(), reinterpret_cast<const int64_t*>(d.ien.data()),
                                 reinterpret_cast<const int32_t*>(d.rng.data()), (int32_t)tags.size(), tags.data(),
                                 ex.data(), &pb_));
        int64_t sz[8];
        check(rsd_problem_sizes(pb_, sz));
        n_free_ = (Index)sz[3];
        std::vector<int32_t> t2a(d.numNodes), ctag(d.numNodes);
        check(rsd_problem_get_dofs(pb_, t2a.data(), nullptr, ctag.data()));
        bool same = n_free_ == topo.numFreeDOFs();
        for (Index i = 0; same && i < d.numNodes; ++i)
            same = topo.isConstrained(i) ? (t2a[i] == -1 && ctag[i] == topo.getConstraintTag(i))
                                         : (t2a[i] == (int32_t)topo.toAlgebraic(i));
        if (!same) {
            rsd_problem_destroy(pb_);
            pb_ = nullptr;
            throw std::invalid_argument(
                "Assembler<Hip>: the TopologicalDOF's numbering is not the interleaved one its constraints give");
        }
        stamp_ = stamp(mesh, topo);
    }
    DeviceProblem(const DeviceProblem&) = delete;
    DeviceProblem& operator=(const DeviceProblem&) = delete;
    ~DeviceProblem() {
        if (scratch_) rsd_matrix_destroy(scratch_);
        if (pb_) {
            detail::forget_problem(pb_);
            rsd_problem_destroy(pb_);
        }
    }

    template<Index numDOFs>
    static std::array<uint64_t, 5> stamp(const mesh::Mesh& m, const topology::TopologicalDOF<numDOFs>& t) {
        return {(uint64_t)m.data.numNodes, (uint64_t)m.data.numElements, (uint64_t)(uintptr_t)m.data.xyz.data(),
                (uint64_t)(uintptr_t)m.data.ien.data(), (uint64_t)t.numFreeDOFs()};
    }
    template<Index numDOFs>
    bool matches(const mesh::Mesh& m, const topology::TopologicalDOF<numDOFs>& t) const {
        return mesh_ == &m && topo_ == (const void*)&t && stamp_ == stamp(m, t);
    }
    bool onMesh(const mesh::Mesh* m) const { return mesh_ == m; }
    bool sameAddresses(const mesh::Mesh* m, const void* t) const { return mesh_ == m && topo_ == t; }

    rsd_problem* get() const { return pb_; }
    Index numFree() const { return n_free_; }

    // the element arithmetic of this problem (RSD_ARITH_REFERENCE, the
    // default and bitwise the CPU tag; RSD_ARITH_CONTRACTED; RSD_ARITH_UNIFORM;
    // RSD_ARITH_FACTORED)
    void setArithmetic(int32_t mode) { check(rsd_problem_set_arithmetic(pb_, mode)); }

    void ensureQuadrature(Index nq) {
        if ((int32_t)nq == nq_) return;
        check(rsd_problem_set_quadrature(Context::get(), pb_, (int32_t)nq));
        nq_ = (int32_t)nq;
    }

    // the material model, set on the device when it changed
    void setModel(const DeviceModel& m) {
        if (model_set_ && m == model_) return;
        const double* tensor = m.anisotropic ? m.tensor.data() : nullptr;
        check(rsd_problem_set_conductivity(pb_, m.k_scalar, m.k_expr.empty() ? nullptr : m.k_expr.c_str(),
                                           m.dk_expr.empty() ? nullptr : m.dk_expr.c_str(), tensor, nullptr));
        check(rsd_problem_set_mass(pb_, m.rho, m.cp));
        model_ = m;
        model_set_ = true;
    }

    // the registry's boundary values at `time` on the constrained nodes
    // (gatherElementVector, ElementMap.tpp:39-72: each entry of the node's
    // constraint tag evaluated at the node, the last one kept)
    template<Index numDOFs>
    void setDirichlet(Real time, const topology::TopologicalDOF<numDOFs>& topo,
                      const fem::boundary::EssentialBoundaryRegistry* reg) {
        if (g_valid_ && g_time_ == time && g_reg_ == (const void*)reg) return;
        std::vector<double> g(cnodes_.size(), 0.0);
        if (reg) {
            const Index SD = mesh_->data.spatialDim;
            for (size_t k = 0; k < cnodes_.size(); ++k) {
                const Index n = cnodes_[k];
                const auto* entries = reg->getEntries(topo.getConstraintTag(n));
                Real v[numDOFs] = {Real(0)};
                Real x[3] = {0, 0, 0};
                for (Index i = 0; i < SD; ++i) x[i] = mesh_->data.xyz[n * SD + i];
                if (entries)
                    for (const auto& e : *entries) e->eval(time, x, v);
                g[k] = v[0];
            }
        }
        check(rsd_problem_set_dirichlet_values(Context::get(), pb_, time, g.data()));
        g_valid_ = true;
        g_time_ = time;
        g_reg_ = reg;
    }

    // the device matrix the element kernels assemble into; the values then go
    // to a CSRMatrix<Real, Hip> in CSR order
    rsd_matrix* scratch() {
        if (!scratch_) check(rsd_matrix_create(pb_, &scratch_));
        return scratch_;
    }

private:
    const mesh::Mesh* mesh_ = nullptr;
    const void* topo_ = nullptr;
    std::array<uint64_t, 5> stamp_{};
    rsd_problem* pb_ = nullptr;
    rsd_matrix* scratch_ = nullptr;
    int32_t dim_ = 0, nq_ = 0;
    Index n_free_ = 0;
    std::vector<Index> cnodes_;
    DeviceModel model_;
    bool model_set_ = false;
    bool g_valid_ = false;
    Real g_time_ = 0;
    const void* g_reg_ = nullptr;
};

// the device problems of this thread, one per (mesh, TopologicalDOF)
class ProblemCache {
public:
    // nq: Gauss points per direction of the rule about to be used, 0 if none
    template<Index numDOFs>
    static DeviceProblem& get(const mesh::Mesh& mesh, const topology::TopologicalDOF<numDOFs>& topo, Index nq = 0) {
        Context::get();  // the context is made first, so it outlives the problems
        auto& v = entries();
        for (auto& p : v)
            if (p->matches(mesh, topo)) {
                if (nq) p->ensureQuadrature(nq);
                return *p;
            }
        for (size_t i = 0; i < v.size();)  // a stale entry for these addresses
            if (v[i]->sameAddresses(&mesh, &topo)) v.erase(v.begin() + (long)i);
            else ++i;
        v.push_back(std::make_unique<DeviceProblem>(mesh, topo, nq));
        return *v.back();
    }
    // Gauss points per direction of a tensor rule
    template<typename QuadratureT>
    static Index pointsPerDirection(const QuadratureT& quad, Index dim) {
        const Index tot = quad.numPointsTotal();
        const Index nq = (Index)std::llround(std::pow((double)tot, 1.0 / (double)dim));
        Index chk = 1;
        for (Index i = 0; i < dim; ++i) chk *= nq;
        if (chk != tot) throw std::invalid_argument("Assembler<Hip>: the same Gauss rule in every direction only");
        return nq;
    }
    // drop the device problems of a mesh
    static void release(const mesh::Mesh& mesh) {
        auto& v = entries();
        for (size_t i = 0; i < v.size();)
            if (v[i]->onMesh(&mesh)) v.erase(v.begin() + (long)i);
            else ++i;
    }

private:
    static std::vector<std::unique_ptr<DeviceProblem>>& entries() {
        thread_local std::vector<std::unique_ptr<DeviceProblem>> v;
        return v;
    }
};

// ---------------------------------------------------------------------------
// forms -> device form kinds
// ---------------------------------------------------------------------------
namespace detail {

template<typename F> struct LeafKind : std::integral_constant<int32_t, 0> {};
template<typename QP> struct LeafKind<equation::heateq::DiffusionForm<QP>> : std::integral_constant<int32_t, RSD_FORM_DIFFUSION> {};
template<typename QP> struct LeafKind<equation::heateq::MassForm<QP>> : std::integral_constant<int32_t, RSD_FORM_MASS> {};
template<typename QP> struct LeafKind<equation::heateq::TangentDiffusionForm<QP>> : std::integral_constant<int32_t, RSD_FORM_TDIFFUSION> {};
template<typename QP> struct LeafKind<equation::heateq::TangentMassForm<QP>> : std::integral_constant<int32_t, RSD_FORM_TMASS> {};

template<typename F> struct IsRegistry : std::false_type {};
template<typename... Fs> struct IsRegistry<fem::form::FormRegistry<Fs...>> : std::true_type {};

template<typename F> struct Scaled : std::false_type {};
template<Index N, typename Inner> struct Scaled<fem::form::ScaledForm<N, Inner>> : std::true_type { using inner = Inner; };

template<typename F> struct SourceCallable { using type = void; };
template<typename QP, Index SD, Index NC, typename C>
struct SourceCallable<equation::heateq::SourceForm<QP, equation::heateq::evaluator::SourceFunction<SD, NC, C>>> { using type = C; };

template<typename F> struct IsNodalSource : std::false_type {};
template<typename QP, typename S, Index NC> struct IsNodalSource<equation::heateq::NodalSourceForm<QP, S, NC>> : std::true_type {};

// the leaf forms a form type holds, counting nested registries
template<typename F>
constexpr int leaf_count() {
    if constexpr (IsRegistry<F>::value) return []<typename... Fs>(fem::form::FormRegistry<Fs...>*) { return (0 + ... + leaf_count<Fs>()); }((F*)nullptr);
    else return 1;
}

// a quadrature point on which every heat form's (0, 0) entry is 1: N = 1,
// dN/dx = e_0, K = dK/dT = I, gradT = e_0, rho = cp = dcp/dT = T' = 1,
// measure = w = 1
template<typename QP, typename EvalEleT>
QP unit_point(const EvalEleT& ele) {
    QP qp(ele);
    constexpr Index SD = QP::SpatialDim;
    const Index npe = qp.nodesPerElement();
    for (Index a = 0; a < npe; ++a) {
        qp.N[a] = 1.0;
        for (Index i = 0; i < SD; ++i) qp.dNdx[a * SD + i] = i == 0 ? 1.0 : 0.0;
    }
    for (Index i = 0; i < SD * SD; ++i) qp.K[i] = qp.dKdT[i] = (i % (SD + 1) == 0) ? 1.0 : 0.0;
    for (Index i = 0; i < SD; ++i) qp.gradT[i] = i == 0 ? 1.0 : 0.0;
    qp.rho = qp.cp = qp.dcpdT = 1.0;
    qp.measure = qp.w = 1.0;
    qp.T.valueData[0] = 1.0;
    qp.T.derivatives[0][0] = 1.0;
    return qp;
}

// ScaledForm keeps its coefficient private: its (0, 0) entry on the unit
// point is coefficient x 1 exactly, its inner form being one heat form
template<typename QP, typename EvalEleT, typename F>
Real scaled_coefficient(const F& f, const EvalEleT& ele) {
    QP qp = unit_point<QP>(ele);
    const Index npe = qp.nodesPerElement();
    std::vector<Real> Ke(npe * npe, 0.0);
    f.computeElementLevelMatrix(qp, Ke.data());
    return Ke[0];
}

template<typename QP, bool InScaled = false, typename EvalEleT, typename F>
void collect(const F& f, const EvalEleT& ele, std::vector<int32_t>& kinds, std::vector<double>& coefs, double coef = 1.0) {
    if constexpr (IsRegistry<F>::value) {
        [&]<std::size_t... I>(std::index_sequence<I...>) {
            (collect<QP, InScaled>(f.template get<I>(), ele, kinds, coefs, coef), ...);
        }(std::make_index_sequence<F::numForms()>{});
    } else if constexpr (Scaled<F>::value) {
        using Inner = typename Scaled<F>::inner;
        static_assert(!InScaled, "Assembler<Hip>: a ScaledForm inside a ScaledForm");
        static_assert(leaf_count<Inner>() == 1, "Assembler<Hip>: a ScaledForm of exactly one heat form");
        static_assert(std::is_default_constructible_v<Inner>, "Assembler<Hip>: stateless forms only");
        collect<QP, true>(Inner{}, ele, kinds, coefs, scaled_coefficient<QP>(f, ele));
    } else {
        static_assert(LeafKind<F>::value != 0,
                      "Assembler<Hip>: a form the device does not evaluate (heat diffusion, mass and their tangents)");
        if (kinds.size() >= 4) throw std::invalid_argument("Assembler<Hip>: at most 4 forms per assembly");
        kinds.push_back(LeafKind<F>::value | (InScaled ? RSD_FORM_SCALED : 0));
        coefs.push_back(InScaled ? coef : 1.0);
    }
}

// the single source form of a load registry, if the forms are a load
template<typename F> struct LoadForm { static constexpr bool value = false; };
template<typename S>
requires(!std::is_void_v<typename SourceCallable<S>::type> || IsNodalSource<S>::value)
struct LoadForm<fem::form::FormRegistry<S>> {
    static constexpr bool value = true;
    static const S& get(const fem::form::FormRegistry<S>& r) { return r.template get<0>(); }
};

}  // namespace detail

}  // namespace hip

namespace fem {
namespace assembly {

template<>
class Assembler<linalg::types::backend::Hip> {
    using Hip = linalg::types::backend::Hip;

public:
    // Assembler.tpp:7-76, built on the device problem (host, threaded), the
    // pattern copied into the reference's size_t layout
    template<Index numDOFs>
    static linalg::types::CSRMatrix<Real, Hip> createMatrix(const mesh::Mesh& mesh,
                                                            const topology::TopologicalDOF<numDOFs>& topoDOF) {
        hip::DeviceProblem& dp = hip::ProblemCache::get(mesh, topoDOF);
        int64_t sz[8];
        hip::check(rsd_problem_sizes(dp.get(), sz));
        const Index n = dp.numFree(), nnz = (Index)sz[4];
        std::vector<int64_t> rp(n + 1), ci(nnz);
        hip::check(rsd_problem_get_pattern(dp.get(), rp.data(), ci.data()));
        linalg::types::CSRMatrix<Real, Hip> K(n, n);
        K.resize(nnz);
        static_assert(sizeof(Index) == sizeof(int64_t));
        Hip::copy<Index>(K.rowPtr(), reinterpret_cast<const Index*>(rp.data()), n + 1,
                         linalg::types::backend::CopyKind::HostToDevice);
        Hip::copy<Index>(K.colIdx(), reinterpret_cast<const Index*>(ci.data()), nnz,
                         linalg::types::backend::CopyKind::HostToDevice);
        K.zero();
        hip::detail::remember_pattern(K.rowPtr(), K.colIdx(), dp.get());
        return K;
    }

    template<Index numDOFs>
    static linalg::types::Vector<Real, Hip> createVector(const mesh::Mesh&,
                                                         const topology::TopologicalDOF<numDOFs>& topoDOF) {
        return linalg::types::Vector<Real, Hip>(topoDOF.numFreeDOFs());
    }

    // Assembler.tpp:86-140
    template<Index numDOFs, evaluator::EvalElement EvalEleT, evaluator::EvalQuadraturePointVolume EvalQPT,
             typename ModelT, typename FormsT, typename QuadratureT, typename StateVectorT>
    requires evaluator::EvalModel<ModelT, EvalQPT> && evaluator::StateVector<StateVectorT, numDOFs, EvalQPT>
    static void assembleMatrix(const mesh::Mesh& mesh, const topology::TopologicalDOF<numDOFs>& topoDOF,
                               const Real time, const ModelT& model, const FormsT& forms, const EvalEleT& evalEle,
                               const QuadratureT& quadrature, StateVectorT& state,
                               linalg::types::CSRMatrix<Real, Hip>& K,
                               const fem::boundary::EssentialBoundaryRegistry* bcRegistry) {
        hip::DeviceProblem& dp = prepare<EvalQPT>(mesh, topoDOF, time, model, evalEle, quadrature, bcRegistry);
        std::vector<int32_t> kinds;
        std::vector<double> coefs;
        hip::detail::collect<EvalQPT>(forms, evalEle, kinds, coefs);
        rsd_matrix* M = dp.scratch();
        hip::check(rsd_assemble_matrix(hip::Context::get(), dp.get(), time, (int32_t)kinds.size(), kinds.data(),
                                       coefs.data(), values(state.valueSource), values(state.rateSource), M));
        hip::check(rsd_matrix_store_csr(hip::Context::get(), M, K.data()));
    }

    // Assembler.tpp:142-197
    template<Index numDOFs, evaluator::EvalElement EvalEleT, evaluator::EvalQuadraturePointVolume EvalQPT,
             typename ModelT, typename FormsT, typename QuadratureT, typename StateVectorT>
    requires evaluator::EvalModel<ModelT, EvalQPT> && evaluator::StateVector<StateVectorT, numDOFs, EvalQPT>
    static void assembleDiagonal(const mesh::Mesh& mesh, const topology::TopologicalDOF<numDOFs>& topoDOF,
                                 const Real time, const ModelT& model, const FormsT& forms, const EvalEleT& evalEle,
                                 const QuadratureT& quadrature, StateVectorT& state,
                                 linalg::types::Vector<Real, Hip>& diag,
                                 const fem::boundary::EssentialBoundaryRegistry* bcRegistry) {
        hip::DeviceProblem& dp = prepare<EvalQPT>(mesh, topoDOF, time, model, evalEle, quadrature, bcRegistry);
        std::vector<int32_t> kinds;
        std::vector<double> coefs;
        hip::detail::collect<EvalQPT>(forms, evalEle, kinds, coefs);
        hip::check(rsd_assemble_diagonal(hip::Context::get(), dp.get(), time, (int32_t)kinds.size(), kinds.data(),
                                         coefs.data(), values(state.valueSource), diag.data()));
    }

    // Assembler.tpp:199-255 with one (forms, selector) pair
    template<Index numDOFs, evaluator::EvalElement EvalEleT, evaluator::EvalQuadraturePointVolume EvalQPT,
             typename ModelT, typename QuadratureT, typename StateVectorT, typename... FormsWithOperandT>
    requires evaluator::EvalModel<ModelT, EvalQPT> && evaluator::StateVector<StateVectorT, numDOFs, EvalQPT>
    static void assembleVector(const mesh::Mesh& mesh, const topology::TopologicalDOF<numDOFs>& topoDOF,
                               const Real time, const ModelT& model, const EvalEleT& evalEle,
                               const QuadratureT& quadrature, StateVectorT& state,
                               linalg::types::Vector<Real, Hip>& F,
                               const fem::boundary::EssentialBoundaryRegistry* bcRegistry,
                               FormsWithOperandT... pairs) {
        static_assert(sizeof...(FormsWithOperandT) == 1, "Assembler<Hip>::assembleVector takes one (forms, selector) pair");
        (assembleOne<EvalQPT>(mesh, topoDOF, time, model, evalEle, quadrature, state, F, bcRegistry, pairs), ...);
    }

private:
    template<typename EvalQPT, Index numDOFs, typename ModelT, typename EvalEleT, typename QuadratureT>
    static hip::DeviceProblem& prepare(const mesh::Mesh& mesh, const topology::TopologicalDOF<numDOFs>& topoDOF,
                                       Real time, const ModelT& model, const EvalEleT& evalEle,
                                       const QuadratureT& quadrature,
                                       const fem::boundary::EssentialBoundaryRegistry* bcRegistry) {
        hip::DeviceProblem& dp =
            hip::ProblemCache::get(mesh, topoDOF, hip::ProblemCache::pointsPerDirection(quadrature, EvalEleT::ParametricDim));
        hip::DeviceModel m = hip::probe_model<EvalQPT>(model, evalEle);
        if (m.cp_dependent) throw std::invalid_argument("Assembler<Hip>: a temperature-dependent specific heat is not on the device");
        dp.setModel(m);
        dp.setDirichlet(time, topoDOF, bcRegistry);
        return dp;
    }

    template<typename VectorT>
    static const double* values(const VectorT* v) {
        return v ? v->data() : nullptr;
    }

    template<typename EvalQPT, Index numDOFs, typename ModelT, typename EvalEleT, typename QuadratureT,
             typename StateVectorT, typename PairT>
    static void assembleOne(const mesh::Mesh& mesh, const topology::TopologicalDOF<numDOFs>& topoDOF, Real time,
                            const ModelT& model, const EvalEleT& evalEle, const QuadratureT& quadrature,
                            StateVectorT& state, linalg::types::Vector<Real, linalg::types::backend::Hip>& F,
                            const fem::boundary::EssentialBoundaryRegistry* bcRegistry, PairT& pair) {
        using FormsT = std::decay_t<decltype(pair.forms)>;
        hip::DeviceProblem& dp = prepare<EvalQPT>(mesh, topoDOF, time, model, evalEle, quadrature, bcRegistry);
        rsd_ctx* ctx = hip::Context::get();
        if constexpr (hip::detail::LoadForm<FormsT>::value) {
            const auto& src = hip::detail::LoadForm<FormsT>::get(pair.forms);
            using S = std::decay_t<decltype(src)>;
            if constexpr (hip::detail::IsNodalSource<S>::value) {
                // NodalSourceForm keeps its nodal data private: each node's value
                // is its interpolant on the unit point with the node gathered at
                // every element slot (N = e_0 there: the value times 1)
                EvalQPT qp = hip::detail::unit_point<EvalQPT>(evalEle);
                const Index npe = qp.nodesPerElement();
                for (Index a = 0; a < npe; ++a) qp.N[a] = a == 0 ? 1.0 : 0.0;
                std::vector<double> nodal(mesh.data.numNodes);
                std::vector<Index> ids(npe);
                std::vector<Real> fe(npe);
                for (Index n = 0; n < mesh.data.numNodes; ++n) {
                    std::fill(ids.begin(), ids.end(), n);
                    std::fill(fe.begin(), fe.end(), 0.0);
                    src.gatherElementData(ids.data(), npe);
                    src.computeElementLevelVector(qp, nullptr, fe.data());
                    nodal[n] = fe[0];
                }
                linalg::types::Vector<Real, linalg::types::backend::Hip> dn(mesh.data.numNodes);
                linalg::types::backend::Hip::copy<Real>(dn.data(), nodal.data(), nodal.size(),
                                                        linalg::types::backend::CopyKind::HostToDevice);
                hip::check(rsd_assemble_load(ctx, dp.get(), time, nullptr, dn.data(), F.data()));
            } else {
                using C = typename hip::detail::SourceCallable<S>::type;
                static_assert(std::is_same_v<C, hip::Expression<EvalQPT::SpatialDim, 1>>,
                              "Assembler<Hip>: a source function the device evaluates is a hip::Expression");
                hip::check(rsd_assemble_load(ctx, dp.get(), time, src.sourceFunction.f.text().c_str(), nullptr,
                                             F.data()));
            }
        } else {
            std::vector<int32_t> kinds;
            std::vector<double> coefs;
            hip::detail::collect<EvalQPT>(pair.forms, evalEle, kinds, coefs);
            // which of the state's buffers the selector reads (HeatStateVector.hpp:28-30)
            const Real* sel = pair.selector(state);
            int32_t which = RSD_SEL_VALUE;
            if (sel == state.rateBuffer()) which = RSD_SEL_RATE;
            else if (sel == state.operandBuffer()) which = RSD_SEL_OPERAND;
            else if (sel != state.valueBuffer())
                throw std::invalid_argument("Assembler<Hip>: the selector reads none of the state's buffers");
            hip::check(rsd_assemble_vector(ctx, dp.get(), time, (int32_t)kinds.size(), kinds.data(), coefs.data(),
                                           values(state.valueSource), values(state.rateSource),
                                           values(state.operandSource), which, F.data()));
        }
    }
};

}  // namespace assembly
}  // namespace fem
}  // namespace residuum
