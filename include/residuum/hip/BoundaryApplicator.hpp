// residuum/hip/BoundaryApplicator.hpp -- fem::boundary::BoundaryApplicator<Hip>.
//
// The reference's boundary seam (fem/boundary/BoundaryApplicator.hpp:26-37,
// CPU specialisation in backend/cpu/BoundaryApplicator.tpp) with the same
// static members on Hip vectors:
//   applyEssentialBCs   the lift F -= A_fc g (BoundaryApplicator.tpp:7-64):
//                       the registry's boundary values at `time`, then the
//                       device's lift over the elements touching them;
//   applyNaturalBCs     the flux terms (BoundaryApplicator.tpp:66-140) on
//                       the device (rsd_apply_natural).
//
// A NaturalBoundaryRegistry holds its entries type-erased, so the device
// cannot see their flux functions; hip::NaturalBoundaryRegistry is that
// registry plus a device description of every entry it registers (a
// hip::Expression flux or nodal flux vectors).  applyNaturalBCs takes the
// reference's registry type and finds the description of a
// hip::NaturalBoundaryRegistry; a plain registry with entries on the mesh's
// tags is refused, one without any is a no-op, as on the CPU.
#pragma once

#include <memory>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

// the Hip overloads of operations::* before any reference template that calls them
#include "residuum/hip/Solvers.hpp"

#include "equation/heateq/boundary/BoundaryFluxFunction.hpp"
#include "equation/heateq/evaluator/DefaultModel.hpp"
#include "equation/heateq/form/FluxBoundaryForm.hpp"
#include "equation/heateq/form/NodalFluxForm.hpp"
#include "fem/boundary/BoundaryApplicator.hpp"
#include "fem/boundary/NaturalBoundaryRegistry.hpp"

#include "residuum/hip/Assembler.hpp"

namespace residuum {
namespace hip {

// nodal flux vectors [n_nodes * SD] as the EvalNodalData a NodalFluxForm reads
// (fem/evaluator/EvalNodalData.hpp:15-20)
template<Index SD>
struct NodalVectors {
    static constexpr Index NumComponents = SD;
    std::shared_ptr<const std::vector<Real>> data;
    void eval(Index node, Real* out) const {
        for (Index i = 0; i < SD; ++i) out[i] = (*data)[node * SD + i];
    }
};

template<typename EvalQPT>
class NaturalBoundaryRegistry : public fem::boundary::NaturalBoundaryRegistry<EvalQPT> {
public:
    using Base = fem::boundary::NaturalBoundaryRegistry<EvalQPT>;
    static constexpr Index SD = EvalQPT::SpatialDim;
    using FluxFunction = equation::heateq::BoundaryFluxFunction<SD, 1, Expression<SD, SD>>;
    using FluxForms = fem::form::FormRegistry<equation::heateq::FluxBoundaryForm<EvalQPT, FluxFunction>>;
    using NodalFluxForms = fem::form::FormRegistry<equation::heateq::NodalFluxForm<EvalQPT, NodalVectors<SD>, 1, SD>>;

    struct Entry {
        Int tag;
        std::vector<std::string> exprs;                    // SD texts, or empty
        std::shared_ptr<const std::vector<Real>> nodal;    // or nodal vectors
    };

    NaturalBoundaryRegistry() {
        std::lock_guard<std::mutex> lk(mutex());
        directory()[static_cast<const Base*>(this)] = this;
    }
    NaturalBoundaryRegistry(const NaturalBoundaryRegistry&) = delete;
    NaturalBoundaryRegistry& operator=(const NaturalBoundaryRegistry&) = delete;
    ~NaturalBoundaryRegistry() {
        std::lock_guard<std::mutex> lk(mutex());
        directory().erase(static_cast<const Base*>(this));
    }

    // FluxBoundaryForm with the flux vector q(t, x) (HeatProblem.tpp:96-110)
    void registerFlux(Int tag, const Expression<SD, SD>& q) {
        auto bc = std::make_shared<fem::boundary::BoundaryCondition<FluxFunction>>(
            fem::boundary::BoundaryCondition<FluxFunction>{tag, {fem::boundary::BCCategory::Natural}, FluxFunction{q}});
        flux_.push_back(std::make_unique<FluxForms>(equation::heateq::FluxBoundaryForm<EvalQPT, FluxFunction>{q}));
        this->template registerBC<FluxFunction>(bc, *flux_.back(), model_);
        Entry e{tag, {}, nullptr};
        for (Index i = 0; i < SD; ++i) e.exprs.push_back(q.text(i));
        entries_.push_back(std::move(e));
    }

    // NodalFluxForm with nodal flux vectors (HeatProblem.tpp:86-94)
    void registerNodalFlux(Int tag, std::vector<Real> nodal) {
        auto data = std::make_shared<const std::vector<Real>>(std::move(nodal));
        auto bc = std::make_shared<fem::boundary::BoundaryCondition<FluxFunction>>(
            fem::boundary::BoundaryCondition<FluxFunction>{tag, {fem::boundary::BCCategory::Natural}, FluxFunction{}});
        NodalVectors<SD> nv{data};
        nodal_.push_back(std::make_unique<NodalFluxForms>(equation::heateq::NodalFluxForm<EvalQPT, NodalVectors<SD>, 1, SD>{nv}));
        this->template registerBC<FluxFunction>(bc, *nodal_.back(), model_);
        entries_.push_back(Entry{tag, {}, data});
    }

    const std::vector<Entry>& deviceEntries() const { return entries_; }

    // the device copy of entry k's nodal vectors
    const Real* deviceNodal(size_t k) const {
        if (dev_.size() < entries_.size()) dev_.resize(entries_.size());
        if (!entries_[k].nodal) return nullptr;
        if (!dev_[k]) {
            const auto& h = *entries_[k].nodal;
            dev_[k] = std::make_unique<HipVector>(h.size());
            linalg::types::backend::Hip::copy<Real>(dev_[k]->data(), h.data(), h.size(),
                                                    linalg::types::backend::CopyKind::HostToDevice);
        }
        return dev_[k]->data();
    }

    static const NaturalBoundaryRegistry* find(const Base* b) {
        std::lock_guard<std::mutex> lk(mutex());
        auto& d = directory();
        auto it = d.find(b);
        return it == d.end() ? nullptr : it->second;
    }

private:
    // every live registry of this type, by its base address (guarded by mutex())
    static std::unordered_map<const Base*, const NaturalBoundaryRegistry*>& directory() {
        static std::unordered_map<const Base*, const NaturalBoundaryRegistry*> d;
        return d;
    }
    static std::mutex& mutex() {
        static std::mutex m;
        return m;
    }
    equation::heateq::evaluator::DefaultModel<EvalQPT> model_;
    std::vector<std::unique_ptr<FluxForms>> flux_;
    std::vector<std::unique_ptr<NodalFluxForms>> nodal_;
    std::vector<Entry> entries_;
    mutable std::vector<std::unique_ptr<HipVector>> dev_;
};

}  // namespace hip

namespace fem {
namespace boundary {

template<>
class BoundaryApplicator<linalg::types::backend::Hip> {
    using Hip = linalg::types::backend::Hip;

public:
    // BoundaryApplicator.tpp:7-64
    template<Index numDOFs, evaluator::EvalElement EvalEleT, evaluator::EvalQuadraturePointVolume EvalQPT,
             typename ModelT, typename FormsT, typename QuadratureT>
    requires evaluator::EvalModel<ModelT, EvalQPT>
    static void applyEssentialBCs(const mesh::Mesh& mesh, const topology::TopologicalDOF<numDOFs>& topoDOF,
                                  const EssentialBoundaryRegistry& bcRegistry, const Real time, const ModelT& model,
                                  const FormsT& forms, const EvalEleT& evalEle, const QuadratureT& quadrature,
                                  linalg::types::Vector<Real, Hip>& F) {
        hip::DeviceProblem& dp = hip::ProblemCache::get(
            mesh, topoDOF, hip::ProblemCache::pointsPerDirection(quadrature, EvalEleT::ParametricDim));
        hip::DeviceModel m = hip::probe_model<EvalQPT>(model, evalEle);
        dp.setModel(m);
        dp.setDirichlet(time, topoDOF, &bcRegistry);
        std::vector<int32_t> kinds;
        std::vector<double> coefs;
        hip::detail::collect<EvalQPT>(forms, evalEle, kinds, coefs);
        hip::check(rsd_apply_essential(hip::Context::get(), dp.get(), time, (int32_t)kinds.size(), kinds.data(),
                                       coefs.data(), F.data()));
    }

    // BoundaryApplicator.tpp:66-140
    template<Index numDOFs, evaluator::EvalElement EvalEleT, typename EvalQPT, typename QuadratureT>
    static void applyNaturalBCs(const mesh::Mesh& mesh, const topology::TopologicalDOF<numDOFs>& topoDOF,
                                const NaturalBoundaryRegistry<EvalQPT>& bcRegistry, const Real time,
                                const EvalEleT& evalEle, const QuadratureT& quadrature,
                                linalg::types::Vector<Real, Hip>& F) {
        (void)evalEle;
        const auto* reg = hip::NaturalBoundaryRegistry<EvalQPT>::find(&bcRegistry);
        if (!reg) {
            std::set<Int> tags(mesh.data.rng.begin(), mesh.data.rng.end());
            for (Int t : tags)
                if (t >= 0 && bcRegistry.hasAny(t))
                    throw std::invalid_argument("BoundaryApplicator<Hip>: natural BCs on the device need a "
                                                "residuum::hip::NaturalBoundaryRegistry (device-evaluable fluxes)");
            return;
        }
        const auto& ents = reg->deviceEntries();
        if (ents.empty()) return;
        if (ents.size() > 8) throw std::invalid_argument("BoundaryApplicator<Hip>: at most 8 natural BC entries");
        hip::DeviceProblem& dp = hip::ProblemCache::get(
            mesh, topoDOF, hip::ProblemCache::pointsPerDirection(quadrature, EvalEleT::ParametricDim - 1));
        constexpr Index SD = EvalQPT::SpatialDim;
        std::vector<int32_t> tags;
        std::vector<const char*> ex(ents.size() * SD, nullptr);
        std::vector<const double*> nod(ents.size(), nullptr);
        for (size_t k = 0; k < ents.size(); ++k) {
            tags.push_back((int32_t)ents[k].tag);
            if (ents[k].nodal) nod[k] = reg->deviceNodal(k);
            else
                for (Index i = 0; i < SD; ++i) ex[k * SD + i] = ents[k].exprs[i].c_str();
        }
        hip::check(rsd_apply_natural(hip::Context::get(), dp.get(), time, (int32_t)ents.size(), tags.data(), ex.data(),
                                     nod.data(), F.data()));
    }
};

}  // namespace boundary
}  // namespace fem
}  // namespace residuum
