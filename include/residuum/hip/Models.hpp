// residuum/hip/Models.hpp -- material models the device can evaluate.
//
// hip::ConductivityModel and hip::SpecificHeatModel have the interface and
// the arithmetic of the reference's ConductivityModel
// (equation/heateq/evaluator/ConductivityModel.hpp:13-96) and
// SpecificHeatModel (SpecificHeatModel.hpp:12-61): eval / evalGradient set
// qp.K = k(T) shape, qp.dKdT = k'(T) shape, qp.cp, qp.dcpdT on the host, with
// k(T), k'(T) as hip::TemperatureExpression instead of exprtk.  On a probe
// point of the Hip assembler (a quadrature point carrying a DeviceModel,
// see probe_model) they also write their device description, which is how
// a temperature-dependent model reaches the device: any other model is read
// by its values at probe points and must then not depend on T.
#pragma once

#include <array>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "core/Types.hpp"
#include "residuum/hip/Expression.hpp"

namespace residuum {
namespace hip {

// what the device element kernels take as their material model
// (rsd_problem_set_conductivity / rsd_problem_set_mass)
struct DeviceModel {
    bool conductivity = false;  // a conductivity was found
    double k_scalar = 1.0;
    bool anisotropic = false;
    std::array<double, 9> tensor{1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::string k_expr, dk_expr;  // empty: k constant (k_scalar), k' = 0
    bool mass = false;            // rho and cp were found
    double rho = 1.0, cp = 1.0;
    bool cp_dependent = false;    // cp(T): not on the device

    bool operator==(const DeviceModel&) const = default;
};

template<typename QuadraturePointT>
struct ConductivityModel {
    enum class Symmetry { Isotropic, Anisotropic };
    enum class Dependence { Constant, TemperatureDependent };

    static constexpr Index SpatialDim = QuadraturePointT::SpatialDim;

    Symmetry symmetry = Symmetry::Isotropic;
    Dependence dependence = Dependence::Constant;
    Real scalar = Real(0);
    std::array<Real, SpatialDim * SpatialDim> tensor{};
    TemperatureExpression valueExpr, gradientExpr;

    void setConstant(Real k) {
        symmetry = Symmetry::Isotropic;
        dependence = Dependence::Constant;
        scalar = k;
    }
    void setAnisotropic(const std::vector<std::vector<Real>>& t) {
        setTensor(t);
        symmetry = Symmetry::Anisotropic;
        dependence = Dependence::Constant;
        scalar = Real(1);
    }
    void setTemperatureDependentIsotropic(const std::string& value, const std::string& gradient) {
        symmetry = Symmetry::Isotropic;
        dependence = Dependence::TemperatureDependent;
        valueExpr = TemperatureExpression(value);
        gradientExpr = TemperatureExpression(gradient);
    }
    void setTemperatureDependentAnisotropic(const std::vector<std::vector<Real>>& t, const std::string& value,
                                            const std::string& gradient) {
        setTensor(t);
        symmetry = Symmetry::Anisotropic;
        dependence = Dependence::TemperatureDependent;
        valueExpr = TemperatureExpression(value);
        gradientExpr = TemperatureExpression(gradient);
    }

    template<typename QP>
    void eval(QP& qp) const {
        if constexpr (requires { qp.hipDeviceModel; }) describe(qp.hipDeviceModel);
        Real k = scalar;
        if (dependence == Dependence::TemperatureDependent) {
            if constexpr (requires { qp.T.value(); }) k = valueExpr(qp.T.value());
        }
        for (Index i = 0; i < SpatialDim; ++i)
            for (Index j = 0; j < SpatialDim; ++j) qp.K[i * SpatialDim + j] = k * shape(i, j);
    }

    template<typename QP>
    void evalGradient(QP& qp) const {
        Real dk = Real(0);
        if (dependence == Dependence::TemperatureDependent) {
            if constexpr (requires { qp.T.value(); }) dk = gradientExpr(qp.T.value());
        }
        for (Index i = 0; i < SpatialDim; ++i)
            for (Index j = 0; j < SpatialDim; ++j) qp.dKdT[i * SpatialDim + j] = dk * shape(i, j);
    }

    void describe(DeviceModel& d) const {
        d.conductivity = true;
        d.k_scalar = scalar;
        d.anisotropic = symmetry == Symmetry::Anisotropic;
        d.tensor = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        if (d.anisotropic)
            for (Index i = 0; i < SpatialDim * SpatialDim; ++i) d.tensor[i] = tensor[i];
        d.k_expr = dependence == Dependence::TemperatureDependent ? valueExpr.text() : std::string();
        d.dk_expr = dependence == Dependence::TemperatureDependent ? gradientExpr.text() : std::string();
    }

private:
    Real shape(Index i, Index j) const {
        return symmetry == Symmetry::Isotropic ? ((i == j) ? Real(1) : Real(0)) : tensor[i * SpatialDim + j];
    }
    void setTensor(const std::vector<std::vector<Real>>& t) {
        if (t.size() != SpatialDim) throw std::runtime_error("ConductivityModel: tensor must be SpatialDim x SpatialDim");
        for (Index i = 0; i < SpatialDim; ++i) {
            if (t[i].size() != SpatialDim)
                throw std::runtime_error("ConductivityModel: tensor must be SpatialDim x SpatialDim");
            for (Index j = 0; j < SpatialDim; ++j) tensor[i * SpatialDim + j] = t[i][j];
        }
    }
};

template<typename QuadraturePointT>
struct SpecificHeatModel {
    enum class Dependence { Constant, TemperatureDependent };

    Dependence dependence = Dependence::Constant;
    Real value = Real(0);
    TemperatureExpression valueExpr, gradientExpr;

    void setConstant(Real cp) {
        dependence = Dependence::Constant;
        value = cp;
    }
    void setTemperatureDependent(const std::string& v, const std::string& g) {
        dependence = Dependence::TemperatureDependent;
        valueExpr = TemperatureExpression(v);
        gradientExpr = TemperatureExpression(g);
    }

    template<typename QP>
    void eval(QP& qp) const {
        if constexpr (requires { qp.hipDeviceModel; })
            qp.hipDeviceModel.cp_dependent = dependence == Dependence::TemperatureDependent;
        Real cpVal = value;
        if (dependence == Dependence::TemperatureDependent) {
            if constexpr (requires { qp.T.value(); }) cpVal = valueExpr(qp.T.value());
        }
        qp.cp = cpVal;
    }

    template<typename QP>
    void evalGradient(QP& qp) const {
        Real dcp = Real(0);
        if (dependence == Dependence::TemperatureDependent) {
            if constexpr (requires { qp.T.value(); }) dcp = gradientExpr(qp.T.value());
        }
        qp.dcpdT = dcp;
    }
};

// The device description of any model: evaluated at a probe point (a
// quadrature point of the evaluator type QP with a DeviceModel slot) at two
// temperatures.  Models of this header describe themselves; for others the
// values they set are taken as constants (K = k_scalar shape or a full
// tensor, rho, cp), and a model whose values change with T is refused.
template<typename QP, typename EvalEleT, typename ModelT>
DeviceModel probe_model(const ModelT& model, const EvalEleT& ele) {
    struct Probe : QP {
        DeviceModel hipDeviceModel;
        explicit Probe(const EvalEleT& e) : QP(e) {}
    };
    constexpr Index SD = QP::SpatialDim;
    const Real nan = std::numeric_limits<Real>::quiet_NaN();
    auto run = [&](Real T, Probe& p) {
        for (Index i = 0; i < SD * SD; ++i) p.K[i] = p.dKdT[i] = nan;
        p.rho = p.cp = p.dcpdT = nan;
        if constexpr (requires { p.T.valueData[0]; }) p.T.valueData[0] = T;
        model.eval(p);
        model.evalGradient(p);
    };
    Probe a(ele), b(ele);
    run(Real(0.37), a);
    run(Real(1.91), b);
    DeviceModel d = a.hipDeviceModel;
    if (!d.conductivity && !std::isnan(a.K[0])) {
        for (Index i = 0; i < SD * SD; ++i)
            if (a.K[i] != b.K[i] || (a.dKdT[i] != Real(0) && !std::isnan(a.dKdT[i])))
                throw std::invalid_argument(
                    "Assembler<Hip>: a temperature-dependent conductivity needs a device description "
                    "(residuum::hip::ConductivityModel)");
        bool iso = true;
        for (Index i = 0; i < SD; ++i)
            for (Index j = 0; j < SD; ++j) iso = iso && (i == j ? a.K[i * SD + j] == a.K[0] : a.K[i * SD + j] == 0.0);
        d.conductivity = true;
        d.anisotropic = !iso;
        d.k_scalar = iso ? a.K[0] : Real(1);
        d.tensor = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        if (!iso)
            for (Index i = 0; i < SD * SD; ++i) d.tensor[i] = a.K[i];
    }
    if (!std::isnan(a.rho) && !std::isnan(a.cp)) {
        if (!d.cp_dependent && a.cp != b.cp)
            throw std::invalid_argument("Assembler<Hip>: a temperature-dependent specific heat is not on the device");
        d.mass = true;
        d.rho = a.rho;
        d.cp = a.cp;
    }
    return d;
}

}  // namespace hip
}  // namespace residuum
