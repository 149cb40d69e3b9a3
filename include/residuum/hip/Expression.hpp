// residuum/hip/Expression.hpp -- functions of (t, x) and of T that run on
// the host and on the device.
//
// The reference evaluates its configured functions with exprtk on the host
// (src/utils/expression/ScalarExpression.cpp:3-30,
// equation/heateq/evaluator/TemperatureExpression.hpp:38-51).  The device
// evaluates the same kind of expression text in libresiduum_hip.so's own
// grammar (variables t x y z T; + - * / ^; sin cos tan exp log sqrt abs pow
// min max; pi e).  These types carry the text, so the Hip assembler can hand
// it to the device, and evaluate it on the host through the library's
// interpreter, so the reference's CPU templates can call them as the
// callables its function wrappers expect:
//   Expression<SD, N>     (time, x[SD], out[N]) -> void: a callable for
//                         SourceFunction, BoundaryValueFunction and
//                         BoundaryFluxFunction (equation/heateq/evaluator/
//                         SourceFunction.hpp:12-32, boundary/*.hpp);
//   TemperatureExpression (T) -> Real, as the reference's class of that name.
#pragma once

#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>

#include "core/Types.hpp"
#include "residuum_hip.h"

namespace residuum {
namespace hip {

namespace detail {

struct ExprDeleter {
    void operator()(rsd_expr* e) const noexcept { rsd_expr_destroy(e); }
};

// one compiled expression, shared by copies
class CompiledExpr {
public:
    CompiledExpr() = default;
    explicit CompiledExpr(std::string text) : text_(std::move(text)) {
        rsd_expr* e = nullptr;
        if (rsd_expr_compile(text_.c_str(), &e) != RSD_OK)
            throw std::runtime_error("Failed to compile expression:\n" + text_ + "\n" + rsd_last_error_string());
        e_.reset(e, ExprDeleter{});
        int32_t ut = 0, uT = 0;
        rsd_expr_uses(e, &ut, &uT);
        uses_t_ = ut != 0;
        uses_T_ = uT != 0;
    }
    Real operator()(Real t, Real x, Real y, Real z, Real T) const {
        double v = 0.0;
        rsd_expr_eval(e_.get(), t, x, y, z, T, &v);
        return v;
    }
    const std::string& text() const { return text_; }
    bool usesTime() const { return uses_t_; }
    bool usesTemperature() const { return uses_T_; }
    explicit operator bool() const { return static_cast<bool>(e_); }

private:
    std::string text_;
    std::shared_ptr<rsd_expr> e_;
    bool uses_t_ = false, uses_T_ = false;
};

}  // namespace detail

template<Index SD, Index N = 1>
class Expression {
public:
    static constexpr Index SpatialDim = SD;
    static constexpr Index NumComponents = N;

    Expression() = default;
    // one text per component
    explicit Expression(const std::array<std::string, N>& texts) {
        for (Index i = 0; i < N; ++i) c_[i] = detail::CompiledExpr(texts[i]);
    }
    explicit Expression(const std::string& text) requires(N == 1) { c_[0] = detail::CompiledExpr(text); }
    template<typename... S>
    requires(sizeof...(S) == N && N > 1) explicit Expression(const S&... texts)
        : Expression(std::array<std::string, N>{std::string(texts)...}) {}

    void operator()(Real t, const Real* x, Real* out) const {
        const Real X = x[0], Y = SD > 1 ? x[1] : Real(0), Z = SD > 2 ? x[2] : Real(0);
        for (Index i = 0; i < N; ++i) out[i] = c_[i](t, X, Y, Z, Real(0));
    }

    const std::string& text(Index i = 0) const { return c_[i].text(); }
    bool usesTime() const {
        for (Index i = 0; i < N; ++i)
            if (c_[i].usesTime()) return true;
        return false;
    }

private:
    std::array<detail::CompiledExpr, N> c_;
};

class TemperatureExpression {
public:
    TemperatureExpression() = default;
    explicit TemperatureExpression(const std::string& text) : c_(text) {}
    Real operator()(Real T) const { return c_(Real(0), Real(0), Real(0), Real(0), T); }
    const std::string& text() const { return c_.text(); }

private:
    detail::CompiledExpr c_;
};

}  // namespace hip
}  // namespace residuum
