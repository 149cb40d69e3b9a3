// The reference's own linear-algebra templates on the Hip backend tag,
// checked against the same templates on the CPU tag.
//
// Built against the reference's headers (the Hip specialisations plug into
// them); each case prints "PASS <name>" or "FAIL <name>: why", and the exit
// status is the number of failures.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "residuum/hip/Solvers.hpp"

#include "linalg/types/CSRMatrix.hpp"
#include "linalg/types/Vector.hpp"

using namespace residuum;
using CPU = linalg::types::backend::CPU;
using Hip = linalg::types::backend::Hip;
using VecC = linalg::types::Vector<double, CPU>;
using VecH = linalg::types::Vector<double, Hip>;
using MatC = linalg::types::CSRMatrix<double, CPU>;
using MatH = linalg::types::CSRMatrix<double, Hip>;
namespace ops = linalg::operations;
namespace pc = linalg::solver::preconditioner;

static int failures = 0;
static void report(const std::string& name, bool ok, const std::string& why = "") {
    std::printf("%s %s%s%s\n", ok ? "PASS" : "FAIL", name.c_str(), ok ? "" : ": ", ok ? "" : why.c_str());
    std::fflush(stdout);
    if (!ok) ++failures;
}

// a logger with the interface the solvers call, recording what they log
struct RecLogger {
    std::vector<std::pair<Index, double>> logs;
    int resets = 0, summaries = 0;
    bool lastConverged = false;
    void reset() { logs.clear(); ++resets; }
    template<typename D>
    std::vector<D> computePerDOFNorms(const D* r, Index n) const {
        D s = D(0);
        for (Index i = 0; i < n; ++i) s += r[i] * r[i];
        return {std::sqrt(s)};
    }
    template<typename D>
    void log(Index k, const std::vector<D>& v, D = D(0)) { logs.push_back({k, (double)v[0]}); }
    void summary(bool c) { ++summaries; lastConverged = c; }
};

// 2D 5-point operator with a variable coefficient and a diagonal shift:
// symmetric positive definite, sorted columns, n = m * m rows
struct HostCSR {
    Index n = 0;
    std::vector<Index> rp, ci;
    std::vector<double> val;
};
static HostCSR laplace2d(int m, double shift, bool nonsym = false) {
    HostCSR A;
    A.n = (Index)m * m;
    A.rp.push_back(0);
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
            const Index r = (Index)j * m + i;
            auto k = [&](int a, int b) { return 1.0 + 0.5 * std::sin(0.1 * (a + 3 * b)) + (nonsym ? 0.2 * (a % 3) : 0.0); };
            double diag = shift;
            std::vector<std::pair<Index, double>> row;
            if (j > 0) { row.push_back({r - m, -k(i, j)}); diag += k(i, j); }
            if (i > 0) { row.push_back({r - 1, -k(i, j + 7)}); diag += k(i, j + 7); }
            row.push_back({r, 0.0});
            if (i < m - 1) { row.push_back({r + 1, -k(i + 1, j + 7)}); diag += k(i + 1, j + 7); }
            if (j < m - 1) { row.push_back({r + m, -k(i, j + 1)}); diag += k(i, j + 1); }
            for (auto& e : row) {
                if (e.first == r) e.second = diag;
                A.ci.push_back(e.first);
                A.val.push_back(e.second);
            }
            A.rp.push_back(A.ci.size());
        }
    return A;
}

static void load(MatC& A, const HostCSR& h) {
    A.resize(h.ci.size());
    std::copy(h.rp.begin(), h.rp.end(), A.rowPtr());
    std::copy(h.ci.begin(), h.ci.end(), A.colIdx());
    std::copy(h.val.begin(), h.val.end(), A.data());
}
static void load(MatH& A, const HostCSR& h) {
    A.resize(h.ci.size());
    Hip::copy<Index>(A.rowPtr(), h.rp.data(), h.rp.size(), linalg::types::backend::CopyKind::HostToDevice);
    Hip::copy<Index>(A.colIdx(), h.ci.data(), h.ci.size(), linalg::types::backend::CopyKind::HostToDevice);
    Hip::copy<double>(A.data(), h.val.data(), h.val.size(), linalg::types::backend::CopyKind::HostToDevice);
}
static std::vector<double> down(const VecH& v) {
    std::vector<double> h(v.size());
    Hip::copy<double>(h.data(), v.data(), v.size(), linalg::types::backend::CopyKind::DeviceToHost);
    return h;
}
static void up(VecH& v, const std::vector<double>& h) {
    Hip::copy<double>(v.data(), h.data(), h.size(), linalg::types::backend::CopyKind::HostToDevice);
}
static std::vector<double> host(const VecC& v) { return std::vector<double>(v.data(), v.data() + v.size()); }
static double max_rel(const std::vector<double>& a, const std::vector<double>& b) {
    double m = 0.0, s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) {
        m = std::max(m, std::fabs(a[i] - b[i]));
        s = std::max(s, std::fabs(b[i]));
    }
    return m / (s > 0 ? s : 1.0);
}

int main() {
    const int m = 160;
    const HostCSR hA = laplace2d(m, 0.05);
    const Index n = hA.n;
    std::vector<double> hx(n), hy(n), hb(n);
    for (Index i = 0; i < n; ++i) {
        hx[i] = std::sin(0.37 * i) + 0.1;
        hy[i] = std::cos(0.11 * i) - 0.2;
        hb[i] = 1.0 + 0.5 * std::sin(0.013 * i);
    }
    VecC cx(n), cy(n), cz(n);
    std::copy(hx.begin(), hx.end(), cx.data());
    std::copy(hy.begin(), hy.end(), cy.data());
    VecH gx(n), gy(n), gz(n);
    up(gx, hx);
    up(gy, hy);

    // Backend tag: set / zero through the reference's Vector
    {
        VecH t(n);
        t.set(2.5);
        auto h = down(t);
        bool ok = true;
        for (double v : h) ok = ok && v == 2.5;
        t.zero();
        h = down(t);
        for (double v : h) ok = ok && v == 0.0;
        report("backend_set_zero", ok);
    }
    // BLAS-1
    {
        const double dc = ops::dot(cx, cy), dg = ops::dot(gx, gy);
        report("dot", std::fabs(dc - dg) <= 1e-13 * std::fabs(dc) + 1e-13, std::to_string(dc) + " vs " + std::to_string(dg));
        const double nc = ops::norm(cx), ng = ops::norm(gx);
        report("norm", std::fabs(nc - ng) <= 1e-14 * nc);
        ops::axpy(0.3, cx, cy);
        ops::axpy(0.3, gx, gy);
        report("axpy_bitwise", down(gy) == host(cy));
        ops::axpby(0.7, -1.3, cx, cy);
        ops::axpby(0.7, -1.3, gx, gy);
        report("axpby_bitwise", down(gy) == host(cy));
        ops::scal(1.7, cy);
        ops::scal(1.7, gy);
        report("scal_bitwise", down(gy) == host(cy));
        ops::multiply(cx, cy, cz);
        ops::multiply(gx, gy, gz);
        report("multiply_bitwise", down(gz) == host(cz));
        ops::copy(cz, cy);
        ops::copy(gz, gy);
        report("copy_bitwise", down(gy) == host(cy));
    }
    // BLAS-1 on float vectors: the reference's float templates on the CPU
    // tag against the float overloads on the Hip tag
    {
        using VecCf = linalg::types::Vector<float, CPU>;
        using VecHf = linalg::types::Vector<float, Hip>;
        VecCf fx(n), fy(n), fz(n);
        VecHf hxf(n), hyf(n), hzf(n);
        std::vector<float> tx(n), ty(n);
        for (Index i = 0; i < n; ++i) {
            tx[i] = (float)(0.5 + 0.25 * std::sin(0.37 * i));
            ty[i] = (float)(0.75 + 0.25 * std::cos(0.11 * i));
        }
        std::copy(tx.begin(), tx.end(), fx.data());
        std::copy(ty.begin(), ty.end(), fy.data());
        Hip::copy<float>(hxf.data(), tx.data(), n, linalg::types::backend::CopyKind::HostToDevice);
        Hip::copy<float>(hyf.data(), ty.data(), n, linalg::types::backend::CopyKind::HostToDevice);
        auto downf = [&](const VecHf& v) {
            std::vector<float> h(v.size());
            Hip::copy<float>(h.data(), v.data(), v.size(), linalg::types::backend::CopyKind::DeviceToHost);
            return h;
        };
        auto hostf = [](const VecCf& v) { return std::vector<float>(v.data(), v.data() + v.size()); };
        // positive terms: the float loop's rounding error stays near n eps
        const float dc = ops::dot(fx, fy), dg = ops::dot(hxf, hyf);
        report("dot_f32", std::fabs(dc - dg) <= 2e-3f * std::fabs(dc), std::to_string(dc) + " vs " + std::to_string(dg));
        const float nc = ops::norm(fx), ng = ops::norm(hxf);
        report("norm_f32", std::fabs(nc - ng) <= 1e-3f * nc);
        ops::axpy(0.3f, fx, fy);
        ops::axpy(0.3f, hxf, hyf);
        report("axpy_f32_bitwise", downf(hyf) == hostf(fy));
        ops::axpby(0.7f, -1.3f, fx, fy);
        ops::axpby(0.7f, -1.3f, hxf, hyf);
        report("axpby_f32_bitwise", downf(hyf) == hostf(fy));
        ops::scal(1.7f, fy);
        ops::scal(1.7f, hyf);
        report("scal_f32_bitwise", downf(hyf) == hostf(fy));
        ops::multiply(fx, fy, fz);
        ops::multiply(hxf, hyf, hzf);
        report("multiply_f32_bitwise", downf(hzf) == hostf(fz));
        ops::copy(fz, fy);
        ops::copy(hzf, hyf);
        report("copy_f32_bitwise", downf(hyf) == hostf(fy));
    }
    MatC Ac(n, n);
    MatH Ag(n, n);
    load(Ac, hA);
    load(Ag, hA);
    // matvec / lump
    {
        ops::matvec(Ac, cx, cy);
        ops::matvec(Ag, gx, gy);
        report("matvec_bitwise", down(gy) == host(cy));
        ops::lump(Ac, cz);
        ops::lump(Ag, gz);
        report("lump_bitwise", down(gz) == host(cz));
    }
    // operator + Jacobi
    linalg::op::CSROperator<MatC> opC(Ac);
    linalg::op::CSROperator<MatH> opG(Ag);
    {
        report("operator_size_flops", opG.size() == opC.size() && opG.flopsPerApply() == opC.flopsPerApply());
        opC.diagonal(cz);
        opG.diagonal(gz);
        report("diagonal_bitwise", down(gz) == host(cz));
        pc::Jacobi<VecC> Mc;
        pc::Jacobi<VecH> Mg;
        Mc.update(opC);
        Mg.update(opG);
        Mc.apply(cx, cy);
        Mg.apply(gx, gy);
        report("jacobi_apply_bitwise", down(gy) == host(cy) && Mg.flopsPerApply() == Mc.flopsPerApply());
    }
    // missing diagonal throws like CSRMatrix::getDataIndex
    {
        HostCSR h2;
        h2.n = 2;
        h2.rp = {0, 1, 2};
        h2.ci = {1, 0};
        h2.val = {1.0, 1.0};
        MatH B(2, 2);
        load(B, h2);
        linalg::op::CSROperator<MatH> opB(B);
        VecH d(2);
        bool threw = false;
        try {
            opB.diagonal(d);
        } catch (const std::out_of_range&) {
            threw = true;
        }
        report("missing_diagonal_throws", threw);
    }
    // CG: the reference's CPU solver vs the Hip specialisation
    for (int tt = 0; tt < 2; ++tt) {
        using CfgC = linalg::solver::iterative::cg::Config<VecC>;
        using CfgH = linalg::solver::iterative::cg::Config<VecH>;
        CfgC cc;
        CfgH ch;
        cc.tol = ch.tol = tt == 0 ? 1e-12 : 1e-9;
        cc.tolType = tt == 0 ? linalg::solver::iterative::cg::ToleranceType::Relative
                             : linalg::solver::iterative::cg::ToleranceType::Absolute;
        ch.tolType = tt == 0 ? linalg::solver::iterative::cg::ToleranceType::Relative
                             : linalg::solver::iterative::cg::ToleranceType::Absolute;
        cc.maxIters = ch.maxIters = 5000;
        VecC bc(n), xc(n);
        VecH bg(n), xg(n);
        std::copy(hb.begin(), hb.end(), bc.data());
        up(bg, hb);
        xc.zero();
        xg.zero();
        pc::Jacobi<VecC> Mc;
        pc::Jacobi<VecH> Mg;
        RecLogger lc, lg;
        linalg::solver::iterative::cg::Workspace<VecC> Wc(n);
        linalg::solver::iterative::cg::Workspace<VecH> Wg(n);
        linalg::solver::SolverReport<VecC> rc;
        linalg::solver::SolverReport<VecH> rg;
        linalg::solver::iterative::cg::Solver<linalg::op::CSROperator<MatC>, VecC, pc::Jacobi<VecC>, RecLogger> sc(cc);
        linalg::solver::iterative::cg::Solver<linalg::op::CSROperator<MatH>, VecH, pc::Jacobi<VecH>, RecLogger> sg(ch);
        const bool okc = sc.solve(rc, lc, Wc, Mc, opC, bc, xc);
        const bool okg = sg.solve(rg, lg, Wg, Mg, opG, bg, xg);
        const std::string tag = tt == 0 ? "cg_relative" : "cg_absolute";
        report(tag + "_converged", okc && okg && rg.converged);
        report(tag + "_iterations", std::llabs((long long)rc.iterations - (long long)rg.iterations) <= 1,
               std::to_string(rc.iterations) + " vs " + std::to_string(rg.iterations));
        report(tag + "_solution", max_rel(down(xg), host(xc)) < 1e-10, std::to_string(max_rel(down(xg), host(xc))));
        bool logs_ok = lg.resets == 1 && lg.summaries == 1 && lg.lastConverged &&
                       lg.logs.size() == (size_t)rg.iterations + 1;
        for (size_t k = 0; k < lg.logs.size() && k < lc.logs.size() && logs_ok; ++k)
            logs_ok = lg.logs[k].first == k && std::fabs(lg.logs[k].second - lc.logs[k].second) <=
                                                   1e-8 * lc.logs[0].second;
        report(tag + "_logger_history", logs_ok);
        if (tt == 1) report("cg_absolute_criterion", rg.finalResidual < ch.tol);
        // the cached solver matrix picks up new values on the next solve:
        // 2A x = 2b has the same solution
        if (tt == 0) {
            std::vector<double> v2(hA.val);
            for (double& v : v2) v *= 2.0;
            Hip::copy<double>(Ag.data(), v2.data(), v2.size(), linalg::types::backend::CopyKind::HostToDevice);
            ops::scal(2.0, bg);
            xg.zero();
            sg.solve(rg, lg, Wg, Mg, opG, bg, xg);
            report("cg_reuse_values", rg.converged && max_rel(down(xg), host(xc)) < 1e-10);
            Hip::copy<double>(Ag.data(), hA.val.data(), hA.val.size(), linalg::types::backend::CopyKind::HostToDevice);
        }
    }
    // GMRES on a non-symmetric operator
    {
        const HostCSR hN = laplace2d(96, 0.02, true);
        const Index nn = hN.n;
        MatC Nc(nn, nn);
        MatH Ng(nn, nn);
        load(Nc, hN);
        load(Ng, hN);
        linalg::op::CSROperator<MatC> oC(Nc);
        linalg::op::CSROperator<MatH> oG(Ng);
        using CfgC = linalg::solver::iterative::gmres::Config<VecC>;
        using CfgH = linalg::solver::iterative::gmres::Config<VecH>;
        CfgC cc;
        CfgH ch;
        cc.tol = ch.tol = 1e-11;
        cc.maxIters = ch.maxIters = 4000;
        cc.krylovDim = ch.krylovDim = 30;
        std::vector<double> h(nn);
        for (Index i = 0; i < nn; ++i) h[i] = std::cos(0.05 * i);
        VecC bc(nn), xc(nn);
        VecH bg(nn), xg(nn);
        std::copy(h.begin(), h.end(), bc.data());
        up(bg, h);
        xc.zero();
        xg.zero();
        pc::Jacobi<VecC> Mc;
        pc::Jacobi<VecH> Mg;
        RecLogger lc, lg;
        linalg::solver::iterative::gmres::Workspace<VecC> Wc(nn, 30);
        linalg::solver::iterative::gmres::Workspace<VecH> Wg(nn, 30);
        linalg::solver::SolverReport<VecC> rc;
        linalg::solver::SolverReport<VecH> rg;
        linalg::solver::iterative::gmres::Solver<linalg::op::CSROperator<MatC>, VecC, pc::Jacobi<VecC>, RecLogger> sc(cc);
        linalg::solver::iterative::gmres::Solver<linalg::op::CSROperator<MatH>, VecH, pc::Jacobi<VecH>, RecLogger> sg(ch);
        const bool okc = sc.solve(rc, lc, Wc, Mc, oC, bc, xc);
        const bool okg = sg.solve(rg, lg, Wg, Mg, oG, bg, xg);
        report("gmres_converged", okc && okg);
        report("gmres_iterations", std::llabs((long long)rc.iterations - (long long)rg.iterations) <= 2,
               std::to_string(rc.iterations) + " vs " + std::to_string(rg.iterations));
        report("gmres_solution", max_rel(down(xg), host(xc)) < 1e-8, std::to_string(max_rel(down(xg), host(xc))));
    }
    std::printf("%d failure(s)\n", failures);
    return failures;
}
