// The reference's own solver stages and runners -- SteadyStage,
// BackwardEulerStage, NewtonRunner, BackwardEulerRunner, the Steady and
// Transient drivers and the linear / time-stepper factories -- driving
// hip::HeatProblem on the CPU tag (the reference's CPU assembly, boundary
// applicator and solvers) and on the Hip tag (Assembler<Hip>,
// BoundaryApplicator<Hip>, the device solvers), compared case by case.
//
// Each case builds the same HeatConfig, as HeatDispatcher.cpp:68-110 does
// for a `backend:` choice, runs it on both tags and compares the linear
// (and Newton) iteration counts per solve, read from the reference's CSV
// loggers, and the final solutions.  Prints "PASS <name>" / "FAIL <name>:
// why"; the exit status is the number of failures.
//
//   hip_stage_test <scratch dir> [cpu]     (cpu: the CPU tag only)
#include "residuum/hip/HeatProblem.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "mesh/generator/BlockMesh2D.hpp"
#include "mesh/generator/BlockMesh3D.hpp"
#include "solver/driver/Steady.hpp"
#include "solver/driver/Transient.hpp"
#include "solver/stage/BackwardEulerStage.hpp"
#include "solver/stage/SteadyStage.hpp"
#include "solver/timestepper/TimeStepperFactory.hpp"

using namespace residuum;
namespace hc = application::heateq::config;
namespace sc = solver::config;
using CPU = linalg::types::backend::CPU;
using Hip = linalg::types::backend::Hip;

static int failures = 0;
static void report(const std::string& name, bool ok, const std::string& why = "") {
    std::printf("%s %s%s%s\n", ok ? "PASS" : "FAIL", name.c_str(), ok ? "" : ": ", ok ? "" : why.c_str());
    std::fflush(stdout);
    if (!ok) ++failures;
}

struct Case {
    std::string name;
    int dim = 2, n[3] = {8, 8, 8}, p = 1, nq = 2;
    hc::HeatConfig cfg;
    double tol = 1e-12;  // solution agreement, relative to max |U|
};

struct Result {
    bool converged = false;
    std::vector<double> U;
    std::vector<long> linIters, newtonIters;  // per solve, from the CSV loggers
    std::vector<double> linStart;              // per linear solve: the initial residual norm
};

// iterations per solve: the largest `iter` logged under each outer tick;
// start (optional) receives the third column of each tick's first row (the
// linear loggers' abs residual at iteration 0)
static std::vector<long> iterations(const std::string& csv, std::vector<double>* start = nullptr) {
    std::ifstream f(csv);
    std::string line;
    std::getline(f, line);
    std::map<long, long> last;
    std::vector<long> order;
    while (std::getline(f, line)) {
        std::stringstream ss(line);
        double tick = 0, it = 0, a = 0;
        char c;
        ss >> tick >> c >> it >> c >> a;
        if (!last.count((long)tick)) {
            order.push_back((long)tick);
            if (start) start->push_back(a);
        }
        last[(long)tick] = (long)it;
    }
    std::vector<long> out;
    for (long t : order) out.push_back(last[t]);
    return out;
}

// Linear iteration counts must be equal on both tags, except for a Newton
// solve that starts below 1e-8 of its sequence's first residual: its
// right-hand side is the residual near convergence, whose rounding (about
// 1e-16 of the assembled terms) differs between the tags because their
// reductions sum in different orders, i.e. it differs from the other tag's
// right-hand side by ~1e-7 relative, and the Krylov solver may need a few
// iterations more or less on it.  Those counts must agree within 20 %.
static bool counts_agree(const Result& a, const Result& b) {
    if (a.linIters.empty() || a.linIters.size() != b.linIters.size() || a.linStart.size() != a.linIters.size())
        return false;
    std::vector<bool> exact(a.linIters.size(), true);
    size_t j = 0;
    for (long k : a.newtonIters) {
        const double first = j < a.linStart.size() ? a.linStart[j] : 0.0;
        for (long i = 0; i < k && j < a.linStart.size(); ++i, ++j)
            if (a.linStart[j] < 1e-8 * first) exact[j] = false;
    }
    for (size_t i = 0; i < a.linIters.size(); ++i) {
        const long d = std::labs(a.linIters[i] - b.linIters[i]);
        if (exact[i] ? d != 0 : 5 * d > std::max(a.linIters[i], b.linIters[i])) return false;
    }
    return true;
}

template<typename BackendT, Index D, mesh::ElementFamily Family>
Result run(const Case& c, const std::string& dir, const std::string& tag) {
    using Bundle = hip::HeatEquation<D, D, Family>;
    using ProblemT = hip::HeatProblem<BackendT, Bundle>;
    hc::HeatConfig cfg = c.cfg;
    cfg.logging.linear.csvFile = dir + "/" + c.name + "_" + tag + "_linear.csv";
    cfg.logging.nonlinear.csvFile = dir + "/" + c.name + "_" + tag + "_newton.csv";
    std::remove(cfg.logging.linear.csvFile.c_str());
    std::remove(cfg.logging.nonlinear.csvFile.c_str());
    Result r;
    {
        mesh::Mesh m = [&] {
            if constexpr (D == 2) return mesh::generator::BlockMesh2D{(Index)c.n[0], (Index)c.n[1], 0.0, 1.0, 0.0, 1.0, (Index)c.p, (Index)c.p}.generate();
            else return mesh::generator::BlockMesh3D{(Index)c.n[0], (Index)c.n[1], (Index)c.n[2], 0.0, 1.0, 0.0, 1.0, 0.0, 1.0, (Index)c.p, (Index)c.p, (Index)c.p}.generate();
        }();
        typename Bundle::Basis basis = [&] {
            if constexpr (D == 2) return typename Bundle::Basis((Index)c.p, (Index)c.p);
            else return typename Bundle::Basis((Index)c.p, (Index)c.p, (Index)c.p);
        }();
        typename Bundle::QuadratureVolumeType qv = [&] {
            if constexpr (D == 2) return typename Bundle::QuadratureVolumeType((Index)c.nq, (Index)c.nq);
            else return typename Bundle::QuadratureVolumeType((Index)c.nq, (Index)c.nq, (Index)c.nq);
        }();
        typename Bundle::QuadratureBoundaryType qb = [&] {
            if constexpr (D == 2) return typename Bundle::QuadratureBoundaryType((Index)c.nq);
            else return typename Bundle::QuadratureBoundaryType((Index)c.nq, (Index)c.nq);
        }();
        ProblemT pb(cfg, std::move(m), std::move(basis), std::move(qv), std::move(qb));
        if (cfg.solver.driver.type == sc::DriverConfig::Type::Steady) {
            using StageT = solver::stage::SteadyStage<ProblemT>;
            StageT stage(pb);
            solver::driver::Steady<StageT> driver;
            r.converged = driver.solve(stage);
        } else {
            r.converged = solver::timestepper::dispatchTransientStage(
                pb, *pb.solverInstance().timestepper, cfg.logging.timestepper, pb.equationLabel(),
                [](auto& stage, auto& stepper) {
                    using StageT = std::decay_t<decltype(stage)>;
                    solver::driver::Transient<StageT> driver;
                    return driver.solve(stage, stepper);
                });
        }
        r.U.resize(pb.U().size());
        BackendT::template copy<Real>(r.U.data(), pb.U().data(), r.U.size(), linalg::types::backend::CopyKind::DeviceToHost);
    }
    r.linIters = iterations(cfg.logging.linear.csvFile, &r.linStart);
    r.newtonIters = iterations(cfg.logging.nonlinear.csvFile);
    return r;
}

template<Index D, mesh::ElementFamily Family>
void compare(const Case& c, const std::string& dir, bool cpu_only) {
    const Result a = run<CPU, D, Family>(c, dir, "cpu");
    report(c.name + "_cpu_converged", a.converged);
    if (cpu_only) {
        std::printf("  %s: %zu linear solves, %zu Newton solves\n", c.name.c_str(), a.linIters.size(), a.newtonIters.size());
        return;
    }
    const Result b = run<Hip, D, Family>(c, dir, "hip");
    report(c.name + "_hip_converged", b.converged);
    auto show = [](const std::vector<long>& v) {
        std::string s;
        for (long x : v) s += std::to_string(x) + " ";
        return s;
    };
    report(c.name + "_linear_iterations", counts_agree(a, b),
           "cpu [" + show(a.linIters) + "] hip [" + show(b.linIters) + "]");
    report(c.name + "_newton_iterations", a.newtonIters == b.newtonIters,
           "cpu [" + show(a.newtonIters) + "] hip [" + show(b.newtonIters) + "]");
    double m = 0.0, s = 0.0;
    for (size_t i = 0; i < a.U.size() && a.U.size() == b.U.size(); ++i) {
        m = std::max(m, std::fabs(a.U[i] - b.U[i]));
        s = std::max(s, std::fabs(a.U[i]));
    }
    const double rel = s > 0 ? m / s : m;
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.3e", rel);
    report(c.name + "_solution", a.U.size() == b.U.size() && rel <= c.tol, std::string("max rel diff ") + buf);
    std::printf("  %s: linear [%s] newton [%s] rel diff %s\n", c.name.c_str(), show(b.linIters).c_str(),
                show(b.newtonIters).c_str(), buf);
}

static sc::LinearSolverConfig linear(sc::LinearSolverConfig::Type t, double tol, Index maxit,
                                     sc::LinearSolverConfig::OperatorType op = sc::LinearSolverConfig::OperatorType::CSR) {
    sc::LinearSolverConfig l;
    l.type = t;
    l.operatorType = op;
    l.tolerance = tol;
    l.maxIterations = maxit;
    l.preconditioner.type = sc::PreconditionerConfig::Type::Jacobi;
    if (t == sc::LinearSolverConfig::Type::GMRES) l.params = sc::GMRESParams{30};
    return l;
}

static hc::BoundaryConditionConfig value_bc(Int tag, const std::string& e) {
    hc::BoundaryConditionConfig bc;
    bc.boundaryID = tag;
    bc.type = hc::BoundaryConditionConfig::Type::Value;
    bc.expression = e;
    bc.forms = {hc::BoundaryConditionConfig::Form::ValueBC};
    return bc;
}

static hc::BoundaryConditionConfig flux_bc(Int tag, std::vector<std::string> q) {
    hc::BoundaryConditionConfig bc;
    bc.boundaryID = tag;
    bc.type = hc::BoundaryConditionConfig::Type::Flux;
    bc.fluxExpression = std::move(q);
    bc.forms = {hc::BoundaryConditionConfig::Form::FluxBC};
    return bc;
}

static hc::HeatConfig base(int nq, bool transient, double dt, double tf) {
    hc::HeatConfig cfg;
    cfg.discretization.quadrature.xi = cfg.discretization.quadrature.eta = cfg.discretization.quadrature.zeta = nq;
    cfg.discretization.dofOrdering = fem::dof::DOFOrdering::Interleaved;
    cfg.logging.linear.type = cfg.logging.nonlinear.type = cfg.logging.timestepper.type = cfg.logging.driver.type =
        sc::LoggerConfig::Type::None;
    cfg.source.read.mode = sc::NodalFieldReadConfig::Mode::Expression;
    cfg.source.read.expression = "0";
    cfg.initialCondition.read.mode = sc::NodalFieldReadConfig::Mode::Expression;
    cfg.initialCondition.read.expression = "0";
    cfg.conductivity.type = hc::ConductivityConfig::Type::Constant;
    cfg.conductivity.value = 1.0;
    cfg.solver.driver.type = transient ? sc::DriverConfig::Type::Transient : sc::DriverConfig::Type::Steady;
    if (transient) {
        sc::TimeStepperConfig ts;
        ts.type = sc::TimeStepperConfig::Type::BackwardEuler;
        ts.t0 = 0.0;
        ts.tf = tf;
        ts.stepSize.mode = sc::TimeStepSizeConfig::Mode::Constant;
        ts.stepSize.dt = dt;
        cfg.solver.timestepper = ts;
        cfg.density = hc::ScalarMaterialPropertyConfig{1.5, "kg/m^3"};
        cfg.specificHeat = hc::SpecificHeatConfig{hc::SpecificHeatConfig::Type::Constant, 2.0, "", "", "J/(kg*K)"};
    }
    return cfg;
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    const bool cpu_only = argc > 2 && std::string(argv[2]) == "cpu";
    try {
        // linear Backward Euler on Q1 hexes, CSR + CG + Jacobi: Dirichlet
        // faces, a flux face, a source, per-step load and lift
        {
            Case c;
            c.name = "be_hex_cg";
            c.dim = 3;
            c.n[0] = 10; c.n[1] = 8; c.n[2] = 6;
            c.cfg = base(2, true, 0.01, 0.05);
            c.cfg.conductivity.value = 2.0;
            c.cfg.source.read.expression = "1 + x*y";
            c.cfg.boundaryConditions = {value_bc(0, "1"), value_bc(1, "0"), flux_bc(4, {"0", "0", "-2*x"})};
            auto l = linear(sc::LinearSolverConfig::Type::CG, 1e-12, 2000);
            c.cfg.solver.linear = l;
            c.cfg.solver.timestepper->linearSolver = l;
            compare<3, mesh::ElementFamily::Hex>(c, dir, cpu_only);
        }
        // the same transient run on Q1 quads with the matrix-free operator
        // (FEMOperator over Assembler<Hip>), anisotropic conductivity
        {
            Case c;
            c.name = "be_quad_fem";
            c.n[0] = 14; c.n[1] = 11;
            c.cfg = base(2, true, 0.02, 0.06);
            c.cfg.conductivity.type = hc::ConductivityConfig::Type::Anisotropic;
            c.cfg.conductivity.tensor = {{1.0, 0.25}, {0.25, 0.5}};
            c.cfg.initialCondition.read.expression = "x*(1-x)";
            c.cfg.boundaryConditions = {value_bc(0, "0"), value_bc(2, "y*0 + t"), flux_bc(1, {"1 + y", "0"})};
            auto l = linear(sc::LinearSolverConfig::Type::CG, 1e-12, 2000, sc::LinearSolverConfig::OperatorType::FEM);
            c.cfg.solver.linear = l;
            c.cfg.solver.timestepper->linearSolver = l;
            compare<2, mesh::ElementFamily::Quad>(c, dir, cpu_only);
        }
        // steady nonlinear heat on Q2 quads: k(T) = (1 + 0.5 T) diag(1, 0.1),
        // f = 10 sin(pi x) sin(pi y), Newton with GMRES(30) + Jacobi
        {
            Case c;
            c.name = "newton_q2_gmres";
            c.n[0] = 12; c.n[1] = 12;
            c.p = 2;
            c.nq = 3;
            c.cfg = base(3, false, 0, 0);
            c.cfg.conductivity.type = hc::ConductivityConfig::Type::TemperatureDependentAnisotropic;
            c.cfg.conductivity.tensor = {{1.0, 0.0}, {0.0, 0.1}};
            c.cfg.conductivity.valueExpression = "1.0 + 0.5*T";
            c.cfg.conductivity.gradientExpression = "0.5";
            c.cfg.source.read.expression = "10*sin(pi*x)*sin(pi*y)";
            c.cfg.boundaryConditions = {value_bc(0, "0"), value_bc(1, "0"), value_bc(2, "0"), value_bc(3, "0")};
            sc::NonlinearSolverConfig nl;
            nl.type = sc::NonlinearSolverConfig::Type::Newton;
            nl.absoluteTolerance = 1e-14;
            nl.relativeTolerance = 1e-11;
            nl.maxIterations = 20;
            nl.linearSolver = linear(sc::LinearSolverConfig::Type::GMRES, 1e-13, 20000);
            c.cfg.solver.nonlinear = nl;
            c.tol = 1e-11;
            compare<2, mesh::ElementFamily::Quad>(c, dir, cpu_only);
        }
        // transient nonlinear: BackwardEulerStage with its NewtonRunner (the
        // Jacobian K + K_T + M/dt + M_T), k(T) = 1 + T, a flux face
        {
            Case c;
            c.name = "be_newton_quad";
            c.n[0] = 10; c.n[1] = 9;
            c.cfg = base(2, true, 0.05, 0.15);
            c.cfg.conductivity.type = hc::ConductivityConfig::Type::TemperatureDependentIsotropic;
            c.cfg.conductivity.valueExpression = "1 + T";
            c.cfg.conductivity.gradientExpression = "1";
            c.cfg.boundaryConditions = {value_bc(0, "1"), flux_bc(1, {"-0.5", "0"})};
            sc::NonlinearSolverConfig nl;
            nl.type = sc::NonlinearSolverConfig::Type::Newton;
            nl.absoluteTolerance = 1e-14;
            nl.relativeTolerance = 1e-10;
            nl.maxIterations = 20;
            nl.linearSolver = linear(sc::LinearSolverConfig::Type::GMRES, 1e-11, 20000);
            c.cfg.solver.nonlinear = nl;
            c.cfg.solver.timestepper->nonlinearSolver = nl;
            c.cfg.solver.timestepper->linearSolver = nl.linearSolver;
            c.tol = 1e-10;
            compare<2, mesh::ElementFamily::Quad>(c, dir, cpu_only);
        }
    } catch (const std::exception& e) {
        report("exception", false, e.what());
    }
    std::printf("%d failure(s)\n", failures);
    return failures;
}
