"""The reference's own linear-algebra templates (Vector, CSRMatrix,
CSROperator, Jacobi, cg / gmres Solver) instantiated on the Hip backend tag of
include/residuum/hip, checked case by case against the same templates on the
CPU tag (tests/cpp/hip_seam_test.cpp).  The executable is compiled where the
reference headers are available and runs wherever the GPU is."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "hip_seam_test")
STAGE_EXE = os.path.join(ROOT, "oracle", "_ref", "hip_stage_test")


@pytest.mark.gpu
def test_reference_templates_on_hip_tag():
    if not os.path.exists(EXE):
        pytest.skip("hip_seam_test is built only where the reference headers are present")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr)
    fails = [ln for ln in r.stdout.splitlines() if ln.startswith("FAIL")]
    assert r.returncode == 0 and not fails, "\n".join(fails) or r.stderr
    passed = [ln for ln in r.stdout.splitlines() if ln.startswith("PASS")]
    assert len(passed) >= 34 and sum("_f32" in ln for ln in passed) == 7


@pytest.mark.gpu
def test_reference_stages_and_runners_on_hip_tag(tmp_path):
    """The reference's SteadyStage / BackwardEulerStage with NewtonRunner,
    BackwardEulerRunner, the Steady / Transient drivers and the solver
    factories over hip::HeatProblem on the Hip tag (Assembler<Hip>,
    BoundaryApplicator<Hip>, device solvers) against the same templates on
    the CPU tag: a linear BE run on hexes (CSR CG), a matrix-free BE run,
    steady Newton with GMRES(30) on Q2 quads and a transient Newton run --
    equal linear and Newton iteration counts per solve, solutions within
    1e-12 (1e-11 / 1e-10 for the Newton runs) of the CPU tag's
    (tests/cpp/hip_stage_test.cpp)."""
    if not os.path.exists(STAGE_EXE):
        pytest.skip("hip_stage_test is built only where the reference headers are present")
    r = subprocess.run([STAGE_EXE, str(tmp_path)], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    print(r.stderr)
    fails = [ln for ln in r.stdout.splitlines() if ln.startswith("FAIL")]
    assert r.returncode == 0 and not fails, "\n".join(fails) or r.stderr
    passed = [ln for ln in r.stdout.splitlines() if ln.startswith("PASS")]
    assert len(passed) == 4 * 5
