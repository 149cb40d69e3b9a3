"""Write golden vectors produced by the reference's own code.

Runs only where oracle/_ref/libref.so was built from /root/reference
(oracle/Makefile `ref` target).  Each fixture is data only: the inputs
(case parameters) and the reference's outputs, saved as .npz next to this
script.  The test suite checks the CPU oracle against them bit for bit.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_py as O  # noqa: E402

# (name, dim, n, p, nq, tags, bc kinds, ext)
CASES = [
    ("fem2d_q1", 2, (4, 4), (1, 1), 2, [0, 1, 3], [1, 1, 1], (0.0, 4.0, 0.0, 4.0)),
    ("fem3d_q1", 3, (4, 4, 4), (1, 1, 1), 2, [0, 1, 2, 3], [1, 1, 1, 1], (0.0, 4.0, 0.0, 4.0, 0.0, 4.0)),
    ("quad_q2", 2, (5, 3), (2, 2), 3, [0, 1, 2, 3], [2, 2, 2, 2], None),
    ("hex_q1", 3, (5, 4, 3), (1, 1, 1), 2, [0, 1], [6, 0], None),
    ("quad_q1_cfg1", 2, (16, 16), (1, 1), 2, [0, 1, 2, 3], [2, 2, 2, 2], None),
]

TENSOR2 = (1.0, 0.2, 0.2, 0.1)
TENSOR3 = (1.0, 0.2, 0.0, 0.2, 0.1, 0.0, 0.0, 0.0, 1.0)


def main():
    if not O.have_ref():
        sys.exit("oracle/_ref/libref.so missing: run `make -C oracle ref` where /root/reference exists")
    for name, dim, n, p, nq, tags, kinds, ext in CASES:
        r = O.RefCase(dim, n, p, nq, tags, kinds, ext=ext)
        rng = np.random.default_rng(12345)
        kappa = np.exp(0.5 * rng.standard_normal(r.n_elem))
        U = rng.standard_normal(r.n_free)
        b = rng.standard_normal(r.n_free)
        nodal = rng.random(r.n_nodes)
        out = dict(dim=dim, n=np.array(n), p=np.array(p), nq=nq, tags=np.array(tags), kinds=np.array(kinds),
                   ext=np.array(ext if ext is not None else (0.0, 1.0) * dim), kappa=kappa, U=U, b=b, nodal=nodal)
        out["xyz"], out["ien"], out["rng"] = r.mesh()
        out["t2a"], out["ctag"] = r.dofs()
        out["row_ptr"], out["col_idx"] = r.pattern()
        out["elem_map"] = r.map()
        out["K"] = r.assemble_matrix(0)
        out["M"] = r.assemble_matrix(1)
        out["A_be"] = r.assemble_matrix(2, 250.0)
        out["K_kappa"] = r.assemble_matrix(0, kappa=kappa)
        out["A_be_kappa"] = r.assemble_matrix(2, 250.0, kappa=kappa)
        out["KU_kappa"] = r.assemble_vector(0, U=U, kappa=kappa)
        out["MU"] = r.assemble_vector(1, U=U)
        src = 5 if dim == 3 else 4
        out["src_kind"] = src
        out["F_load"] = r.load(src)
        out["F_nodal"] = r.load(0, nodal=nodal)
        out["F_lift"] = r.lift(0, out["F_load"])
        out["F_lift_be_kappa"] = r.lift(2, out["F_load"], 250.0, kappa=kappa)
        r.set_model(tdep=1, ka=1.0, kb=0.5, tensor=TENSOR3 if dim == 3 else TENSOR2 + (0,) * 5)
        out["J_tdep"] = r.assemble_matrix(3, U=U)
        r.set_model()
        A = out["A_be_kappa"]
        for pc in (0, 1):
            x, ok, it = r.solve(0, pc, A, b, tol=1e-10, max_iters=2000)
            out[f"cg_pc{pc}_x"], out[f"cg_pc{pc}_iters"] = x, it
            x, ok, it = r.solve(1, pc, A, b, tol=1e-10, max_iters=2000, krylov=7)
            out[f"gmres_pc{pc}_x"], out[f"gmres_pc{pc}_iters"] = x, it
        Ube, steps, its = r.be_run(0.0, 0.05, 0.01, tol=1e-9, kappa=kappa)
        out["be_U"], out["be_steps"], out["be_iters"] = Ube, steps, its
        if dim == 2:
            r.set_model(tdep=1, ka=1.0, kb=0.5, tensor=TENSOR2 + (0,) * 5)
            rz = O.RefCase(dim, n, p, nq, tags, [0] * len(tags), ext=ext)
            rz.set_model(tdep=1, ka=1.0, kb=0.5, tensor=TENSOR2 + (0,) * 5)
            Un, its_n, lin = rz.newton(src_kind=3, abs_tol=1e-10, rel_tol=1e-10, lin_tol=1e-12, krylov=30)
            out["newton_U"], out["newton_iters"], out["newton_lin_iters"] = Un, its_n, lin
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"wrote {path}: n_free={r.n_free} nnz={r.nnz}")


if __name__ == "__main__":
    main()
