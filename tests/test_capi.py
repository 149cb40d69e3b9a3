"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, rejects bad arguments with a status code and a message,
and its host-side generators match the oracle.  No GPU is touched."""
import ctypes as C

import numpy as np

import oracle_py as O
import residuum_hip as rh


def test_library_exports_every_declared_symbol():
    L = rh.lib()
    names = rh.exported_symbols()
    assert len(names) > 30
    missing = [s for s in names if not hasattr(L, s)]
    assert missing == []
    assert L.rsd_version().decode().startswith("residuum-hip")


def test_null_arguments_map_to_status_codes():
    L = rh.lib()
    assert L.rsd_ctx_create(0, None) == -1
    assert b"null" in L.rsd_last_error_string()
    assert L.rsd_memcpy(None, None, None, C.c_size_t(8), 1) == -1
    assert L.rsd_problem_sizes(None, None) == -1
    assert L.rsd_cg_solve_f64(None, None, None, None, C.c_double(1e-8), 0, C.c_int64(10), 1, None, None) == -1
    assert L.rsd_philox_normal(C.c_uint64(1), C.c_uint32(0), C.c_int64(-1), None) == -1


def test_philox_matches_oracle():
    for seed, stream in ((20261201, 0), (7, 3), (2 ** 40 + 5, 11)):
        assert np.array_equal(rh.philox_normal(seed, stream, 5000), O.normal_field(seed, stream, 5000))
        assert np.array_equal(rh.philox_uniform(seed, stream, 5000), O.uniform_field(seed, stream, 5000))
    z = rh.philox_normal(1, 0, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
