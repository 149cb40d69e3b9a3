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


# Random123 known-answer vectors for philox4x32-10 (Salmon et al., SC'11;
# the kat_vectors file of the Random123 distribution): counter, key, output
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


def test_philox_known_answer_vectors():
    """The oracle's Philox4x32-10 block against the published vectors, and the
    library's uniform draw for (i, stream, seed) = 0 against the first one:
    u = ((w0 << 32 | w1) >> 11 + 0.5) / 2^53."""
    for ctr, key, want in PHILOX_KAT:
        out = (C.c_uint32 * 4)()
        O.lib().orc_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
        assert tuple(out) == want
    w = PHILOX_KAT[0][2]
    u = ((((w[0] << 32) | w[1]) >> 11) + 0.5) / 2.0 ** 53
    assert rh.philox_uniform(0, 0, 1)[0] == u
    assert O.uniform_field(0, 0, 1)[0] == u


def test_binding_declares_every_function_from_the_header():
    """Every rsd_* declaration of include/residuum_hip.h is exported and bound
    with its argument types: a call with the wrong arity or a scalar of the
    wrong kind raises before reaching the library."""
    import pytest
    from residuum_hip import _abi
    L = rh.lib()
    decls = _abi.declarations()
    assert len(decls) >= 90 and all(hasattr(L, n) for n in decls)
    assert L.rsd_philox_normal.argtypes == [C.c_uint64, C.c_uint32, C.c_int64, C.c_void_p]
    assert L.rsd_last_error_string.restype is C.c_char_p
    with pytest.raises(TypeError):
        L.rsd_philox_normal(1, 0, 5)
    with pytest.raises(C.ArgumentError):
        L.rsd_philox_normal(1, 0, 5.0, None)


def test_stale_library_is_refused(tmp_path):
    """The library reports the SHA-256 prefix of the header it was built
    against; binding it to a header with other contents is refused."""
    import pytest
    from residuum_hip import _abi
    L = rh.lib()
    assert L.rsd_abi_hash().decode() == _abi.header_hash()
    other = tmp_path / "residuum_hip.h"
    other.write_text(open(_abi.HEADER).read() + "\n/* one more line */\n")
    with pytest.raises(RuntimeError, match="another residuum_hip.h"):
        _abi.bind(C.CDLL(rh.LIB_PATH), path=str(other))
