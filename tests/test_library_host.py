"""libmrcc.so on the CPU side: it loads, exports every symbol of include/mrcc.h, its host helpers
agree with the reference's order / counts, and its error contract holds (no device work here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import O, pkg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "mrcc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mrcc_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_header_symbol():
    L = pkg.lib()
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    # and the ctypes table covers exactly the header
    assert sorted(pkg._lib.SIGNATURES) == names


def test_abi_version():
    assert pkg.lib().mrcc_abi_version() == 3


def test_host_bfs_indices_match_reference_order():
    L = pkg.lib()
    for n in list(range(1, 200)) + [511, 512, 1000, 4097]:
        out = np.zeros(n, dtype=np.int32)
        assert L.mrcc_bfs_indices(n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))) == 0
        assert tuple(out) == pkg.generate_binary_search_indices(n)


def test_host_state_counts_match_oracle():
    L = pkg.lib()
    rng = np.random.default_rng(0)
    for _ in range(3000):
        N = int(rng.integers(1, 600))
        s = int(rng.integers(0, N + 1))
        m = int(rng.choice([-1, s + int(rng.integers(0, 40)), N + 5]))
        incl = int(rng.integers(0, 2))
        assert L.mrcc_edge_num_states(N, s, m, incl) == O.edge_num_states(N, s, None if m < 0 else m, incl)
    assert L.mrcc_edge_num_states(5, 6, -1, 0) == -1  # N_start > N: the reference asserts


def test_error_contract_and_injection():
    L = pkg.lib()
    h = ctypes.c_void_p()
    r = np.array([0.1, 0.1])
    # invalid arguments are rejected before any device work
    rc = L.mrcc_scene_create_disks(0, 2, pkg._lib.dptr(r), 0, None, None, 0, None, None, ctypes.byref(h))
    assert rc == pkg._lib.MRCC_ERR_INVALID
    assert b"robots" in L.mrcc_last_error()
    # injected failures surface as status codes with a message, then clear
    L.mrcc_debug_inject_failure(pkg._lib.MRCC_ERR_NOMEM, 2)
    for _ in range(2):
        rc = L.mrcc_scene_create_disks(2, 2, pkg._lib.dptr(r), 0, None, None, 0, None, None, ctypes.byref(h))
        assert rc == pkg._lib.MRCC_ERR_NOMEM
        assert L.mrcc_last_error() == b"injected failure"
    rc = L.mrcc_scene_create_disks(0, 2, pkg._lib.dptr(r), 0, None, None, 0, None, None, ctypes.byref(h))
    assert rc == pkg._lib.MRCC_ERR_INVALID
    with pytest.raises(pkg.MrccError):
        pkg._lib.check(rc)


def test_edge_num_states_helper():
    assert pkg.edge_num_states(np.array([-1, 0, 1, 1.0]), np.array([-1, 1, 1, 0.0]), 0.5) == 3
    assert pkg.edge_num_states(np.array([-1, 0, 1, 1.0]), np.array([-1, 1, 1, 0.0]), 0.1) == 11
