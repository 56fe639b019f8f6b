"""C-ABI checks that need no GPU: every declared entry point is exported, host helpers agree with the oracle, and
argument errors are reported as status codes (these return before any launch)."""
import ctypes
import os
import re

import numpy as np
import pytest

from neat_amd import _abi, workloads as W
from neat_amd._abi import NEAT_E_ARG, NEAT_E_OVERFLOW, NEAT_E_UNSUPPORTED, NEAT_OK

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "neat_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(neat_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = _abi.lib()
    names = declared_functions()
    assert len(names) >= 24
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes table covers exactly the declared set
    assert sorted(_abi.SIGNATURES) == names


def test_abi_version():
    assert _abi.lib().neat_abi_version() == 1


@pytest.mark.parametrize("depth", [0, 1, 2, 3, 5])
def test_tree_build_matches_oracle(oracle, depth):
    t = W.build_tree(depth)
    o = oracle.build_tree(depth)
    for k in ("parent", "children", "pos_min", "pos_max", "scale", "depth"):
        assert np.array_equal(t[k], o[k]), k
    assert _abi.lib().neat_tree_validate_host(t["pos_min"].ctypes.data, t["pos_max"].ctypes.data,
                                              t["children"].ctypes.data, t["pos_min"].shape[0], depth) == NEAT_OK


def test_tree_validate_rejects_non_canonical_tables():
    t = W.build_tree(2)
    L = _abi.lib()
    pm = t["pos_max"].copy()
    pm[20, 1] += 0.125
    assert L.neat_tree_validate_host(t["pos_min"].ctypes.data, pm.ctypes.data, t["children"].ctypes.data, 73, 2) \
        == NEAT_E_ARG
    assert "node 20" in _abi.last_error()
    ch = t["children"].copy()
    ch[3, [0, 1]] = ch[3, [1, 0]]
    assert L.neat_tree_validate_host(t["pos_min"].ctypes.data, t["pos_max"].ctypes.data, ch.ctypes.data, 73, 2) \
        == NEAT_E_ARG
    assert L.neat_tree_validate_host(t["pos_min"].ctypes.data, t["pos_max"].ctypes.data, t["children"].ctypes.data,
                                     72, 2) == NEAT_E_ARG


def test_update_active_and_flags(oracle):
    t = W.build_tree(3)
    rng = np.random.default_rng(0)
    act = (rng.random(585) < 0.2).astype(np.int32)
    ids, prefix = W.update_active(act)
    oids, oprefix = oracle.update_active(act)
    assert np.array_equal(ids, oids) and np.array_equal(prefix, oprefix)
    from neat_amd import ops
    flags = ops.tree_flags(act, 3).numpy()
    # brute force: bit 1 iff a strict descendant is active
    desc = np.zeros(585, bool)
    for n in range(585):
        stack = [8 * n + 1 + i for i in range(8)] if t["depth"][n] < 3 else []
        while stack:
            c = stack.pop()
            desc[n] |= act[c] == 1
            if t["depth"][c] < 3:
                stack.extend(8 * c + 1 + i for i in range(8))
    assert np.array_equal(flags & 1, act) and np.array_equal((flags >> 1) & 1, desc.astype(np.uint8))


def test_argument_errors_are_status_codes():
    L = _abi.lib()
    # unsupported decoder width (no kernel), before any device work
    assert L.neat_decoder_fwd(None, None, 10, 8, 48, 1, 1, None, 1, 3, 2.0, None) == NEAT_E_UNSUPPORTED
    assert "hidden=48" in _abi.last_error()
    assert L.neat_decoder_fwd(None, None, 10, 8, 64, 1, 9, None, 1, 3, 2.0, None) == NEAT_E_ARG
    # negative sizes and null pointers
    st = np.zeros(5, np.int64)
    assert L.neat_gs3d_fwd(None, ctypes.c_void_p(8), 1, 8, 3, 3, 3, st.ctypes.data, None, None, -1, None) == NEAT_E_ARG
    assert L.neat_gs3d_fwd(None, None, 1, 8, 3, 3, 3, st.ctypes.data, None, None, 5, None) == NEAT_E_ARG
    # sampler: node count inconsistent with the depth, max_samples_per_node <= 0
    assert L.neat_raysample_count(None, ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8), 70, 2, None,
                                  None, 0, 32, None, None, None, None, 0) == NEAT_E_ARG
    assert L.neat_raysample_count(None, ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8), 73, 2, None,
                                  None, 0, 0, None, None, None, None, 0) == NEAT_E_ARG
    # too-small workspace
    assert L.neat_decoder_bwd(None, ctypes.c_void_p(8), 100, 8, 32, 1, 1, ctypes.c_void_p(8), 1, 3, 2.0,
                              ctypes.c_void_p(8), ctypes.c_void_p(8), None, ctypes.c_void_p(8), 16) == NEAT_E_OVERFLOW
    assert L.neat_mse_fp64(None, None, None, 4, 1.0, ctypes.c_void_p(8), None, None, 0) == NEAT_E_ARG


def test_workspace_queries_scale():
    L = _abi.lib()
    assert L.neat_raysample_workspace_bytes(1 << 18) >= 8 * ((1 << 18) // 1024)
    assert L.neat_decoder_workspace_bytes(1 << 20, 8, 64, 1, 1) >= 4 * 4801
    assert L.neat_decoder_num_params(8, 64, 1, 1) == 4801
    assert L.neat_decoder_num_params(8, 32, 1, 1) == 1377


def test_tv_loss_refuses_more_active_nodes_than_grid_rows():
    """The total variation reads and writes the grid rows its slots name: more active nodes than rows is an argument
    error, returned before any launch (the reference's index_select raises there too)."""
    L = _abi.lib()
    st = np.array([8 * 27, 1, 8 * 9, 8 * 3, 8], np.int64)
    dummy = ctypes.c_void_p(8)
    ws = ctypes.c_void_p(256)
    assert L.neat_tv_loss_fwd(None, dummy, 4, 8, 3, st.ctypes.data, dummy, 5, 1.0, dummy, ws, 1 << 20) == NEAT_E_ARG
    assert "5 active nodes for 4 grid rows" in _abi.last_error()
    assert L.neat_tv_loss_bwd(None, dummy, 4, 8, 3, st.ctypes.data, dummy, 5, 1.0, dummy, dummy) == NEAT_E_ARG
    assert L.neat_tv_loss_fwd_bwd(None, dummy, 4, 8, 3, st.ctypes.data, dummy, 5, 1.0, dummy, dummy, dummy, 1, ws,
                                  1 << 20) == NEAT_E_ARG
