"""The C-ABI library: it exists, exports every symbol include/pgn.h declares,
its host-only pattern builder agrees with the oracle's assembly pattern, and
it refuses to run (loudly) without a gfx950 device."""
import os
import re
import ctypes

import numpy as np
import pytest

import physim_amd as pg
from oracle import orc
from tests.helpers import oracle_model, small_beam

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "pgn.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pgn_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = pg.lib()
    names = _declared()
    assert len(names) >= 30
    for name in names:
        assert hasattr(L, name), name
    assert sorted(pg.EXPORTS) == names


def test_host_pattern_matches_oracle_assembly_pattern():
    s = small_beam(dims=(2, 3, 4))
    rp, col = pg.host_bsr_pattern(s.n, s.T)
    M = oracle_model(s)
    lrp, lci, _ = M.hessian_lower()
    rows = np.repeat(np.arange(3 * s.n), np.diff(lrp))
    oracle_blocks = set(zip((rows // 3).tolist(), (lci // 3).tolist()))
    brow = np.repeat(np.arange(s.n), np.diff(rp))
    ours = set((int(i), int(j)) for i, j in zip(brow, col) if j <= i)
    assert ours == oracle_blocks
    # full pattern is symmetric and rows are sorted
    full = set((int(i), int(j)) for i, j in zip(brow, col))
    assert all((j, i) in full for i, j in full)
    for i in range(s.n):
        assert np.all(np.diff(col[rp[i]:rp[i + 1]]) > 0)


def test_host_pattern_rejects_bad_node_ids():
    T = np.array([[0, 1, 2, 7]], dtype=np.int32)
    with pytest.raises(pg.PgnError, match="out of range"):
        pg.host_bsr_pattern(4, T)


def _has_gpu():
    return os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK)


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_context_fails_loudly_without_device():
    h = ctypes.c_void_p()
    rc = pg.lib().pgn_ctx_create(0, ctypes.byref(h))
    assert rc == -4
    assert b"gfx950" in pg.lib().pgn_last_error() or b"device" in pg.lib().pgn_last_error()
    with pytest.raises(pg.PgnError):
        pg.Context(0)
