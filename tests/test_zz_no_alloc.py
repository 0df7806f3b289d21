"""pgn_csr_linear_solve_noalloc (no counterpart in the reference; for callers
that cannot allocate device memory) against pgn_csr_linear_solve on the same
full CSR systems: the same kernels over the same inputs, so the same
iteration counts, regularization tries, solutions and shifted diagonals,
bit for bit.  Two calls in a row share the static work space, the second on
a smaller system (and a third needing regularization after a larger one), so
state left by one call must not leak into the next.  Device buffers are
allocated by the test with the HIP runtime libpgn itself links (plumbing
only); every array passed is a device address."""
import ctypes
import sys

import numpy as np
import pytest
import scipy.sparse

import physim_amd as pg
from physim_amd import scenes
from tests.helpers import oracle_model, perturbed_x

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _full_system(dims, seed):
    s = scenes.cantilever(dims=dims, scale=(0.1, 0.1, 0.4))
    O = oracle_model(s)
    for _ in range(s.stages):
        O.step_load()
    O.set_x(perturbed_x(s, 0.002, seed=seed))
    rp, ci, v = O.hessian_lower()
    n = rp.size - 1
    L = scipy.sparse.csr_matrix((v, ci, rp), shape=(n, n))
    A = (scipy.sparse.tril(L) + scipy.sparse.tril(L, -1).T).tocsr()
    A.sort_indices()
    b = np.random.default_rng(seed).standard_normal(n)
    return A.indptr.astype(np.int64), A.indices.astype(np.int32), A.data.copy(), b


_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        pg.lib()  # libpgn first: its HIP runtime is the one loaded below
        _HIP = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        _HIP.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _HIP.hipFree.argtypes = [ctypes.c_void_p]
        _HIP.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_int]
        _HIP.hipDeviceSynchronize.argtypes = []
    return _HIP


class _Dev:
    """A device copy of a numpy array (hipMalloc / hipMemcpy by the test)."""

    H2D, D2H = 1, 2

    def __init__(self, a):
        self.a = np.ascontiguousarray(a)
        self.ptr = ctypes.c_void_p()
        assert _hip().hipMalloc(ctypes.byref(self.ptr), max(self.a.nbytes, 8)) == 0
        if self.a.nbytes:
            assert _hip().hipMemcpy(self.ptr, self.a.ctypes.data, self.a.nbytes, self.H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        if out.nbytes:
            assert _hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        _hip().hipFree(self.ptr)


def _both(ctx, rp, ci, v, b, x0, **opts):
    vh = v.copy()
    xh, rh = pg.csr_linear_solve(ctx, rp, ci, vh, b, x0=x0, layout=pg.LAYOUT_FULL, **opts)
    d = [_Dev(rp), _Dev(ci), _Dev(v.copy()), _Dev(b), _Dev(np.array(x0, dtype=np.float64))]
    assert _hip().hipDeviceSynchronize() == 0
    rd = pg.csr_linear_solve_noalloc(ctx, rp.size - 1, *[q.ptr.value for q in d], **opts)
    out = (d[4].get(), rd, d[2].get())
    for q in d:
        q.free()
    return (xh, rh, vh), out


def _same(h, d):
    (xh, rh, vh), (xd, rd, vd) = h, d
    assert rh == rd, (rh, rd)
    assert np.array_equal(xh, xd)
    assert np.array_equal(vh, vd)


def test_noalloc_matches_allocating_entry_point(ctx):
    opts = dict(tol=1e-10, max_iters=5000, reg_iters=10, precond=0)
    for dims, seed in [((6, 6, 24), 1), ((2, 2, 6), 2)]:
        _say(f"noalloc vs allocating: beam {dims}")
        rp, ci, v, b = _full_system(dims, seed)
        h, d = _both(ctx, rp, ci, v, b, np.zeros(rp.size - 1), **opts)
        _say(f"  {h[1]}")
        assert h[1]["status"] == 0 and h[1]["iterations"] > 0
        _same(h, d)
    # regularization after a larger call: the shifted diagonal and the tries
    _say("noalloc vs allocating: indefinite diagonal")
    n = 6
    rp = np.arange(n + 1, dtype=np.int64)
    ci = np.arange(n, dtype=np.int32)
    v = np.array([4.0, 3.0, 2.0, 1.0, -0.5, -1.0])
    b = np.array([0.0, 0.0, 0.0, 0.0, 1.0, 1.0])
    h, d = _both(ctx, rp, ci, v, b, np.zeros(n), tol=1e-12, max_iters=100, reg_iters=10)
    assert h[1]["reg_used"] > 0
    _same(h, d)


def test_noalloc_rejects_oversized_and_malformed_input(ctx):
    t = _Dev(np.zeros(8))
    with pytest.raises(pg.PgnError, match="PGN_CSR_NOALLOC_MAX_ROWS"):
        pg.csr_linear_solve_noalloc(ctx, 262145, *([t.ptr.value] * 5))
    t.free()
    # a column index out of range is caught on the device before any product
    d = [_Dev(np.array([0, 1, 2], np.int64)), _Dev(np.array([0, 5], np.int32)),
         _Dev(np.ones(2)), _Dev(np.ones(2)), _Dev(np.ones(2))]
    with pytest.raises(pg.PgnError, match="malformed"):
        pg.csr_linear_solve_noalloc(ctx, 2, *[q.ptr.value for q in d])
    for q in d:
        q.free()
