"""physim_amd -- MI355X-native PhySim Newton inner loop.

The product is libpgn.so (HIP kernels for gfx950 + C++ host driver behind the
C ABI of include/pgn.h).  This module is a thin ctypes binding of that ABI for
tests and benchmarks; it has no compute of its own and raises if the library
is missing or no gfx950 device is present.
"""
import ctypes
import weakref
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpgn.so")

CONH, STVK = 0, 1
FP64, MIXED = 0, 1
LS_EIGENCG, LS_EIGENLDLT = 1, 2  # LSSolverType values (BasicTypes.h:36-46)
LAYOUT_LOWER, LAYOUT_FULL, LAYOUT_BSR3 = 0, 1, 2
# Newton options of a solve "to residual 1e-8": the reference's MINIMPR /
# MINSTEP exits and its energy-only line search would stop short of it
QP_STEEPEST, QP_NEWTON, QP_LBFGS = 0, 1, 2  # QPSolverType (BasicTypes.h:48-55)
CONVERGE_OPTS = dict(tolMaxError=1e-8, tolMinImpr=0.0, tolMinStep=0.0,
                     ls_noise_guard=1e-12)
OS_NAMES = ["ONGOING", "SUCCESS", "MINSTEP", "MINIMPR", "FAILURE", "MAXITER",
            "NONDESC"]

# every symbol include/pgn.h declares
EXPORTS = [
    "pgn_last_error", "pgn_ctx_create", "pgn_ctx_destroy", "pgn_ctx_stream",
    "pgn_ctx_synchronize", "pgn_model_create_tet", "pgn_model_destroy",
    "pgn_model_num_dofs", "pgn_model_num_elements", "pgn_model_device_bytes",
    "pgn_model_set_fixed", "pgn_model_set_gravity", "pgn_model_set_forces",
    "pgn_model_step_load", "pgn_model_is_loaded", "pgn_model_reset",
    "pgn_model_set_x", "pgn_model_get_x", "pgn_model_set_velocity",
    "pgn_model_get_velocity", "pgn_dynamic_step", "pgn_model_get_energy",
    "pgn_model_get_gradient", "pgn_model_hessian_nnz",
    "pgn_model_hessian_pattern", "pgn_model_get_hessian",
    "pgn_model_element_egh", "pgn_linear_default_opts", "pgn_linear_solve",
    "pgn_model_hessian_product", "pgn_model_precond_apply", "pgn_model_coarse_space",
    "pgn_rod_bt_eval",
    "pgn_newton_default_opts", "pgn_newton_solve", "pgn_time_kernel",
    "pgn_kernel_bytes", "pgn_host_bsr_pattern", "pgn_model_set_profiling",
    "pgn_model_get_profile", "pgn_model_create_shell", "pgn_batch_create",
    "pgn_batch_destroy", "pgn_batch_add_tet", "pgn_batch_add_shell", "pgn_batch_set_fixed",
    "pgn_batch_set_gravity", "pgn_batch_set_forces", "pgn_batch_build",
    "pgn_batch_num_instances", "pgn_batch_device_bytes", "pgn_batch_reset",
    "pgn_batch_newton", "pgn_batch_get_x", "pgn_batch_get_energy", "pgn_server_default_opts",
    "pgn_server_create", "pgn_server_destroy", "pgn_server_submit_tet",
    "pgn_server_submit_shell", "pgn_server_poll", "pgn_server_wait", "pgn_server_stats",
    "pgn_rods_create", "pgn_rods_destroy",
    "pgn_rods_num_dofs", "pgn_rods_set_fixed", "pgn_rods_set_gravity",
    "pgn_rods_reset", "pgn_rods_set_x", "pgn_rods_get_x", "pgn_rods_get_energy",
    "pgn_rods_get_gradient", "pgn_rods_get_hessian_dense", "pgn_rods_get_frames",
    "pgn_rods_set_frames", "pgn_rods_update_kinematics", "pgn_rods_element_bt",
    "pgn_rods_newton", "pgn_csr_linear_solve", "pgn_csr_linear_solve_noalloc",
    "pgn_tetgen_sizes", "pgn_tetgen_read", "pgn_scene_load", "pgn_scene_destroy",
    "pgn_scene_sizes", "pgn_scene_mesh", "pgn_scene_boundary", "pgn_scene_solver",
    "pgn_scene_create_model",
]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_ucp = ctypes.POINTER(ctypes.c_ubyte)
_vp = ctypes.c_void_p


class LinearOpts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("max_iters", ctypes.c_int),
                ("reg_iters", ctypes.c_int), ("precond", ctypes.c_int),
                ("inner_tol", ctypes.c_double), ("solver", ctypes.c_int)]


class LinearResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_int),
                ("reg_used", ctypes.c_int), ("refinements", ctypes.c_int),
                ("rel_residual", ctypes.c_double)]


class NewtonOpts(ctypes.Structure):
    _fields_ = [("tolMaxError", ctypes.c_double), ("tolMinStep", ctypes.c_double),
                ("tolMinImpr", ctypes.c_double), ("minStepMaxCount", ctypes.c_int),
                ("minImprMaxCount", ctypes.c_int), ("numMaxIters", ctypes.c_int),
                ("maxStepSize", ctypes.c_double), ("trySteepest", ctypes.c_int),
                ("lsearch_numTry", ctypes.c_int), ("lSearch_alpha", ctypes.c_double),
                ("lsNumMaxIters", ctypes.c_int), ("lsTolMaxError", ctypes.c_double),
                ("lsNumRegTrials", ctypes.c_int), ("precond", ctypes.c_int),
                ("inner_tol", ctypes.c_double), ("adaptive_tol", ctypes.c_int),
                ("ls_noise_guard", ctypes.c_double), ("ew_gamma", ctypes.c_double),
                ("ew_max", ctypes.c_double), ("lsSolverType", ctypes.c_int),
                ("qpSolverType", ctypes.c_int), ("coarse_giveup", ctypes.c_int)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("lanes", ctypes.c_int),
                ("max_batch", ctypes.c_int), ("precision", ctypes.c_int),
                ("newton", NewtonOpts), ("group_by_size", ctypes.c_int),
                ("max_wait_s", ctypes.c_double), ("cohorts_per_lane", ctypes.c_int),
                ("solo_nodes", ctypes.c_long)]


class NewtonResult(ctypes.Structure):
    _fields_ = [("result", ctypes.c_int), ("iterations", ctypes.c_int),
                ("pcg_iterations", ctypes.c_int),
                ("line_search_bisections", ctypes.c_int),
                ("energy_evals", ctypes.c_int), ("hessian_evals", ctypes.c_int),
                ("objective", ctypes.c_double), ("optimality", ctypes.c_double),
                ("seconds", ctypes.c_double)]


_LIB = None


def lib():
    """Loads libpgn.so (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: build it with "
                               "`python -m physim_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.pgn_last_error.restype = ctypes.c_char_p
        L.pgn_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
        L.pgn_ctx_destroy.argtypes = [_vp]
        L.pgn_ctx_destroy.restype = None
        L.pgn_ctx_stream.argtypes = [_vp]
        L.pgn_ctx_stream.restype = _vp
        L.pgn_ctx_synchronize.argtypes = [_vp]
        L.pgn_model_create_tet.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_int, _ip,
                                           _ucp, _dp, _dp, _dp, ctypes.c_int,
                                           ctypes.POINTER(_vp)]
        L.pgn_model_destroy.argtypes = [_vp]
        L.pgn_model_destroy.restype = None
        L.pgn_model_num_dofs.argtypes = [_vp]
        L.pgn_model_num_elements.argtypes = [_vp]
        L.pgn_model_device_bytes.argtypes = [_vp]
        L.pgn_model_device_bytes.restype = ctypes.c_size_t
        L.pgn_model_set_fixed.argtypes = [_vp, _ucp]
        L.pgn_model_set_gravity.argtypes = [_vp, _dp, _dp, ctypes.c_int]
        L.pgn_model_set_forces.argtypes = [_vp, _dp, _dp, ctypes.c_int]
        L.pgn_model_step_load.argtypes = [_vp, _ip]
        L.pgn_model_is_loaded.argtypes = [_vp]
        L.pgn_model_reset.argtypes = [_vp]
        L.pgn_model_set_x.argtypes = [_vp, _dp]
        L.pgn_model_get_x.argtypes = [_vp, _dp]
        L.pgn_model_set_velocity.argtypes = [_vp, _dp]
        L.pgn_model_get_velocity.argtypes = [_vp, _dp]
        L.pgn_model_get_energy.argtypes = [_vp, _dp]
        L.pgn_model_get_gradient.argtypes = [_vp, _dp]
        L.pgn_model_hessian_nnz.argtypes = [_vp, ctypes.c_int, _lp]
        L.pgn_model_hessian_pattern.argtypes = [_vp, ctypes.c_int, _lp, _ip]
        L.pgn_model_get_hessian.argtypes = [_vp, ctypes.c_int, _dp]
        L.pgn_model_element_egh.argtypes = [_vp, ctypes.c_int, _dp, _dp, _dp]
        L.pgn_model_hessian_product.argtypes = [_vp, ctypes.c_int, _dp, _dp]
        L.pgn_model_precond_apply.argtypes = [_vp, ctypes.c_int, _dp, _dp]
        L.pgn_model_coarse_space.argtypes = [_vp, _ip, _ip, _ucp, _dp, _dp]
        L.pgn_rod_bt_eval.argtypes = [_vp, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int,
                                      _dp, _dp, _dp]
        L.pgn_linear_default_opts.argtypes = [ctypes.POINTER(LinearOpts)]
        L.pgn_linear_default_opts.restype = None
        L.pgn_linear_solve.argtypes = [_vp, _dp, _dp, ctypes.POINTER(LinearOpts),
                                       ctypes.POINTER(LinearResult)]
        L.pgn_csr_linear_solve.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _lp, _ip, _dp, _dp,
                                           _dp, ctypes.POINTER(LinearOpts),
                                           ctypes.POINTER(LinearResult)]
        L.pgn_csr_linear_solve_noalloc.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                                   ctypes.POINTER(LinearOpts),
                                                   ctypes.POINTER(LinearResult)]
        _cp = ctypes.c_char_p
        L.pgn_tetgen_sizes.argtypes = [_cp, _ip, _ip]
        L.pgn_tetgen_read.argtypes = [_cp, _dp, _ip]
        L.pgn_scene_load.argtypes = [_cp, ctypes.POINTER(_vp)]
        L.pgn_scene_destroy.argtypes = [_vp]
        L.pgn_scene_destroy.restype = None
        L.pgn_scene_sizes.argtypes = [_vp, _ip, _ip]
        L.pgn_scene_mesh.argtypes = [_vp, _dp, _dp, _ip, _ucp, _dp, _dp, _dp]
        L.pgn_scene_boundary.argtypes = [_vp, _ucp, _dp, _dp, _ip, _ip]
        L.pgn_scene_solver.argtypes = [_vp, ctypes.POINTER(NewtonOpts), _ip, _dp]
        L.pgn_scene_create_model.argtypes = [_vp, _vp, ctypes.c_int, ctypes.POINTER(_vp)]
        L.pgn_newton_default_opts.argtypes = [ctypes.POINTER(NewtonOpts)]
        L.pgn_newton_default_opts.restype = None
        L.pgn_newton_solve.argtypes = [_vp, ctypes.POINTER(NewtonOpts),
                                       ctypes.POINTER(NewtonResult)]
        L.pgn_dynamic_step.argtypes = [_vp, ctypes.c_double, ctypes.POINTER(NewtonOpts),
                                       ctypes.POINTER(NewtonResult)]
        L.pgn_time_kernel.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp]
        L.pgn_kernel_bytes.argtypes = [_vp, ctypes.c_int, _dp]
        L.pgn_host_bsr_pattern.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _lp, _lp, _ip]
        L.pgn_model_set_profiling.argtypes = [_vp, ctypes.c_int]
        L.pgn_model_create_shell.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp,
                                             _dp, _dp, _dp, ctypes.c_int, _ip, _dp,
                                             ctypes.c_int, ctypes.POINTER(_vp)]
        L.pgn_batch_create.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp)]
        L.pgn_batch_destroy.argtypes = [_vp]
        L.pgn_batch_destroy.restype = None
        L.pgn_batch_add_tet.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_int, _ip, _ucp, _dp,
                                        _dp, _dp, _ip]
        L.pgn_batch_add_shell.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, _dp,
                                          _dp, _dp, ctypes.c_int, _ip, _dp, _ip]
        L.pgn_batch_set_fixed.argtypes = [_vp, ctypes.c_int, _ucp]
        L.pgn_batch_set_gravity.argtypes = [_vp, ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.pgn_batch_set_forces.argtypes = [_vp, ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.pgn_batch_build.argtypes = [_vp]
        L.pgn_batch_num_instances.argtypes = [_vp]
        L.pgn_batch_device_bytes.argtypes = [_vp]
        L.pgn_batch_device_bytes.restype = ctypes.c_size_t
        L.pgn_batch_reset.argtypes = [_vp]
        L.pgn_batch_newton.argtypes = [_vp, ctypes.POINTER(NewtonOpts),
                                       ctypes.POINTER(NewtonResult)]
        L.pgn_batch_get_x.argtypes = [_vp, ctypes.c_int, _dp]
        L.pgn_batch_get_energy.argtypes = [_vp, _dp]
        L.pgn_server_default_opts.argtypes = [ctypes.POINTER(ServerOpts)]
        L.pgn_server_default_opts.restype = None
        L.pgn_server_create.argtypes = [ctypes.c_int, ctypes.POINTER(ServerOpts),
                                        ctypes.POINTER(_vp)]
        L.pgn_server_destroy.argtypes = [_vp]
        L.pgn_server_destroy.restype = None
        L.pgn_server_submit_tet.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_int, _ip, _ucp,
                                            _dp, _dp, _dp, _ucp, _dp, _dp, ctypes.c_int, _lp]
        L.pgn_server_submit_shell.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp,
                                              _dp, _dp, _dp, ctypes.c_int, _ip, _dp, _ucp, _dp,
                                              _dp, ctypes.c_int, _lp]
        L.pgn_server_poll.argtypes = [_vp, ctypes.c_long, _ip]
        L.pgn_server_wait.argtypes = [_vp, ctypes.c_long, _dp, ctypes.POINTER(NewtonResult),
                                      _dp, _dp]
        L.pgn_server_stats.argtypes = [_vp, _lp, _lp, ctypes.POINTER(ctypes.c_size_t)]
        L.pgn_rods_create.argtypes = [_vp, ctypes.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp,
                                      _dp, ctypes.c_int, ctypes.POINTER(_vp)]
        L.pgn_rods_destroy.argtypes = [_vp]
        L.pgn_rods_destroy.restype = None
        L.pgn_rods_num_dofs.argtypes = [_vp]
        L.pgn_rods_num_dofs.restype = ctypes.c_long
        L.pgn_rods_set_fixed.argtypes = [_vp, _ucp]
        L.pgn_rods_set_gravity.argtypes = [_vp, _dp, _dp, ctypes.c_int]
        L.pgn_rods_reset.argtypes = [_vp]
        L.pgn_rods_set_x.argtypes = [_vp, _dp]
        L.pgn_rods_get_x.argtypes = [_vp, _dp]
        L.pgn_rods_get_energy.argtypes = [_vp, _dp]
        L.pgn_rods_get_gradient.argtypes = [_vp, _dp]
        L.pgn_rods_get_hessian_dense.argtypes = [_vp, ctypes.c_int, _dp]
        L.pgn_rods_get_frames.argtypes = [_vp, _dp, _dp, _dp]
        L.pgn_rods_set_frames.argtypes = [_vp, _dp, _dp, _dp]
        L.pgn_rods_update_kinematics.argtypes = [_vp]
        L.pgn_rods_element_bt.argtypes = [_vp, _dp, _dp, _dp]
        L.pgn_rods_newton.argtypes = [_vp, ctypes.POINTER(NewtonOpts),
                                      ctypes.POINTER(NewtonResult)]
        L.pgn_model_get_profile.argtypes = [_vp, ctypes.c_int, _lp, _dp]
        _LIB = L
    return _LIB


def _newton_opts(kw):
    o = NewtonOpts()
    lib().pgn_newton_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        if k not in dict(NewtonOpts._fields_):
            raise AttributeError(f"pgn_newton_opts has no field {k}")
        setattr(o, k, v)
    return o


def _linear_opts(kw):
    o = LinearOpts()
    lib().pgn_linear_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        if k not in dict(LinearOpts._fields_):
            raise AttributeError(f"pgn_linear_opts has no field {k}")
        setattr(o, k, v)
    return o


def _linear_result(r):
    return {"status": r.status, "iterations": r.iterations, "reg_used": r.reg_used,
            "refinements": r.refinements, "rel_residual": r.rel_residual}


def csr_linear_solve(ctx, rowptr, col, val, b, x0=None, layout=None, **opts):
    """ILinearSolver::Solve(A, b, x) on a caller's CSR matrix (host arrays):
    pgn_csr_linear_solve.  layout: LAYOUT_LOWER (default) or LAYOUT_FULL.
    val is updated in place with the regularization shifts, as the reference
    updates mA.  Returns (x, result dict)."""
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    ci = np.ascontiguousarray(col, dtype=np.int32)
    if not (isinstance(val, np.ndarray) and val.dtype == np.float64 and val.flags.c_contiguous):
        raise TypeError("val must be a C-contiguous float64 array (it is updated in place)")
    n = rp.size - 1
    b = _f64(b).reshape(-1)
    x = np.zeros(n) if x0 is None else _f64(x0).reshape(-1).copy()
    o = _linear_opts(opts)
    r = LinearResult()
    _ok(lib().pgn_csr_linear_solve(ctx.handle, n, LAYOUT_LOWER if layout is None else layout,
                                   _p(rp, _lp), _p(ci, _ip), _p(val), _p(b), _p(x),
                                   ctypes.byref(o), ctypes.byref(r)))
    return x, _linear_result(r)


def csr_linear_solve_noalloc(ctx, n, rowptr_ptr, col_ptr, val_ptr, b_ptr, x_ptr, **opts):
    """pgn_csr_linear_solve_noalloc: every array is a device address (int) on
    the context's device, the matrix in full; x is updated on the device."""
    o = _linear_opts(opts)
    r = LinearResult()
    _ok(lib().pgn_csr_linear_solve_noalloc(ctx.handle, int(n), rowptr_ptr, col_ptr, val_ptr,
                                           b_ptr, x_ptr, ctypes.byref(o), ctypes.byref(r)))
    return _linear_result(r)


def read_tetgen(base):
    """TetGen base.node / base.ele -> (X (n, 3), T (m, 4)) (pgn_tetgen_read)."""
    b = str(base).encode()
    n, m = ctypes.c_int(), ctypes.c_int()
    _ok(lib().pgn_tetgen_sizes(b, ctypes.byref(n), ctypes.byref(m)))
    X = np.zeros((n.value, 3))
    T = np.zeros((m.value, 4), dtype=np.int32)
    _ok(lib().pgn_tetgen_read(b, _p(X), _p(T, _ip)))
    return X, T


class Scene:
    """A scene document in the reference's XML vocabulary (pgn_scene_load):
    mesh, materials, boundary conditions, problem and solver options."""

    def __init__(self, path):
        h = _vp()
        _ok(lib().pgn_scene_load(str(path).encode(), ctypes.byref(h)))
        self._h = h
        n, m = ctypes.c_int(), ctypes.c_int()
        _ok(lib().pgn_scene_sizes(h, ctypes.byref(n), ctypes.byref(m)))
        self.n, self.m = n.value, m.value
        self.X0 = np.zeros((self.n, 3))
        self.X = np.zeros((self.n, 3))
        self.T = np.zeros((self.m, 4), dtype=np.int32)
        self.material = np.zeros(self.m, dtype=np.uint8)
        self.lame1 = np.zeros(self.m)
        self.lame2 = np.zeros(self.m)
        self.density = np.zeros(self.m)
        _ok(lib().pgn_scene_mesh(h, _p(self.X0), _p(self.X), _p(self.T, _ip),
                                 _p(self.material, _ucp), _p(self.lame1), _p(self.lame2),
                                 _p(self.density)))
        self.fixed = np.zeros((self.n, 3), dtype=np.uint8)
        gi, ge = np.zeros(3), np.zeros(3)
        st, hg = ctypes.c_int(), ctypes.c_int()
        _ok(lib().pgn_scene_boundary(h, _p(self.fixed, _ucp), _p(gi), _p(ge), ctypes.byref(st),
                                     ctypes.byref(hg)))
        self.gravity = ge if hg.value else None
        self.gravity_ini = gi
        self.stages = st.value
        o = NewtonOpts()
        dyn, dt = ctypes.c_int(), ctypes.c_double()
        _ok(lib().pgn_scene_solver(h, ctypes.byref(o), ctypes.byref(dyn), ctypes.byref(dt)))
        self.opts = {k: getattr(o, k) for k, _ in NewtonOpts._fields_}
        self.dynamic, self.dt = bool(dyn.value), dt.value

    def close(self):
        if getattr(self, "_h", None):
            lib().pgn_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def create_model(self, ctx, precision=MIXED):
        """pgn_scene_create_model: a TetModel with the scene's boundary
        conditions applied."""
        h = _vp()
        _ok(lib().pgn_scene_create_model(ctx.handle, self._h, precision, ctypes.byref(h)))
        return TetModel._adopt_handle(ctx, h, self.n, self.m, precision)


class PgnError(RuntimeError):
    pass


def _ok(rc):
    if rc != 0:
        raise PgnError(f"libpgn error {rc}: {lib().pgn_last_error().decode()}")


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """A libpgn context (one device, one HIP stream).  Models and rod batches
    created on it register here, and close() destroys them before the
    context: the library requires every model to go before its context."""

    def __init__(self, device=0):
        h = _vp()
        _ok(lib().pgn_ctx_create(device, ctypes.byref(h)))
        self._h = h
        self._children = weakref.WeakSet()

    @property
    def handle(self):
        return self._h

    def _adopt(self, child):
        self._children.add(child)

    def stream(self):
        return lib().pgn_ctx_stream(self._h)

    def synchronize(self):
        _ok(lib().pgn_ctx_synchronize(self._h))

    def close(self):
        if getattr(self, "_h", None):
            for c in list(getattr(self, "_children", ())):
                c.close()
            lib().pgn_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class TetModel:
    """A device-resident Tet4 FEM model (ISimulable + IOptimProblem)."""

    def __init__(self, ctx, X0, T, material, lame1, lame2, density, precision=FP64):
        X0 = _f64(X0).reshape(-1)
        T = np.ascontiguousarray(T, dtype=np.int32).reshape(-1)
        m = T.size // 4
        mat = np.ascontiguousarray(np.broadcast_to(material, (m,)), dtype=np.uint8)
        l1 = _f64(np.broadcast_to(lame1, (m,)))
        l2 = _f64(np.broadcast_to(lame2, (m,)))
        rho = _f64(np.broadcast_to(density, (m,)))
        self.ctx = ctx
        self.n = X0.size // 3
        self.m = m
        self.N = 3 * self.n
        self.precision = precision
        h = _vp()
        _ok(lib().pgn_model_create_tet(ctx.handle, self.n, _p(X0), m, _p(T, _ip),
                                       _p(mat, _ucp), _p(l1), _p(l2), _p(rho),
                                       precision, ctypes.byref(h)))
        self._h = h
        ctx._adopt(self)

    @classmethod
    def _adopt_handle(cls, ctx, h, n, m, precision):
        """Wraps a model the library created (e.g. pgn_scene_create_model)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.n, self.m, self.N = n, m, 3 * n
        self.precision = precision
        self._h = h
        ctx._adopt(self)
        return self

    def close(self):
        if getattr(self, "_h", None):
            lib().pgn_model_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def device_bytes(self):
        return lib().pgn_model_device_bytes(self._h)

    def set_fixed(self, mask):
        mk = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
        assert mk.size == self.N
        _ok(lib().pgn_model_set_fixed(self._h, _p(mk, _ucp)))

    def set_gravity(self, g_end, stages, g_ini=(0.0, 0.0, 0.0)):
        _ok(lib().pgn_model_set_gravity(self._h, _p(_f64(g_ini)), _p(_f64(g_end)),
                                        int(stages)))

    def set_forces(self, f_end, stages, f_ini=None):
        f_end = _f64(f_end).reshape(-1)
        f_ini = np.zeros_like(f_end) if f_ini is None else _f64(f_ini).reshape(-1)
        _ok(lib().pgn_model_set_forces(self._h, _p(f_ini), _p(f_end), int(stages)))

    def step_load(self):
        s = ctypes.c_int()
        _ok(lib().pgn_model_step_load(self._h, ctypes.byref(s)))
        return bool(s.value)

    def is_loaded(self):
        return bool(lib().pgn_model_is_loaded(self._h))

    def reset(self):
        _ok(lib().pgn_model_reset(self._h))

    def set_x(self, x):
        x = _f64(x).reshape(-1)
        _ok(lib().pgn_model_set_x(self._h, _p(x)))

    def get_x(self):
        x = np.zeros(self.N)
        _ok(lib().pgn_model_get_x(self._h, _p(x)))
        return x

    def set_velocity(self, v):
        v = _f64(v).reshape(-1)
        _ok(lib().pgn_model_set_velocity(self._h, _p(v)))

    def get_velocity(self):
        v = np.zeros(self.N)
        _ok(lib().pgn_model_get_velocity(self._h, _p(v)))
        return v

    def energy(self):
        e = ctypes.c_double()
        _ok(lib().pgn_model_get_energy(self._h, ctypes.byref(e)))
        return e.value

    def gradient(self):
        g = np.zeros(self.N)
        _ok(lib().pgn_model_get_gradient(self._h, _p(g)))
        return g

    def hessian(self, layout=LAYOUT_LOWER):
        nnz = ctypes.c_long()
        _ok(lib().pgn_model_hessian_nnz(self._h, layout, ctypes.byref(nnz)))
        rows = self.n + 1 if layout == LAYOUT_BSR3 else self.N + 1
        rp = np.zeros(rows, dtype=np.int64)
        ci = np.zeros(nnz.value, dtype=np.int32)
        _ok(lib().pgn_model_hessian_pattern(self._h, layout, _p(rp, _lp), _p(ci, _ip)))
        nv = nnz.value * (9 if layout == LAYOUT_BSR3 else 1)
        v = np.zeros(nv)
        _ok(lib().pgn_model_get_hessian(self._h, layout, _p(v)))
        return rp, ci, v

    def element_egh(self, precision=None):
        prec = self.precision if precision is None else precision
        E = np.zeros(self.m)
        g = np.zeros((self.m, 12))
        H = np.zeros((self.m, 12, 12))
        _ok(lib().pgn_model_element_egh(self._h, prec, _p(E), _p(g), _p(H)))
        return E, g, H

    def hessian_product(self, p, matrix_free=False):
        """y = H p at the current x (pgn_model_hessian_product)."""
        p = _f64(p).reshape(-1)
        y = np.zeros(self.N)
        _ok(lib().pgn_model_hessian_product(self._h, 1 if matrix_free else 0, _p(p), _p(y)))
        return y

    def precond_apply(self, r, precond):
        """z = M^-1 r with the preconditioner of the Hessian at the current x
        (pgn_model_precond_apply: 0 Jacobi, 1 block Jacobi, 2 two-level)."""
        r = _f64(r).reshape(-1)
        z = np.zeros(self.N)
        _ok(lib().pgn_model_precond_apply(self._h, int(precond), _p(r), _p(z)))
        return z

    def coarse_space(self, matrices=False):
        """The two-level preconditioner's aggregation (pgn_model_coarse_space):
        (agg_of, cmask) and, with matrices, also (Ac, Ac^-1) of the Hessian at
        the current x."""
        na = ctypes.c_int(0)
        _ok(lib().pgn_model_coarse_space(self._h, ctypes.byref(na), None, None, None, None))
        k = na.value
        agg = np.zeros(self.N // 3, dtype=np.int32)
        cm = np.zeros(max(k, 1), dtype=np.uint8)
        if not matrices:
            _ok(lib().pgn_model_coarse_space(self._h, ctypes.byref(na), _p(agg, _ip),
                                             _p(cm, _ucp), None, None))
            return agg, cm[:k]
        Ac = np.zeros((6 * k, 6 * k))
        Ai = np.zeros((6 * k, 6 * k))
        _ok(lib().pgn_model_coarse_space(self._h, ctypes.byref(na), _p(agg, _ip), _p(cm, _ucp),
                                         _p(Ac), _p(Ai)))
        return agg, cm[:k], Ac, Ai

    def linear_solve(self, b, x0=None, **kw):
        o = _linear_opts(kw)
        b = _f64(b).reshape(-1)
        x = np.zeros(self.N) if x0 is None else _f64(x0).reshape(-1).copy()
        r = LinearResult()
        _ok(lib().pgn_linear_solve(self._h, _p(b), _p(x), ctypes.byref(o), ctypes.byref(r)))
        return x, {"status": r.status, "iterations": r.iterations,
                   "reg_used": r.reg_used, "refinements": r.refinements,
                   "rel_residual": r.rel_residual}

    def newton(self, **kw):
        o = _newton_opts(kw)
        r = NewtonResult()
        _ok(lib().pgn_newton_solve(self._h, ctypes.byref(o), ctypes.byref(r)))
        return {"result": OS_NAMES[r.result], "iterations": r.iterations,
                "pcg_iterations": r.pcg_iterations,
                "bisections": r.line_search_bisections,
                "energy_evals": r.energy_evals, "hessian_evals": r.hessian_evals,
                "objective": r.objective, "optimality": r.optimality,
                "seconds": r.seconds}

    def dynamic_step(self, dt, **kw):
        """One backward-Euler time step of length dt (OptimProblem_BasicDynamic
        minimised by SolveFull); updates x and the velocity."""
        o = _newton_opts(kw)
        r = NewtonResult()
        _ok(lib().pgn_dynamic_step(self._h, float(dt), ctypes.byref(o), ctypes.byref(r)))
        return {"result": OS_NAMES[r.result], "iterations": r.iterations,
                "pcg_iterations": r.pcg_iterations,
                "bisections": r.line_search_bisections,
                "energy_evals": r.energy_evals, "hessian_evals": r.hessian_evals,
                "objective": r.objective, "optimality": r.optimality,
                "seconds": r.seconds}

    def time_kernel(self, which, reps=20):
        ms = ctypes.c_double()
        _ok(lib().pgn_time_kernel(self._h, which, reps, ctypes.byref(ms)))
        return ms.value

    def set_profiling(self, on=True):
        _ok(lib().pgn_model_set_profiling(self._h, 1 if on else 0))

    def profile(self, which):
        """(launches, total ms) of PCG kernel `which` (0 = SpMV, 1 = update);
        2-4: (lanes per row in use, SpMV us at 4 / 8 / 16 lanes); 5: (bytes per
        SELL-64 block column, 0 without the SELL-64 copy; 0); 6-8: (tuned tet
        element tile, its us) of the Hessian, FP64 E+g and E+g+H launches;
        9 + t: (t, the Hessian launch's us with tile t)."""
        c = ctypes.c_long()
        ms = ctypes.c_double()
        _ok(lib().pgn_model_get_profile(self._h, which, ctypes.byref(c), ctypes.byref(ms)))
        return c.value, ms.value

    def kernel_bytes(self, which):
        b = ctypes.c_double()
        _ok(lib().pgn_kernel_bytes(self._h, which, ctypes.byref(b)))
        return b.value


class ShellModel(TetModel):
    """A device-resident thin shell (Simulable_ThinShell): StVK membrane
    triangles + discrete-shell hinges.  Every TetModel method applies; the
    element outputs list triangles first, then hinges."""

    def __init__(self, ctx, X0, F, lame1, lame2, thickness, density, hinges, bending_k,
                 precision=FP64):
        X0 = _f64(X0).reshape(-1)
        F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1)
        Hn = np.ascontiguousarray(hinges, dtype=np.int32).reshape(-1)
        nt, nh = F.size // 3, Hn.size // 4
        b = lambda a, k: _f64(np.broadcast_to(a, (k,)))
        self.ctx = ctx
        self.n = X0.size // 3
        self.m = nt + nh
        self.n_tris, self.n_hinges = nt, nh
        self.N = 3 * self.n
        self.precision = precision
        h = _vp()
        _ok(lib().pgn_model_create_shell(ctx.handle, self.n, _p(X0), nt, _p(F, _ip),
                                         _p(b(lame1, nt)), _p(b(lame2, nt)),
                                         _p(b(thickness, nt)), _p(b(density, nt)), nh,
                                         _p(Hn, _ip), _p(b(bending_k, nh)), precision,
                                         ctypes.byref(h)))
        self._h = h
        ctx._adopt(self)


def _result_dict(r):
    return {"result": OS_NAMES[r.result], "iterations": r.iterations,
            "pcg_iterations": r.pcg_iterations, "bisections": r.line_search_bisections,
            "energy_evals": r.energy_evals, "hessian_evals": r.hessian_evals,
            "objective": r.objective, "optimality": r.optimality, "seconds": r.seconds}


class Batch:
    """A fused batch of independent tet / thin-shell solves (pgn_batch):
    stage instances with add_tet / add_shell (+ boundary conditions), then
    build() uploads them as one device model."""

    def __init__(self, ctx, precision=FP64):
        h = _vp()
        _ok(lib().pgn_batch_create(ctx.handle, precision, ctypes.byref(h)))
        self._h = h
        self.ctx = ctx
        self.n_nodes = []
        ctx._adopt(self)

    def close(self):
        if getattr(self, "_h", None):
            lib().pgn_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def add_tet(self, X0, T, material, lame1, lame2, density):
        X0 = _f64(X0).reshape(-1)
        T = np.ascontiguousarray(T, dtype=np.int32).reshape(-1)
        m = T.size // 4
        b = lambda a: _f64(np.broadcast_to(a, (m,)))
        mat = np.ascontiguousarray(np.broadcast_to(material, (m,)), dtype=np.uint8)
        k = ctypes.c_int()
        _ok(lib().pgn_batch_add_tet(self._h, X0.size // 3, _p(X0), m, _p(T, _ip),
                                    _p(mat, _ucp), _p(b(lame1)), _p(b(lame2)), _p(b(density)),
                                    ctypes.byref(k)))
        self.n_nodes.append(X0.size // 3)
        return k.value

    def add_shell(self, X0, F, lame1, lame2, thickness, density, hinges, bending_k):
        X0 = _f64(X0).reshape(-1)
        F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1)
        Hn = np.ascontiguousarray(hinges, dtype=np.int32).reshape(-1)
        nt, nh = F.size // 3, Hn.size // 4
        b = lambda a, c: _f64(np.broadcast_to(a, (c,)))
        k = ctypes.c_int()
        _ok(lib().pgn_batch_add_shell(self._h, X0.size // 3, _p(X0), nt, _p(F, _ip),
                                      _p(b(lame1, nt)), _p(b(lame2, nt)), _p(b(thickness, nt)),
                                      _p(b(density, nt)), nh, _p(Hn, _ip), _p(b(bending_k, nh)),
                                      ctypes.byref(k)))
        self.n_nodes.append(X0.size // 3)
        return k.value

    def add_scene(self, scene):
        """Stage a scenes.Scene (tets) or scenes.ShellScene with its fixation
        and loads; returns the instance index."""
        if hasattr(scene, "hinges"):
            k = self.add_shell(scene.X0, scene.F, scene.lame1, scene.lame2, scene.thickness,
                               scene.density, scene.hinges, scene.bending_k)
        else:
            k = self.add_tet(scene.X0, scene.T, scene.material, scene.lame1, scene.lame2,
                             scene.density)
        mk = np.ascontiguousarray(scene.fixed, dtype=np.uint8).reshape(-1)
        _ok(lib().pgn_batch_set_fixed(self._h, k, _p(mk, _ucp)))
        _ok(lib().pgn_batch_set_gravity(self._h, k, _p(_f64((0.0, 0.0, 0.0))),
                                        _p(_f64(scene.gravity)), int(scene.stages)))
        if getattr(scene, "forces", None) is not None:
            fe = _f64(scene.forces).reshape(-1)
            _ok(lib().pgn_batch_set_forces(self._h, k, _p(np.zeros_like(fe)), _p(fe),
                                           int(scene.stages)))
        return k

    def build(self):
        _ok(lib().pgn_batch_build(self._h))

    @property
    def K(self):
        return lib().pgn_batch_num_instances(self._h)

    def device_bytes(self):
        return lib().pgn_batch_device_bytes(self._h)

    def reset(self):
        _ok(lib().pgn_batch_reset(self._h))

    def newton(self, **kw):
        o = _newton_opts(kw)
        res = (NewtonResult * self.K)()
        _ok(lib().pgn_batch_newton(self._h, ctypes.byref(o), res))
        return [_result_dict(r) for r in res]

    def get_x(self, k):
        x = np.zeros(3 * self.n_nodes[k])
        _ok(lib().pgn_batch_get_x(self._h, k, _p(x)))
        return x.reshape(-1, 3)

    def energy(self):
        e = np.zeros(self.K)
        _ok(lib().pgn_batch_get_energy(self._h, _p(e)))
        return e


class Server:
    """The request loop (pgn_server): submit scenes from any thread, wait for
    their results.  Continuous batching: each lane keeps up to
    cohorts_per_lane fused batches resident on its own stream and advances
    them one Newton iteration at a time; between iterations it admits queued
    requests (by estimated device footprint against hbm_budget_bytes, similar
    footprints cheapest first under group_by_size, else arrival order) and
    hands back each request as soon as its own solve ends."""

    def __init__(self, device=0, hbm_budget_bytes=64 << 30, lanes=4, max_batch=64,
                 precision=MIXED, group_by_size=1, max_wait_s=2.0, cohorts_per_lane=4,
                 solo_nodes=32768, **newton):
        o = ServerOpts()
        lib().pgn_server_default_opts(ctypes.byref(o))
        o.hbm_budget_bytes = int(hbm_budget_bytes)
        o.lanes = lanes
        o.max_batch = max_batch
        o.precision = precision
        o.group_by_size = int(group_by_size)
        o.max_wait_s = float(max_wait_s)
        o.cohorts_per_lane = int(cohorts_per_lane)
        o.solo_nodes = int(solo_nodes)
        o.newton = _newton_opts(newton)
        h = _vp()
        _ok(lib().pgn_server_create(device, ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._n = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().pgn_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, scene):
        """Queue a scenes.Scene (tets) or scenes.ShellScene; returns a ticket."""
        t = ctypes.c_long()
        X0 = _f64(scene.X0).reshape(-1)
        n = X0.size // 3
        fixed = np.ascontiguousarray(scene.fixed, dtype=np.uint8).reshape(-1)
        grav = _f64(scene.gravity)
        forces = getattr(scene, "forces", None)
        fp = _p(_f64(forces).reshape(-1)) if forces is not None else None
        if hasattr(scene, "hinges"):
            F = np.ascontiguousarray(scene.F, dtype=np.int32).reshape(-1)
            Hn = np.ascontiguousarray(scene.hinges, dtype=np.int32).reshape(-1)
            nt, nh = F.size // 3, Hn.size // 4
            b = lambda a, c: _f64(np.broadcast_to(a, (c,)))
            _ok(lib().pgn_server_submit_shell(
                self._h, n, _p(X0), nt, _p(F, _ip), _p(b(scene.lame1, nt)),
                _p(b(scene.lame2, nt)), _p(b(scene.thickness, nt)), _p(b(scene.density, nt)),
                nh, _p(Hn, _ip), _p(b(scene.bending_k, nh)), _p(fixed, _ucp), _p(grav), fp,
                int(scene.stages), ctypes.byref(t)))
        else:
            T = np.ascontiguousarray(scene.T, dtype=np.int32).reshape(-1)
            m = T.size // 4
            b = lambda a: _f64(np.broadcast_to(a, (m,)))
            mat = np.ascontiguousarray(np.broadcast_to(scene.material, (m,)), dtype=np.uint8)
            _ok(lib().pgn_server_submit_tet(
                self._h, n, _p(X0), m, _p(T, _ip), _p(mat, _ucp), _p(b(scene.lame1)),
                _p(b(scene.lame2)), _p(b(scene.density)), _p(fixed, _ucp), _p(grav), fp,
                int(scene.stages), ctypes.byref(t)))
        self._n[t.value] = n
        return t.value

    def poll(self, ticket):
        d = ctypes.c_int()
        _ok(lib().pgn_server_poll(self._h, ticket, ctypes.byref(d)))
        return bool(d.value)

    def wait(self, ticket):
        """(x (n, 3), result dict, latency s, queueing s)"""
        x = np.zeros(3 * self._n.pop(ticket))
        r = NewtonResult()
        lat = ctypes.c_double()
        q = ctypes.c_double()
        _ok(lib().pgn_server_wait(self._h, ticket, _p(x), ctypes.byref(r), ctypes.byref(lat),
                                  ctypes.byref(q)))
        return x.reshape(-1, 3), _result_dict(r), lat.value, q.value

    def stats(self):
        b = ctypes.c_long()
        sv = ctypes.c_long()
        u = ctypes.c_size_t()
        _ok(lib().pgn_server_stats(self._h, ctypes.byref(b), ctypes.byref(sv), ctypes.byref(u)))
        return {"batches": b.value, "solved": sv.value, "bytes_in_use": u.value}


def host_bsr_pattern(n_nodes, T):
    """BSR3 pattern the library builds for a tet mesh (host logic only)."""
    T = np.ascontiguousarray(T, dtype=np.int32).reshape(-1)
    m = T.size // 4
    nb = ctypes.c_long()
    _ok(lib().pgn_host_bsr_pattern(n_nodes, m, _p(T, _ip), ctypes.byref(nb), None, None))
    rp = np.zeros(n_nodes + 1, dtype=np.int64)
    col = np.zeros(nb.value, dtype=np.int32)
    _ok(lib().pgn_host_bsr_pattern(n_nodes, m, _p(T, _ip), ctypes.byref(nb),
                                   _p(rp, _lp), _p(col, _ip)))
    return rp, col


def rod_bt_eval(ctx, prm, e1, e2, theta, precision=FP64):
    """Bend-twist elements on given states through the device kernel
    (pgn_rod_bt_eval): prm (k, 27), e1 / e2 (k, 3), theta (k, 2) ->
    E (k), g (k, 8), H (k, 8, 8)."""
    prm = _f64(prm).reshape(-1, 27)
    k = prm.shape[0]
    E = np.zeros(k)
    g = np.zeros((k, 8))
    H = np.zeros((k, 8, 8))
    _ok(lib().pgn_rod_bt_eval(ctx.handle, k, _p(prm), _p(_f64(e1).reshape(-1)),
                              _p(_f64(e2).reshape(-1)), _p(_f64(theta).reshape(-1)),
                              precision, _p(E), _p(g), _p(H)))
    return E, g, H


class Rods:
    """A batch of discrete elastic rods on the device (Simulable_DER)."""

    def __init__(self, ctx, rods, frame0_normals, young, shear, radius_h, radius_w,
                 density, precision=FP64):
        nverts = np.array([len(r) for r in rods], dtype=np.int32)
        X0 = _f64(np.concatenate([np.asarray(r).reshape(-1, 3) for r in rods]))
        R = len(rods)
        b = lambda a: _f64(np.broadcast_to(a, (R,)))
        self.nverts = nverts
        self.R = R
        self.ctx = ctx
        h = _vp()
        _ok(lib().pgn_rods_create(ctx.handle, R, _p(nverts, _ip), _p(X0),
                                  _p(_f64(np.broadcast_to(frame0_normals, (R, 3)))),
                                  _p(b(young)), _p(b(shear)), _p(b(radius_h)),
                                  _p(b(radius_w)), _p(b(density)), precision,
                                  ctypes.byref(h)))
        self._h = h
        ctx._adopt(self)
        self.N = lib().pgn_rods_num_dofs(self._h)
        self.offsets = np.concatenate([[0], np.cumsum(4 * nverts - 1)])

    def close(self):
        if getattr(self, "_h", None):
            lib().pgn_rods_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_fixed(self, mask):
        mk = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
        assert mk.size == self.N
        _ok(lib().pgn_rods_set_fixed(self._h, _p(mk, _ucp)))

    def set_gravity(self, g_end, stages, g_ini=(0.0, 0.0, 0.0)):
        _ok(lib().pgn_rods_set_gravity(self._h, _p(_f64(g_ini)), _p(_f64(g_end)), int(stages)))

    def reset(self):
        _ok(lib().pgn_rods_reset(self._h))

    def set_x(self, x):
        _ok(lib().pgn_rods_set_x(self._h, _p(_f64(x).reshape(-1))))

    def get_x(self):
        x = np.zeros(self.N)
        _ok(lib().pgn_rods_get_x(self._h, _p(x)))
        return x

    def energy(self):
        e = np.zeros(self.R)
        _ok(lib().pgn_rods_get_energy(self._h, _p(e)))
        return e

    def gradient(self):
        g = np.zeros(self.N)
        _ok(lib().pgn_rods_get_gradient(self._h, _p(g)))
        return g

    def hessian_dense(self, rod):
        n = 4 * int(self.nverts[rod]) - 1
        H = np.zeros((n, n))
        _ok(lib().pgn_rods_get_hessian_dense(self._h, rod, _p(H)))
        return H

    def frames(self):
        ne = int((self.nverts - 1).sum())
        t = np.zeros((ne, 3))
        n = np.zeros((ne, 3))
        rt = np.zeros(int(self.nverts.sum()))
        _ok(lib().pgn_rods_get_frames(self._h, _p(t), _p(n), _p(rt)))
        return t, n, rt

    def set_frames(self, tan, nor, rt):
        _ok(lib().pgn_rods_set_frames(self._h, _p(_f64(tan).reshape(-1)),
                                      _p(_f64(nor).reshape(-1)), _p(_f64(rt).reshape(-1))))

    def update_kinematics(self):
        _ok(lib().pgn_rods_update_kinematics(self._h))

    def element_bt(self):
        nv = int(self.nverts.sum())
        E = np.zeros(nv)
        g = np.zeros((nv, 8))
        H = np.zeros((nv, 8, 8))
        _ok(lib().pgn_rods_element_bt(self._h, _p(E), _p(g), _p(H)))
        return E, g, H

    def newton(self, **kw):
        o = _newton_opts(kw)
        res = (NewtonResult * self.R)()
        _ok(lib().pgn_rods_newton(self._h, ctypes.byref(o), res))
        return [{"result": OS_NAMES[r.result], "iterations": r.iterations,
                 "bisections": r.line_search_bisections, "energy_evals": r.energy_evals,
                 "hessian_evals": r.hessian_evals, "objective": r.objective,
                 "optimality": r.optimality, "seconds": r.seconds} for r in res]
