"""ctypes wrapper for the FP64 CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline -- never as the product path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

CONH, STVK = 0, 1
OS_NAMES = ["ONGOING", "SUCCESS", "MINSTEP", "MINIMPR", "FAILURE", "MAXITER",
            "NONDESC"]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_ucp = ctypes.POINTER(ctypes.c_ubyte)


class NewtonOpts(ctypes.Structure):
    _fields_ = [("tolMaxError", ctypes.c_double),
                ("tolMinStep", ctypes.c_double),
                ("tolMinImpr", ctypes.c_double),
                ("minStepMaxCount", ctypes.c_int),
                ("minImprMaxCount", ctypes.c_int),
                ("numMaxIters", ctypes.c_int),
                ("maxStepSize", ctypes.c_double),
                ("trySteepest", ctypes.c_int),
                ("lsearch_numTry", ctypes.c_int),
                ("lSearch_alpha", ctypes.c_double),
                ("lsNumMaxIters", ctypes.c_int),
                ("lsTolMaxError", ctypes.c_double),
                ("lsNumRegTrials", ctypes.c_int),
                ("ls_noise_guard", ctypes.c_double),
                ("ls_solver", ctypes.c_int),
                ("qpSolverType", ctypes.c_int)]


class NewtonResult(ctypes.Structure):
    _fields_ = [("result", ctypes.c_int),
                ("iterations", ctypes.c_int),
                ("cg_iterations", ctypes.c_int),
                ("line_search_bisections", ctypes.c_int),
                ("objective", ctypes.c_double),
                ("optimality", ctypes.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_psi.restype = ctypes.c_double
        L.orc_psi.argtypes = [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double]
        for f in ("orc_pk1", "orc_dpdf"):
            getattr(L, f).argtypes = [ctypes.c_int, _dp, ctypes.c_double,
                                      ctypes.c_double, _dp]
        L.orc_tet_rest.argtypes = [_dp, _dp, _dp]
        L.orc_tet_egh.restype = ctypes.c_int
        L.orc_tet_egh.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.orc_tet_grid.argtypes = [_dp, _ip, _dp, _ip]
        L.orc_model_create.restype = ctypes.c_void_p
        L.orc_model_create.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, _ucp,
                                       _dp, _dp, _dp]
        L.orc_model_destroy.argtypes = [ctypes.c_void_p]
        L.orc_model_set_fixed.argtypes = [ctypes.c_void_p, _ucp]
        L.orc_model_set_gravity.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.orc_model_set_forces.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.orc_model_step_load.argtypes = [ctypes.c_void_p]
        L.orc_model_fully_loaded.argtypes = [ctypes.c_void_p]
        L.orc_model_set_x.argtypes = [ctypes.c_void_p, _dp]
        L.orc_model_get_x.argtypes = [ctypes.c_void_p, _dp]
        L.orc_model_get_mass.argtypes = [ctypes.c_void_p, _dp]
        L.orc_model_energy.restype = ctypes.c_double
        L.orc_model_energy.argtypes = [ctypes.c_void_p]
        L.orc_model_gradient.argtypes = [ctypes.c_void_p, _dp]
        L.orc_model_nnz_lower.restype = ctypes.c_long
        L.orc_model_nnz_lower.argtypes = [ctypes.c_void_p]
        L.orc_model_hessian_lower.argtypes = [ctypes.c_void_p, _lp, _ip, _dp]
        L.orc_model_element_egh.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_lbfgs_direction.restype = None
        L.orc_lbfgs_direction.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_cg.restype = ctypes.c_int
        L.orc_cg.argtypes = [ctypes.c_int, _lp, _ip, _dp, _dp, _dp,
                             ctypes.c_double, ctypes.c_int, _dp]
        L.orc_linear_solve.restype = ctypes.c_int
        L.orc_linear_solve.argtypes = [ctypes.c_int, _lp, _ip, _dp, _dp, _dp,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       _ip, _ip]
        L.orc_newton_default_opts.argtypes = [ctypes.POINTER(NewtonOpts)]
        L.orc_newton_solve.argtypes = [ctypes.c_void_p, ctypes.POINTER(NewtonOpts),
                                       ctypes.POINTER(NewtonResult)]
        L.orc_model_create_mixed.restype = ctypes.c_void_p
        L.orc_model_create_mixed.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, _ucp,
                                             _dp, _dp, _dp, _dp]
        for f in ("orc_hinge_theta", "orc_hinge_shape"):
            getattr(L, f).restype = ctypes.c_double
            getattr(L, f).argtypes = [_dp]
        L.orc_hinge_dtheta.argtypes = [_dp, _dp]
        L.orc_hinge_egh.restype = ctypes.c_double
        L.orc_hinge_egh.argtypes = [_dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    _dp, _dp]
        L.orc_membrane_rest.argtypes = [_dp, _dp, _dp]
        L.orc_membrane_egh.restype = ctypes.c_double
        L.orc_membrane_egh.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, _dp, _dp]
        L.orc_mem_psi.restype = ctypes.c_double
        L.orc_mem_psi.argtypes = [_dp, ctypes.c_double, ctypes.c_double]
        L.orc_mem_pk1.argtypes = [_dp, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_mem_dpdf.argtypes = [_dp, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_model_energy_scale.restype = ctypes.c_double
        L.orc_model_energy_scale.argtypes = [ctypes.c_void_p]
        L.orc_rod_bt.restype = ctypes.c_double
        L.orc_rod_bt.argtypes = [_dp, _dp, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_rod_stretch.restype = ctypes.c_double
        L.orc_rod_stretch.argtypes = [_dp, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_reference_twist_change.restype = ctypes.c_double
        L.orc_reference_twist_change.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_double]
        L.orc_rod_create.restype = ctypes.c_void_p
        L.orc_rod_create.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_rod_destroy.argtypes = [ctypes.c_void_p]
        L.orc_rod_num_dofs.argtypes = [ctypes.c_void_p]
        L.orc_rod_set_fixed.argtypes = [ctypes.c_void_p, _ucp]
        L.orc_rod_set_gravity.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
        L.orc_rod_step_load.argtypes = [ctypes.c_void_p]
        L.orc_rod_fully_loaded.argtypes = [ctypes.c_void_p]
        L.orc_rod_get_x.argtypes = [ctypes.c_void_p, _dp]
        L.orc_rod_set_x.argtypes = [ctypes.c_void_p, _dp]
        L.orc_rod_get_frames.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_rod_energy.restype = ctypes.c_double
        L.orc_rod_energy.argtypes = [ctypes.c_void_p]
        L.orc_rod_gradient.argtypes = [ctypes.c_void_p, _dp]
        L.orc_rod_hessian.argtypes = [ctypes.c_void_p, _dp]
        L.orc_rod_update_kinematics.argtypes = [ctypes.c_void_p]
        L.orc_rod_newton.argtypes = [ctypes.c_void_p, ctypes.POINTER(NewtonOpts),
                                     ctypes.POINTER(NewtonResult)]
        L.orc_dynamic_create.restype = ctypes.c_void_p
        L.orc_dynamic_create.argtypes = [ctypes.c_void_p]
        L.orc_dynamic_destroy.argtypes = [ctypes.c_void_p]
        L.orc_dynamic_set_velocity.argtypes = [ctypes.c_void_p, _dp]
        L.orc_dynamic_get_velocity.argtypes = [ctypes.c_void_p, _dp]
        L.orc_dynamic_begin.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_dynamic_energy.restype = ctypes.c_double
        L.orc_dynamic_energy.argtypes = [ctypes.c_void_p]
        L.orc_dynamic_gradient.argtypes = [ctypes.c_void_p, _dp]
        L.orc_dynamic_hessian_lower.argtypes = [ctypes.c_void_p, _lp, _ip, _dp]
        L.orc_dynamic_solve.argtypes = [ctypes.c_void_p, ctypes.POINTER(NewtonOpts),
                                        ctypes.POINTER(NewtonResult)]
        L.orc_dynamic_end.argtypes = [ctypes.c_void_p]
        L.orc_dynamic_step.argtypes = [ctypes.c_void_p, ctypes.c_double,
                                       ctypes.POINTER(NewtonOpts), ctypes.POINTER(NewtonResult)]
        L.orc_dense_ldlt_solve.restype = ctypes.c_int
        L.orc_dense_ldlt_solve.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double,
                                           ctypes.c_int, _ip]
        L.orc_sparse_ldlt_solve.restype = ctypes.c_int
        L.orc_sparse_ldlt_solve.argtypes = [ctypes.c_int, _lp, _ip, _dp, _dp, _dp,
                                            ctypes.c_double, ctypes.c_int, _ip]
        _LIB = L
    return _LIB


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---- element-level --------------------------------------------------------

def psi(kind, F, lame1, lame2):
    Fv = f64(np.asarray(F).T.reshape(-1))
    return lib().orc_psi(kind, _p(Fv), lame1, lame2)


def pk1(kind, F, lame1, lame2):
    Fv = f64(np.asarray(F).T.reshape(-1))
    out = np.zeros(9)
    lib().orc_pk1(kind, _p(Fv), lame1, lame2, _p(out))
    return out


def dpdf(kind, F, lame1, lame2):
    Fv = f64(np.asarray(F).T.reshape(-1))
    out = np.zeros(81)
    lib().orc_dpdf(kind, _p(Fv), lame1, lame2, _p(out))
    return out.reshape(9, 9)


def tet_rest(X):
    X = f64(np.asarray(X).reshape(12))
    D = np.zeros(108)
    V0 = ctypes.c_double()
    lib().orc_tet_rest(_p(X), _p(D), ctypes.byref(V0))
    return D.reshape(9, 12), V0.value


def tet_egh(x, DFDx, V0, kind, lame1, lame2):
    x = f64(np.asarray(x).reshape(12))
    D = f64(DFDx)
    E = ctypes.c_double()
    g = np.zeros(12)
    H = np.zeros(144)
    inv = lib().orc_tet_egh(_p(x), _p(D), V0, kind, lame1, lame2,
                            ctypes.byref(E), _p(g), _p(H))
    return E.value, g, H.reshape(12, 12), bool(inv)


def tet_grid(scale, dims):
    dims = np.asarray(dims, dtype=np.int32)
    n = int(np.prod(dims + 1))
    m = int(5 * np.prod(dims))
    V = np.zeros(3 * n)
    T = np.zeros(4 * m, dtype=np.int32)
    lib().orc_tet_grid(_p(f64(scale)), _p(dims, _ip), _p(V), _p(T, _ip))
    return V.reshape(n, 3), T.reshape(m, 4)


def lame(young, poisson):
    l1, l2 = ctypes.c_double(), ctypes.c_double()
    t0 = 1.0 + poisson
    t1 = 1.0 - 2.0 * poisson
    return (young * poisson) / (t0 * t1), young / (2.0 * t0)


# ---- model-level ----------------------------------------------------------

class Model:
    """Simulable_FEM (Tet4 + lumped mass) with hard-fixed DoFs, gravity and
    nodal-force boundary conditions, restated on the CPU in FP64."""

    def __init__(self, X0, T, kind, lame1, lame2, rho, extra=None):
        """extra: per-element thickness (membranes, kind 2) or bending
        stiffness (hinges, kind 3); None for tet-only models."""
        X0 = f64(X0).reshape(-1)
        T = np.ascontiguousarray(T, dtype=np.int32).reshape(-1)
        m = T.size // 4
        kind = np.ascontiguousarray(np.broadcast_to(kind, (m,)), dtype=np.uint8)
        l1 = f64(np.broadcast_to(lame1, (m,)))
        l2 = f64(np.broadcast_to(lame2, (m,)))
        rh = f64(np.broadcast_to(rho, (m,)))
        self.n = X0.size // 3
        self.m = m
        self.N = 3 * self.n
        if extra is None:
            self._h = lib().orc_model_create(self.n, _p(X0), m, _p(T, _ip),
                                             _p(kind, _ucp), _p(l1), _p(l2), _p(rh))
        else:
            ex = f64(np.broadcast_to(extra, (m,)))
            self._h = lib().orc_model_create_mixed(self.n, _p(X0), m, _p(T, _ip),
                                                   _p(kind, _ucp), _p(l1), _p(l2), _p(rh),
                                                   _p(ex))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_model_destroy(h)
            self._h = None

    def set_fixed(self, mask):
        mk = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
        lib().orc_model_set_fixed(self._h, _p(mk, _ucp))

    def set_gravity(self, g_end, stages, g_ini=(0.0, 0.0, 0.0)):
        lib().orc_model_set_gravity(self._h, _p(f64(g_ini)), _p(f64(g_end)),
                                    int(stages))

    def set_forces(self, f_end, stages, f_ini=None):
        f_end = f64(f_end).reshape(-1)
        f_ini = np.zeros_like(f_end) if f_ini is None else f64(f_ini).reshape(-1)
        lib().orc_model_set_forces(self._h, _p(f_ini), _p(f_end), int(stages))

    def step_load(self):
        return lib().orc_model_step_load(self._h)

    def fully_loaded(self):
        return bool(lib().orc_model_fully_loaded(self._h))

    def set_x(self, x):
        x = f64(x).reshape(-1)
        lib().orc_model_set_x(self._h, _p(x))

    def get_x(self):
        x = np.zeros(self.N)
        lib().orc_model_get_x(self._h, _p(x))
        return x

    def mass(self):
        m = np.zeros(self.N)
        lib().orc_model_get_mass(self._h, _p(m))
        return m

    def energy(self):
        return lib().orc_model_energy(self._h)

    def gradient(self):
        g = np.zeros(self.N)
        lib().orc_model_gradient(self._h, _p(g))
        return g

    def hessian_lower(self):
        nnz = lib().orc_model_nnz_lower(self._h)
        rp = np.zeros(self.N + 1, dtype=np.int64)
        ci = np.zeros(nnz, dtype=np.int32)
        v = np.zeros(nnz)
        lib().orc_model_hessian_lower(self._h, _p(rp, _lp), _p(ci, _ip), _p(v))
        return rp, ci, v

    def element_egh(self):
        E = np.zeros(self.m)
        g = np.zeros((self.m, 12))
        H = np.zeros((self.m, 12, 12))
        lib().orc_model_element_egh(self._h, _p(E), _p(g), _p(H))
        return E, g, H

    def newton(self, **kw):
        o = NewtonOpts()
        lib().orc_newton_default_opts(ctypes.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        r = NewtonResult()
        lib().orc_newton_solve(self._h, ctypes.byref(o), ctypes.byref(r))
        return {"result": OS_NAMES[r.result], "iterations": r.iterations,
                "cg_iterations": r.cg_iterations,
                "bisections": r.line_search_bisections,
                "objective": r.objective, "optimality": r.optimality}


class Dynamic:
    """OptimProblem_BasicDynamic over a Model: one backward-Euler time step
    per step(); keeps the velocity and the solver's previous x."""

    def __init__(self, model):
        self.model = model  # keeps the model alive
        self.N = model.N
        self._h = lib().orc_dynamic_create(model._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_dynamic_destroy(h)
            self._h = None

    def set_velocity(self, v):
        v = f64(v).reshape(-1)
        lib().orc_dynamic_set_velocity(self._h, _p(v))

    def get_velocity(self):
        v = np.zeros(self.N)
        lib().orc_dynamic_get_velocity(self._h, _p(v))
        return v

    def begin(self, dt):
        lib().orc_dynamic_begin(self._h, float(dt))

    def energy(self):
        return lib().orc_dynamic_energy(self._h)

    def gradient(self):
        g = np.zeros(self.N)
        lib().orc_dynamic_gradient(self._h, _p(g))
        return g

    def hessian_lower(self):
        nnz = lib().orc_model_nnz_lower(self.model._h)
        rp = np.zeros(self.N + 1, dtype=np.int64)
        ci = np.zeros(nnz, dtype=np.int32)
        v = np.zeros(nnz)
        lib().orc_dynamic_hessian_lower(self._h, _p(rp, _lp), _p(ci, _ip), _p(v))
        return rp, ci, v

    def solve(self, **kw):
        o = NewtonOpts()
        lib().orc_newton_default_opts(ctypes.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        r = NewtonResult()
        lib().orc_dynamic_solve(self._h, ctypes.byref(o), ctypes.byref(r))
        return {"result": OS_NAMES[r.result], "iterations": r.iterations,
                "cg_iterations": r.cg_iterations,
                "bisections": r.line_search_bisections,
                "objective": r.objective, "optimality": r.optimality}

    def end(self):
        lib().orc_dynamic_end(self._h)

    def step(self, dt, **kw):
        self.begin(dt)
        r = self.solve(**kw)
        self.end()
        return r


def cg(rowptr, col, val, b, x0, tol, maxit):
    n = rowptr.size - 1
    x = f64(x0).copy()
    err = ctypes.c_double()
    it = lib().orc_cg(n, _p(np.ascontiguousarray(rowptr, dtype=np.int64), _lp),
                      _p(np.ascontiguousarray(col, dtype=np.int32), _ip),
                      _p(f64(val)), _p(f64(b)), _p(x), tol, maxit, ctypes.byref(err))
    return x, it, err.value


def lower_to_dense(rowptr, col, val):
    """Symmetric dense matrix from the reference-layout lower CSR (entries with
    row < col, i.e. the upper part of stored diagonal node blocks, are
    ignored like Eigen's selfadjointView<Lower>)."""
    n = rowptr.size - 1
    A = np.zeros((n, n))
    for r in range(n):
        for p in range(rowptr[r], rowptr[r + 1]):
            c = col[p]
            if c <= r:
                A[r, c] = val[p]
                A[c, r] = val[p]
    return A


# ---- discrete elastic rods ---------------------------------------------------

def rod_bt(prm, e1, e2, th1, th2):
    """Bend-twist element: (E, g8, H 8x8); prm in the getBT*OPT argument order."""
    g = np.zeros(8)
    H = np.zeros(64)
    E = lib().orc_rod_bt(_p(f64(prm)), _p(f64(e1)), _p(f64(e2)), th1, th2, _p(g), _p(H))
    return E, g, H.reshape(8, 8)


def rod_stretch(x12, L0, k):
    g = np.zeros(6)
    H = np.zeros(36)
    E = lib().orc_rod_stretch(_p(f64(x12)), L0, k, _p(g), _p(H))
    return E, g, H.reshape(6, 6)


def reference_twist_change(t1, n1, t2, n2, rt):
    return lib().orc_reference_twist_change(_p(f64(t1)), _p(f64(n1)), _p(f64(t2)),
                                            _p(f64(n2)), rt)


class Rod:
    """One open rod (Simulable_DER restated on the CPU)."""

    def __init__(self, X0, frame0_normal, young, shear, radius_h, radius_w, density):
        X0 = f64(X0).reshape(-1)
        self.n = X0.size // 3
        self._h = lib().orc_rod_create(self.n, _p(X0), _p(f64(frame0_normal)), young, shear,
                                       radius_h, radius_w, density)
        self.N = lib().orc_rod_num_dofs(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_rod_destroy(h)
            self._h = None

    def set_fixed(self, mask):
        mk = np.ascontiguousarray(mask, dtype=np.uint8).reshape(-1)
        lib().orc_rod_set_fixed(self._h, _p(mk, _ucp))

    def set_gravity(self, g_end, stages, g_ini=(0.0, 0.0, 0.0)):
        lib().orc_rod_set_gravity(self._h, _p(f64(g_ini)), _p(f64(g_end)), int(stages))

    def step_load(self):
        return lib().orc_rod_step_load(self._h)

    def get_x(self):
        x = np.zeros(self.N)
        lib().orc_rod_get_x(self._h, _p(x))
        return x

    def set_x(self, x):
        lib().orc_rod_set_x(self._h, _p(f64(x)))

    def frames(self):
        t = np.zeros((self.n - 1, 3))
        nn = np.zeros((self.n - 1, 3))
        rt = np.zeros(self.n)
        lib().orc_rod_get_frames(self._h, _p(t), _p(nn), _p(rt))
        return t, nn, rt

    def energy(self):
        return lib().orc_rod_energy(self._h)

    def gradient(self):
        g = np.zeros(self.N)
        lib().orc_rod_gradient(self._h, _p(g))
        return g

    def hessian(self):
        H = np.zeros(self.N * self.N)
        lib().orc_rod_hessian(self._h, _p(H))
        return H.reshape(self.N, self.N)

    def update_kinematics(self):
        lib().orc_rod_update_kinematics(self._h)

    def newton(self, **kw):
        o = NewtonOpts()
        lib().orc_newton_default_opts(ctypes.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        r = NewtonResult()
        lib().orc_rod_newton(self._h, ctypes.byref(o), ctypes.byref(r))
        return {"result": OS_NAMES[r.result], "iterations": r.iterations,
                "bisections": r.line_search_bisections, "objective": r.objective,
                "optimality": r.optimality}


def lbfgs_direction(S, Y, g, count, newest):
    """orc_lbfgs_direction: S, Y are (M, N) arrays (pair i in row i), the
    `count` newest pairs ending at row `newest` are used; returns -H g."""
    S = f64(S)
    Y = f64(Y)
    M, N = S.shape
    dx = np.zeros(N)
    lib().orc_lbfgs_direction(N, M, int(count), int(newest), _p(S), _p(Y), _p(f64(g)),
                              _p(dx))
    return dx


def dense_ldlt_solve(A, b, x, max_error, reg_iters):
    """LinearSolver_EigenLDLT::SolveInternal restated on a dense matrix (lower
    part used, regularization applied in place to a copy).  Returns
    (status 0 = SUCCESS, regularization try used); x is written."""
    A = np.array(A, dtype=np.float64, order="C")
    reg = ctypes.c_int(0)
    st = lib().orc_dense_ldlt_solve(A.shape[0], _p(A), _p(f64(b)), _p(x), max_error,
                                    int(reg_iters), ctypes.byref(reg))
    return st, reg.value


def sparse_ldlt_solve(rowptr, col, val, b, x, max_error, reg_iters):
    """LinearSolver_EigenLDLT::SolveInternal on a lower CSR matrix as
    orc_model_hessian_lower stores it (normA over every stored entry).
    Returns (status 0 = SUCCESS, regularization try used); x is written."""
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    ci = np.ascontiguousarray(col, dtype=np.int32)
    reg = ctypes.c_int(0)
    st = lib().orc_sparse_ldlt_solve(rp.size - 1, _p(rp, _lp), _p(ci, _ip), _p(f64(val)),
                                     _p(f64(b)), _p(x), max_error, int(reg_iters),
                                     ctypes.byref(reg))
    return st, reg.value


# ---- thin shells ---------------------------------------------------------------

def hinge_theta(x):
    return lib().orc_hinge_theta(_p(f64(x).reshape(-1)))


def hinge_shape(x):
    return lib().orc_hinge_shape(_p(f64(x).reshape(-1)))


def hinge_dtheta(x):
    d = np.zeros(12)
    lib().orc_hinge_dtheta(_p(f64(x).reshape(-1)), _p(d))
    return d


def hinge_egh(x, kB, theta0, shape0):
    g = np.zeros(12)
    H = np.zeros(144)
    E = lib().orc_hinge_egh(_p(f64(x).reshape(-1)), kB, theta0, shape0, _p(g), _p(H))
    return E, g, H.reshape(12, 12)


def membrane_rest(X):
    D = np.zeros(4)
    A = np.zeros(1)
    lib().orc_membrane_rest(_p(f64(X).reshape(-1)), _p(D), _p(A))
    return D, A[0]


def membrane_egh(x, Dminv, A0t, lame1, lame2):
    g = np.zeros(9)
    H = np.zeros(81)
    E = lib().orc_membrane_egh(_p(f64(x).reshape(-1)), _p(f64(Dminv)), A0t, lame1, lame2,
                               _p(g), _p(H))
    return E, g, H.reshape(9, 9)


def mem_psi(Fv, l1, l2):
    return lib().orc_mem_psi(_p(f64(Fv)), l1, l2)


def mem_pk1(Fv, l1, l2):
    P = np.zeros(6)
    lib().orc_mem_pk1(_p(f64(Fv)), l1, l2, _p(P))
    return P


def mem_dpdf(Fv, l1, l2):
    C = np.zeros(36)
    lib().orc_mem_dpdf(_p(f64(Fv)), l1, l2, _p(C))
    return C.reshape(6, 6)
