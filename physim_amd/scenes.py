"""Synthetic scenes of the benchmark workloads (host-side numpy).

tet_grid restates CreateTetMesh_Grid (MeshUtils.cpp:927-991): nodes on a
regular grid, x fastest, and the same 5-tet split of every hex cell.
Lame parameters follow ParameterSet::InitLinearFromYoungPoisson
(ParameterSet.cpp:54-70).
"""
import numpy as np

G_EARTH = (0.0, -9.81, 0.0)

# local tets of one hex cell; corner c = bit0 (+x) | bit1 (+y) | bit2 (+z)
_HEX_TETS = np.array([[0, 1, 2, 4], [6, 7, 4, 2], [5, 4, 7, 1], [3, 2, 1, 7],
                      [1, 2, 4, 7]], dtype=np.int64)


def tet_grid(scale, dims):
    dx, dy, dz = (int(d) for d in dims)
    nx, ny, nz = dx + 1, dy + 1, dz + 1
    h = np.asarray(scale, dtype=np.float64) / np.asarray(dims, dtype=np.float64)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    V = np.stack([i.ravel() * h[0], j.ravel() * h[1], k.ravel() * h[2]], axis=1)
    ck, cj, ci = np.meshgrid(np.arange(dz), np.arange(dy), np.arange(dx), indexing="ij")
    o = (nx * ny * ck + nx * cj + ci).ravel()
    corner = np.array([0, 1, nx, nx + 1, nx * ny, nx * ny + 1, nx * ny + nx,
                       nx * ny + nx + 1], dtype=np.int64)
    T = (o[:, None, None] + corner[_HEX_TETS][None, :, :]).reshape(-1, 4)
    return V, T.astype(np.int32)


def lame(young, poisson):
    young = np.asarray(young, dtype=np.float64)
    t0 = 1.0 + poisson
    t1 = 1.0 - 2.0 * poisson
    return young * poisson / (t0 * t1), young / (2.0 * t0)


class Scene:
    """Inputs of one quasi-static tet solve."""

    def __init__(self, name, X0, T, material, lame1, lame2, density, fixed,
                 gravity, stages, forces=None, tol=1e-8):
        self.name = name
        self.X0 = X0
        self.T = T
        self.material = material
        self.lame1 = lame1
        self.lame2 = lame2
        self.density = density
        self.fixed = fixed
        self.gravity = gravity
        self.stages = stages
        self.forces = forces
        self.tol = tol

    @property
    def n(self):
        return self.X0.shape[0]

    @property
    def m(self):
        return self.T.shape[0]

    def bbox_diag(self):
        return float(np.linalg.norm(self.X0.max(0) - self.X0.min(0)))

    def apply(self, model):
        """Sets fixation and loads on an oracle or libpgn model."""
        model.set_fixed(self.fixed)
        model.set_gravity(self.gravity, self.stages)
        if self.forces is not None:
            model.set_forces(self.forces, self.stages)


def cantilever(dims=(10, 10, 40), scale=(0.1, 0.1, 0.4), young=1e5, poisson=0.3,
               density=1000.0, material=0, stages=10, tol=1e-8, young_sigma=0.0,
               jitter=0.0, seed=0, name=None):
    """Beam along z clamped at z = 0 under gravity -9.81 y, ramped over
    `stages` load increments.  young_sigma > 0 draws per-tet Young's moduli
    lognormal(ln young, sigma); jitter > 0 moves every node by U(-j, j) cell
    sizes per component (no z jitter on the clamped face)."""
    V, T = tet_grid(scale, dims)
    rng = np.random.default_rng(seed)
    m = T.shape[0]
    if young_sigma > 0:
        E = np.exp(np.log(young) + young_sigma * rng.standard_normal(m))
    else:
        E = np.full(m, float(young))
    if jitter > 0:
        h = np.asarray(scale, dtype=np.float64) / np.asarray(dims, dtype=np.float64)
        J = rng.uniform(-jitter, jitter, size=V.shape) * h[None, :]
        J[V[:, 2] == 0.0, 2] = 0.0
        V = V + J
    l1, l2 = lame(E, poisson)
    fixed = np.zeros((V.shape[0], 3), dtype=np.uint8)
    fixed[np.abs(V[:, 2]) < 1e-12] = 1
    mat = np.full(m, material, dtype=np.uint8)
    return Scene(name or f"cantilever{dims}", V, T, mat, l1, l2,
                 np.full(m, float(density)), fixed, G_EARTH, stages, tol=tol)


def config1():
    """10x10x40 CoNH cantilever, 20,000 tets, FP64 reference case."""
    return cantilever(name="beam_10x10x40_conh")


def config2(seed=2):
    """40x40x160 CoNH cantilever, 1.28M tets, lognormal E (sigma 0.3), 5%
    node jitter."""
    return cantilever(dims=(40, 40, 160), young_sigma=0.3, jitter=0.05, seed=seed,
                      name="beam_40x40x160_conh_lognormal")


# ---- discrete elastic rods -----------------------------------------------------

def helix(n, radius, pitch, seg_angle=2 * np.pi / 24):
    """n vertices on a helix about z; consecutive vertices seg_angle apart."""
    t = np.arange(n) * seg_angle
    return np.stack([radius * np.cos(t), radius * np.sin(t), pitch * t / (2 * np.pi)], axis=1)


class RodScene:
    """Inputs of a batch of clamped rods under gravity."""

    def __init__(self, rods, young, shear, radius, density, stages, name):
        self.rods = rods
        self.young = young
        self.shear = shear
        self.radius = radius
        self.density = density
        self.stages = stages
        self.name = name

    @property
    def nverts(self):
        return np.array([len(r) for r in self.rods])

    def frame0_normals(self):
        out = []
        for r in self.rods:
            t = r[1] - r[0]
            a = np.array([0.0, 0.0, 1.0]) if abs(t[2]) < 0.9 * np.linalg.norm(t) else \
                np.array([1.0, 0.0, 0.0])
            out.append(a)
        return np.array(out)

    def fixed_mask(self):
        """First two vertices and the first edge angle of every rod."""
        parts = []
        for r in self.rods:
            n = len(r)
            m = np.zeros(4 * n - 1, dtype=np.uint8)
            m[:6] = 1
            m[3 * n] = 1
            parts.append(m)
        return np.concatenate(parts)

    def bbox_diags(self):
        return np.array([np.linalg.norm(r.max(0) - r.min(0)) for r in self.rods])


def rod_batch(n_rods, n_min=50, n_max=500, seed=4, stages=10, young=1e7,
              radius=1e-3, density=None, name=None):
    """Helices with vertex counts U{n_min..n_max}, radius U(0.01, 0.05) m,
    pitch U(0.005, 0.05) m, E = 1e7 Pa, G = E/3, circular section r = 1e-3 m;
    density is kg per metre of rod (default: 1000 kg/m^3 x pi r^2)."""
    rng = np.random.default_rng(seed)
    rods = []
    for _ in range(n_rods):
        n = int(rng.integers(n_min, n_max + 1))
        R = rng.uniform(0.01, 0.05)
        p = rng.uniform(0.005, 0.05)
        rods.append(helix(n, R, p))
    dens = 1000.0 * np.pi * radius ** 2 if density is None else density
    return RodScene(rods, young, young / 3.0, radius, dens, stages,
                    name or f"helices_{n_rods}")
