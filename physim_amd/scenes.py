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


def batch_beam(rng, cell=0.0025, name=None, dmin=(4, 4, 16), dmax=(24, 24, 96)):
    """One beam of the 256-instance batch: hex-grid dims U{4..24} x U{4..24} x
    U{16..96} of `cell`-sized cells, CoNH or StVK (50/50), Young's modulus
    lognormal(ln 1e5, 0.5), Poisson U(0.25, 0.45), density 1000 kg/m^3; one of
    the 6 faces (uniform) clamped; gravity -9.81 y plus a traction of
    U(0, 50) Pa in a random direction on the opposite face, lumped to its
    nodes (each face cell gives a quarter of its force to each corner); both
    ramped over 10 load stages."""
    dims = tuple(int(rng.integers(lo, hi + 1)) for lo, hi in zip(dmin, dmax))
    scale = tuple(d * cell for d in dims)
    V, T = tet_grid(scale, dims)
    m = T.shape[0]
    E = np.exp(np.log(1e5) + 0.5 * rng.standard_normal(m))
    nu = rng.uniform(0.25, 0.45)
    l1, l2 = lame(E, nu)
    mat = np.full(m, int(rng.integers(0, 2)), dtype=np.uint8)
    face = int(rng.integers(0, 6))
    ax, hi = face // 2, face % 2
    lo_v, hi_v = V[:, ax].min(), V[:, ax].max()
    on_clamp = np.abs(V[:, ax] - (hi_v if hi else lo_v)) < 1e-12
    on_load = np.abs(V[:, ax] - (lo_v if hi else hi_v)) < 1e-12
    fixed = np.zeros((V.shape[0], 3), dtype=np.uint8)
    fixed[on_clamp] = 1
    # lumped traction: tributary area of each node on the loaded face
    others = [a for a in range(3) if a != ax]
    area = np.ones(V.shape[0])
    for a in others:
        c = V[:, a]
        edge = (np.abs(c - c.min()) < 1e-12) | (np.abs(c - c.max()) < 1e-12)
        area *= np.where(edge, 0.5 * cell, cell)
    t = rng.uniform(0.0, 50.0)
    d = rng.standard_normal(3)
    d /= np.linalg.norm(d)
    forces = np.zeros_like(V)
    forces[on_load] = t * area[on_load, None] * d[None, :]
    return Scene(name or f"beam{dims}_{'stvk' if mat[0] else 'conh'}", V, T, mat, l1, l2,
                 np.full(m, 1000.0), fixed, G_EARTH, 10, forces=forces)


def batch_beams(count=256, seed=3, cell=0.0025):
    """The 256-instance fused-batch workload (seeded)."""
    rng = np.random.default_rng(seed)
    return [batch_beam(rng, cell) for _ in range(count)]


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


# ---- thin shells ---------------------------------------------------------------

def grid_tri3(N, M):
    """create2DGridIndex_Tri3 (MeshUtils.cpp:61-71): node (i, j) = i*M + j,
    two triangles per cell."""
    F = []
    for i in range(N - 1):
        for j in range(M - 1):
            F.append((i * M + j, i * M + j + 1, (i + 1) * M + j))
            F.append(((i + 1) * M + j, i * M + j + 1, (i + 1) * M + j + 1))
    return np.array(F, dtype=np.int32)


def shell_hinges(F):
    """Interior edges of a triangle mesh as hinges, in the order and node
    order Simulable_ThinShell / EnergyElement_DiscreteShells use: edges in
    creation order of Mesh_Face's half-edge build (faces in order, edges
    (k, k+1) of each face), each hinge [target of half0, target of half1,
    vertex opposite in half0's face, vertex opposite in half1's face], where
    half0 is the edge's first half-edge (Mesh_Face.cpp:560-636,
    EnergyElement_DiscreteShells.cpp:30-34)."""
    first = {}    # undirected edge -> (a, b, opposite) of its first half-edge
    order = []
    second = {}
    for f in F:
        for k in range(3):
            a, b, c = int(f[k]), int(f[(k + 1) % 3]), int(f[(k + 2) % 3])
            key = (min(a, b), max(a, b))
            if key not in first:
                first[key] = (a, b, c)
                order.append(key)
            else:
                second[key] = (a, b, c)
    H = []
    for key in order:
        if key not in second:
            continue  # boundary edge: no bending element
        a0, b0, c0 = first[key]
        a1, b1, c1 = second[key]
        H.append((b0, b1, c0, c1))
    return np.array(H, dtype=np.int32).reshape(-1, 4)


class ShellScene:
    """Inputs of one quasi-static thin-shell solve."""

    def __init__(self, name, X0, F, hinges, lame1, lame2, thickness, density, bending_k,
                 fixed, gravity, stages):
        self.name = name
        self.X0 = X0
        self.F = F
        self.hinges = hinges
        self.lame1 = lame1
        self.lame2 = lame2
        self.thickness = thickness
        self.density = density
        self.bending_k = bending_k
        self.fixed = fixed
        self.gravity = gravity
        self.stages = stages

    @property
    def n(self):
        return self.X0.shape[0]

    def bbox_diag(self):
        return float(np.linalg.norm(self.X0.max(0) - self.X0.min(0)))

    def apply(self, model):
        model.set_fixed(self.fixed)
        model.set_gravity(self.gravity, self.stages)

    def oracle_arrays(self):
        """Element arrays of the oracle's mixed model: triangles (kind 2, row
        v0 v1 v2 v2) then hinges (kind 3)."""
        nt, nh = len(self.F), len(self.hinges)
        T = np.concatenate([np.concatenate([self.F, self.F[:, 2:3]], axis=1), self.hinges])
        kind = np.concatenate([np.full(nt, 2, np.uint8), np.full(nh, 3, np.uint8)])
        l1 = np.concatenate([np.broadcast_to(self.lame1, (nt,)), np.zeros(nh)])
        l2 = np.concatenate([np.broadcast_to(self.lame2, (nt,)), np.zeros(nh)])
        rho = np.concatenate([np.broadcast_to(self.density, (nt,)), np.zeros(nh)])
        extra = np.concatenate([np.broadcast_to(self.thickness, (nt,)),
                                np.broadcast_to(self.bending_k, (nh,))])
        return T.astype(np.int32), kind, l1, l2, rho, extra


def cloth(n, size=1.0, thickness=1e-3, young=1e6, poisson=0.3, density=0.2, stages=10,
          jitter=0.0, seed=0, name=None, vertical=False):
    """n x n node sheet of side `size` in the horizontal xz-plane, pinned at
    the two corners of its x = 0 edge, under gravity -9.81 y; or, vertical,
    hanging in the xy-plane below its top edge (y = 0) from that edge's two
    corners.  Membrane:
    2D StVK with the Lame parameters of (young, poisson); bending stiffness
    kB = young t^3 / (12 (1 - poisson^2)) per hinge (passed explicitly, as
    Param_BendingK); density in kg/m^2 (the reference lumps density x area,
    with no thickness factor).  jitter > 0 moves nodes by U(-j, j) cells
    along the sheet's normal (a non-flat rest shape)."""
    h = size / (n - 1)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    if vertical:
        X0 = np.stack([i.ravel() * h, -j.ravel() * h, np.zeros(n * n)], axis=1)
        normal = 2
    else:
        X0 = np.stack([i.ravel() * h, np.zeros(n * n), j.ravel() * h], axis=1)
        normal = 1
    if jitter > 0:
        rng = np.random.default_rng(seed)
        X0[:, normal] += rng.uniform(-jitter, jitter, n * n) * h
    F = grid_tri3(n, n)
    Hn = shell_hinges(F)
    l1, l2 = lame(young, poisson)
    kB = young * thickness ** 3 / (12.0 * (1.0 - poisson ** 2))
    fixed = np.zeros((n * n, 3), dtype=np.uint8)
    fixed[0] = 1                                 # node (0, 0)
    fixed[(n - 1) * n if vertical else n - 1] = 1  # the other end of the pinned edge
    return ShellScene(name or f"cloth_{n}x{n}", X0, F, Hn, float(l1), float(l2), thickness,
                      density, kB, fixed, G_EARTH, stages)
