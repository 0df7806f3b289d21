"""Shared test helpers: small meshes, perturbations and the oracle twin of a
scene."""
import numpy as np

from oracle import orc
from physim_amd import scenes


def small_beam(dims=(2, 2, 5), material=0, seed=0, mixed_materials=False):
    s = scenes.cantilever(dims=dims, scale=(0.1, 0.1, 0.25), young_sigma=0.3,
                          jitter=0.05, seed=seed, material=material)
    if mixed_materials:
        rng = np.random.default_rng(seed + 1)
        s.material = rng.integers(0, 2, s.m).astype(np.uint8)
    return s


def perturbed_x(scene, amp=0.01, seed=0):
    rng = np.random.default_rng(seed)
    x = scene.X0 + amp * rng.standard_normal(scene.X0.shape)
    x[scene.fixed.astype(bool)] = scene.X0[scene.fixed.astype(bool)]
    return x


def oracle_model(scene):
    M = orc.Model(scene.X0, scene.T, scene.material, scene.lame1, scene.lame2,
                  scene.density)
    scene.apply(M)
    return M


def pgn_model(ctx, scene, precision):
    import physim_amd as pg
    M = pg.TetModel(ctx, scene.X0, scene.T, scene.material, scene.lame1,
                    scene.lame2, scene.density, precision)
    scene.apply(M)
    return M


def lower_csr_to_dense(rp, ci, v, n):
    return orc.lower_to_dense(rp, ci, v) if n <= 3000 else None
