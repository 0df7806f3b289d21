"""Fused batches (pgn_batch) on the GPU: every instance keeps its own Newton
state, so its result matches the FP64 oracle's solo solve and the same
problem solved alone, within 1e-6 x its bounding-box diagonal; per-instance
energies match single models to 1e-12 relative."""
import sys

import numpy as np
import pytest

import physim_amd as pg
from oracle import orc
from physim_amd import scenes

pytestmark = pytest.mark.gpu

OPTS = dict(pg.CONVERGE_OPTS, numMaxIters=200)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _small_mix(seed):
    rng = np.random.default_rng(seed)
    tets = [scenes.batch_beam(rng, cell=0.01, dmin=(2, 2, 4), dmax=(4, 4, 10)) for _ in range(5)]
    shells = [scenes.cloth(int(n), size=0.2, stages=10) for n in (5, 7)]
    return tets + shells


def _oracle(scene):
    if hasattr(scene, "hinges"):
        T, kind, l1, l2, rho, extra = scene.oracle_arrays()
        O = orc.Model(scene.X0, T, kind, l1, l2, rho, extra)
    else:
        O = orc.Model(scene.X0, scene.T, scene.material, scene.lame1, scene.lame2,
                      scene.density)
    scene.apply(O)
    return O


def _solo(ctx, scene, precision):
    if hasattr(scene, "hinges"):
        G = pg.ShellModel(ctx, scene.X0, scene.F, scene.lame1, scene.lame2, scene.thickness,
                          scene.density, scene.hinges, scene.bending_k, precision)
    else:
        G = pg.TetModel(ctx, scene.X0, scene.T, scene.material, scene.lame1, scene.lame2,
                        scene.density, precision)
    scene.apply(G)
    return G


def test_batch_energy_matches_single_models(ctx):
    _say("batch energies vs single models")
    sc = _small_mix(1)
    B = pg.Batch(ctx, pg.FP64)
    for s in sc:
        B.add_scene(s)
    B.build()
    E = B.energy()
    for k, s in enumerate(sc):
        G = _solo(ctx, s, pg.FP64)
        e = G.energy()
        assert abs(E[k] - e) <= 1e-12 * (abs(e) + 1e-12), k
        G.close()


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_batch_newton_matches_oracle(ctx, precision):
    _say(f"batch newton precision={precision}")
    sc = _small_mix(2)
    B = pg.Batch(ctx, precision)
    for s in sc:
        B.add_scene(s)
    B.build()
    res = B.newton(**OPTS)
    for k, s in enumerate(sc):
        O = _oracle(s)
        ro = O.newton(**OPTS)
        _say(f"  {s.name}: gpu {res[k]['result']} {res[k]['iterations']} its, oracle "
             f"{ro['result']} {ro['iterations']} its")
        assert ro["result"] == "SUCCESS", (k, ro)
        assert res[k]["result"] == "SUCCESS" and res[k]["optimality"] <= 1e-8, (k, res[k])
        err = np.abs(B.get_x(k) - O.get_x().reshape(-1, 3)).max()
        assert err <= 1e-6 * s.bbox_diag(), (k, err)


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_batch_16bit_columns_match_32bit(ctx, precision, monkeypatch):
    """The segmented SpMV with block columns as 16-bit offsets from their
    rows (default where they fit) and as 32-bit columns (PGN_SELL_C16=0,
    read at a batch's first solve): the same products in the same order, so
    the same iteration counts and bit-identical positions."""
    _say(f"batch 16-bit columns precision={precision}")
    sc = _small_mix(4)
    out = []
    for c16 in ("1", "0"):
        monkeypatch.setenv("PGN_SELL_C16", c16)
        B = pg.Batch(ctx, precision)
        for s in sc:
            B.add_scene(s)
        B.build()
        res = B.newton(**OPTS)
        assert all(r["result"] == "SUCCESS" for r in res), res
        out.append(([(r["iterations"], r["pcg_iterations"]) for r in res],
                    [B.get_x(k).copy() for k in range(len(sc))]))
        B.close()
    assert out[0][0] == out[1][0]
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)


def test_batch_result_same_alone_and_in_mix(ctx):
    _say("batch alone vs mix")
    sc = _small_mix(3)
    B = pg.Batch(ctx, pg.MIXED)
    for s in sc:
        B.add_scene(s)
    B.build()
    B.newton(**OPTS)
    for k in (1, 5):
        A = pg.Batch(ctx, pg.MIXED)
        A.add_scene(sc[k])
        A.build()
        r = A.newton(**OPTS)
        assert r[0]["result"] == "SUCCESS"
        err = np.abs(A.get_x(0) - B.get_x(k)).max()
        assert err <= 1e-6 * sc[k].bbox_diag(), (k, err)
        A.close()
    # the batch can be solved again from rest with the same result
    x1 = B.get_x(2).copy()
    B.reset()
    B.newton(**OPTS)
    assert np.abs(B.get_x(2) - x1).max() <= 1e-6 * sc[2].bbox_diag()


@pytest.mark.parametrize("precision,nodes", [(pg.FP64, "1000"), (pg.MIXED, "1000"),
                                             (pg.MIXED, "40")])
def test_batch_twolevel_matches_block_jacobi(ctx, precision, nodes, monkeypatch):
    """The segmented PCG with each instance's own coarse level (precond 2;
    PGN_SEG_COARSE_NODES=40 gives these small instances 3-8 aggregates
    instead of 2) converges every instance to the block-Jacobi batch's
    result and the oracle's within 1e-6 x its bounding-box diagonal, in fewer
    PCG iterations over the batch."""
    _say(f"batch two-level precision={precision} nodes={nodes}")
    monkeypatch.setenv("PGN_SEG_COARSE_NODES", nodes)
    sc = _small_mix(5)
    out = {}
    for pc in (1, 2):
        B = pg.Batch(ctx, precision)
        for s in sc:
            B.add_scene(s)
        B.build()
        res = B.newton(**OPTS, precond=pc)
        assert all(r["result"] == "SUCCESS" for r in res), (pc, res)
        out[pc] = ([B.get_x(k).copy() for k in range(len(sc))],
                   sum(r["pcg_iterations"] for r in res))
        B.close()
    _say(f"  PCG iterations: block Jacobi {out[1][1]}, two-level {out[2][1]}")
    assert out[2][1] < out[1][1]
    for k, s in enumerate(sc):
        O = _oracle(s)
        ro = O.newton(**OPTS)
        assert ro["result"] == "SUCCESS"
        xo = O.get_x().reshape(-1, 3)
        for pc in (1, 2):
            err = np.abs(out[pc][0][k] - xo).max()
            assert err <= 1e-6 * s.bbox_diag(), (k, pc, err)


def test_batch_coarse_giveup_matches_oracle(ctx):
    """Two-level batch with coarse_giveup low enough that some instances
    drop their coarse level mid-run (solves then mix instances with and
    without it, and later run block Jacobi outright): every instance still
    converges to the oracle's equilibrium within 1e-6 x its bounding-box
    diagonal."""
    _say("batch coarse give-up")
    sc = _small_mix(6)
    B = pg.Batch(ctx, pg.MIXED)
    for s in sc:
        B.add_scene(s)
    B.build()
    res = B.newton(**OPTS, precond=2, coarse_giveup=15)
    _say("  " + ", ".join(f"{r['result']}/{r['iterations']}/{r['pcg_iterations']}" for r in res))
    for k, s in enumerate(sc):
        assert res[k]["result"] == "SUCCESS", (k, res[k])
        O = _oracle(s)
        ro = O.newton(**OPTS)
        assert ro["result"] == "SUCCESS"
        err = np.abs(B.get_x(k) - O.get_x().reshape(-1, 3)).max()
        assert err <= 1e-6 * s.bbox_diag(), (k, err)
    B.close()
