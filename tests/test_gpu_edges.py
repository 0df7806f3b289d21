"""Edge cases of the C ABI on the GPU: the smallest model (one tet), a model
with nodes and no elements, every DoF fixed, a ragged batch whose instances
range from one tet to a cloth, and inputs the library must reject with an
error instead of computing on them.  Results are compared with the FP64
oracle (1e-6 x bounding-box diagonal) or with what the fixed-DoF projection
fixes exactly (FixVectorDoFs / FixMatrixDoFs, Simulable.cpp:735-760)."""
import sys

import numpy as np
import pytest

import physim_amd as pg
from oracle import orc
from physim_amd import scenes
from tests.helpers import oracle_model, pgn_model, small_beam

pytestmark = pytest.mark.gpu

OPTS = dict(pg.CONVERGE_OPTS, numMaxIters=200)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _one_tet(kind):
    X0 = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0], [0.02, 0.03, 0.1]])
    T = np.array([[0, 1, 2, 3]], dtype=np.int32)
    l1, l2 = orc.lame(1e5, 0.3)
    fixed = np.zeros((4, 3), dtype=np.uint8)
    fixed[:3] = 1
    return scenes.Scene(f"one_tet_{kind}", X0, T, np.array([kind], dtype=np.uint8),
                        np.array([l1]), np.array([l2]), np.array([1000.0]), fixed,
                        scenes.G_EARTH, 5)


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
@pytest.mark.parametrize("kind", [orc.CONH, orc.STVK])
def test_single_tet_newton_matches_oracle(ctx, precision, kind):
    _say(f"single tet kind={kind} precision={precision}")
    s = _one_tet(kind)
    O = oracle_model(s)
    ro = O.newton(**OPTS)
    G = pgn_model(ctx, s, precision)
    rg = G.newton(**OPTS)
    assert ro["result"] == "SUCCESS" and rg["result"] == "SUCCESS", (ro, rg)
    assert np.abs(G.get_x() - O.get_x()).max() <= 1e-6 * s.bbox_diag()
    assert np.abs(G.get_x()[9:] - s.X0[3]).max() > 1e-4  # the free node moved
    G.close()


def test_model_without_elements(ctx):
    _say("model without elements")
    X0 = np.random.default_rng(1).random((5, 3))
    T = np.zeros((0, 4), dtype=np.int32)
    G = pg.TetModel(ctx, X0, T, np.zeros(0, np.uint8), np.zeros(0), np.zeros(0),
                    np.zeros(0), pg.FP64)
    G.set_gravity([0.0, -9.81, 0.0], 3)
    assert G.energy() == 0.0
    assert np.all(G.gradient() == 0.0)  # no element, no lumped mass
    r = G.newton(**OPTS)
    assert r["result"] == "SUCCESS" and r["pcg_iterations"] == 0, r
    assert np.array_equal(G.get_x(), X0.reshape(-1))
    rp, ci, v = G.hessian(pg.LAYOUT_LOWER)
    # only the node blocks, all zero; AMatrixSd stores whole blocks with
    # row >= col (AddStaticBlock, BasicTypes.cpp:64-77): 9 entries per node
    assert rp[-1] == 5 * 9 and np.all(v == 0.0)
    G.close()


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_every_dof_fixed(ctx, precision):
    _say(f"every DoF fixed precision={precision}")
    s = small_beam(dims=(2, 2, 3))
    s.fixed = np.ones_like(s.fixed)
    G = pgn_model(ctx, s, precision)
    assert np.all(G.gradient() == 0.0)
    r = G.newton(**OPTS)
    assert r["result"] == "SUCCESS" and r["pcg_iterations"] == 0, r
    assert np.array_equal(G.get_x(), s.X0.reshape(-1))
    rp, ci, v = G.hessian(pg.LAYOUT_LOWER)
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    assert np.all(v[rows == ci] == 1.0) and np.all(v[rows != ci] == 0.0)
    G.close()


def test_ragged_batch_same_alone_and_in_mix(ctx):
    """One tet, a beam and a cloth in one fused batch: each instance's result
    is the one it gets alone and matches the oracle."""
    _say("ragged batch")
    rng = np.random.default_rng(7)
    mix = [_one_tet(orc.CONH),
           scenes.batch_beam(rng, cell=0.01, dmin=(2, 2, 4), dmax=(3, 3, 6)),
           scenes.cloth(5, size=0.2, stages=10)]
    B = pg.Batch(ctx, pg.MIXED)
    for s in mix:
        B.add_scene(s)
    B.build()
    res = B.newton(**OPTS)
    for k, s in enumerate(mix):
        assert res[k]["result"] == "SUCCESS", (k, res[k])
        A = pg.Batch(ctx, pg.MIXED)
        A.add_scene(s)
        A.build()
        ra = A.newton(**OPTS)
        assert ra[0]["result"] == "SUCCESS", (k, ra)
        assert np.abs(A.get_x(0) - B.get_x(k)).max() <= 1e-6 * s.bbox_diag(), k
        A.close()
    O = oracle_model(mix[0])
    O.newton(**OPTS)
    assert np.abs(B.get_x(0) - O.get_x().reshape(-1, 3)).max() <= 1e-6 * mix[0].bbox_diag()
    B.close()


def test_create_rejects_bad_input(ctx):
    _say("bad inputs")
    s = _one_tet(orc.CONH)
    args = [s.material, s.lame1, s.lame2, s.density]
    for T in (np.array([[0, 1, 2, 4]]), np.array([[0, 1, -1, 3]])):  # node id out of range
        with pytest.raises(pg.PgnError):
            pg.TetModel(ctx, s.X0, T.astype(np.int32), *args, pg.FP64)
    with pytest.raises(pg.PgnError):  # unknown element kind
        pg.TetModel(ctx, s.X0, s.T, np.array([7], np.uint8), *args[1:], pg.FP64)
    with pytest.raises(pg.PgnError):  # unknown precision
        pg.TetModel(ctx, s.X0, s.T, *args, 9)
    with pytest.raises(pg.PgnError):  # no node
        pg.TetModel(ctx, np.zeros((0, 3)), np.zeros((0, 4), np.int32), np.zeros(0, np.uint8),
                    np.zeros(0), np.zeros(0), np.zeros(0), pg.FP64)
    G = pgn_model(ctx, s, pg.FP64)
    x = G.get_x()
    with pytest.raises(pg.PgnError):  # the model is unchanged after a rejected call
        G.dynamic_step(-1.0)
    assert np.array_equal(G.get_x(), x)
    G.close()
