"""The tet element kernel's tiles (workgroup size x elements per lane, chosen
per model by the in-library autotuner): every tile computes each element the
same way and takes energy partials per 64 consecutive elements, so energies,
gradients, assembled Hessians and per-element outputs are bit-identical
whichever tile runs.  A model of at least 16,384 tets is tuned at its first
launch and reports the chosen tile and its timing."""
import sys

import numpy as np
import pytest

import physim_amd as pg
from physim_amd import scenes
from tests.helpers import perturbed_x, pgn_model

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _scene():
    s = scenes.cantilever(dims=(12, 12, 40), scale=(0.1, 0.1, 0.3), young_sigma=0.3,
                          jitter=0.05, seed=3)
    s.material[::2] = 1  # both materials
    return s


def _outputs(ctx, s, precision):
    G = pgn_model(ctx, s, precision)
    for _ in range(s.stages):
        G.step_load()
    G.set_x(perturbed_x(s, 2e-4))
    out = (G.energy(), G.gradient(), G.hessian(pg.LAYOUT_BSR3)[2]) + tuple(
        G.element_egh(precision))
    tiles = (G.profile(6), G.profile(7))
    G.close()
    return out, tiles


@pytest.mark.parametrize("precision", [pg.FP64, pg.MIXED])
def test_every_tile_gives_the_same_bits(ctx, precision, monkeypatch):
    _say(f"element tiles precision={precision}")
    s = _scene()
    assert s.m >= 16384
    monkeypatch.delenv("PGN_TET_TILE", raising=False)
    ref, tiles = _outputs(ctx, s, precision)
    _say(f"  tuned tiles (Hessian, E+g): {tiles}")
    assert 0 <= tiles[0][0] < 6 and tiles[0][1] > 0
    assert 0 <= tiles[1][0] < 6 and tiles[1][1] > 0
    for t in range(6):
        monkeypatch.setenv("PGN_TET_TILE", str(t))
        out, forced = _outputs(ctx, s, precision)
        assert forced[0][0] == t
        for a, b in zip(out, ref):
            assert np.array_equal(np.asarray(a), np.asarray(b)), t
