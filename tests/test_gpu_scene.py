"""A scene document in the reference's XML vocabulary, loaded, turned into a
device model (pgn_scene_create_model) and solved with its own solver options
(<solver>): the equilibrium matches the oracle's for the same model, loads
and stages within 1e-6 x bounding-box diagonal.  Gravity from XML has no
animator, so it applies in full from the first stage (BCondition.h:140-155);
the oracle is set up the same way."""
import sys

import numpy as np
import pytest

import physim_amd as pg
from oracle import orc
from tests.test_scene import CONFIG1

pytestmark = pytest.mark.gpu


def test_scene_model_solves_like_the_oracle(ctx, tmp_path):
    print("scene -> model -> newton", flush=True)
    xml = CONFIG1.replace("<dims>10,10,40</dims>", "<dims>3,3,12</dims>")
    p = tmp_path / "beam.xml"
    p.write_text(xml)
    S = pg.Scene(p)
    O = orc.Model(S.X0, S.T, S.material, S.lame1, S.lame2, S.density)
    O.set_fixed(S.fixed)
    O.set_gravity(S.gravity, S.stages, g_ini=S.gravity_ini)
    opts = dict(S.opts)
    extra = dict(tolMinImpr=0.0, tolMinStep=0.0, ls_noise_guard=1e-12)
    ro = O.newton(**{k: opts[k] for k in ("tolMaxError", "numMaxIters", "lsNumMaxIters",
                                           "lsTolMaxError")}, **extra)
    for prec in (pg.FP64, pg.MIXED):
        G = S.create_model(ctx, prec)
        r = G.newton(**dict(opts, **extra))
        print(f"  precision {prec}: {r['result']} {r['iterations']} its; oracle {ro}", flush=True)
        assert ro["result"] == "SUCCESS" and r["result"] == "SUCCESS"
        diag = np.linalg.norm(S.X0.max(0) - S.X0.min(0))
        assert np.abs(G.get_x() - O.get_x()).max() <= 1e-6 * diag
        G.close()
    S.close()
