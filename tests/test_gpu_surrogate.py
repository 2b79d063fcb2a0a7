"""The committed warm-start weights on held-out channels (seed 9001, never
trained on): the learned start is closer to the steady state than the same
network with its random initial weights (the trained-vs-untrained ablation), and
single-lattice runs from it reach the 1e-6 residual in fewer steps than runs
from rest, ending within a few per cent of the rest-started state (the
Zou-He channel's steady state depends on the start)."""
import os

import numpy as np
import pytest

from thesis_amd import lattices as LT
from thesis_amd import lbm as L
from thesis_amd import surrogate as S

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(S.__file__), "surrogate_weights.npz")


def test_learned_start_beats_untrained_and_rest():
    print("case warm start, 4 held-out channels", flush=True)
    net = S.Net.load(WEIGHTS)
    untrained = S.Net()
    test = LT.surrogate_channels(4, 9001)
    pred = S.predict_fields(net, test)
    base = S.predict_fields(untrained, test)
    steps_r, steps_w = [], []
    for d, p, b0 in zip(test, pred, base):
        n_r, _, s_r = S.rest_start_single(d)
        _, u_r, v_r, _ = (a.reshape(d["node_class"].shape) for a in s_r.macros())
        s_r.close()
        fl = d["node_class"] != L.SOLID
        e_learned = S.rel_err((p[1], p[2]), (u_r, v_r), fl)[1]
        e_untrained = S.rel_err((b0[1], b0[2]), (u_r, v_r), fl)[1]
        print("start L2 error learned", e_learned, "untrained", e_untrained, flush=True)
        assert e_learned < 0.8 * e_untrained
        n_w, _, s_w = S.warm_start_single(net, d)
        _, u_w, v_w, _ = (a.reshape(d["node_class"].shape) for a in s_w.macros())
        s_w.close()
        e_final = S.rel_err((u_w, v_w), (u_r, v_r), fl)[0]
        print("steps rest", n_r, "warm", n_w, "final max difference", e_final, flush=True)
        assert e_final < 0.05
        steps_r.append(n_r)
        steps_w.append(n_w)
    assert sum(steps_r) > 1.2 * sum(steps_w)
