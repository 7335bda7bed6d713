"""Model sweeps on the CPU side: the seeded parameter draw of the 256-model
sweep and its packing into batches under the per-batch node limit."""
import collections

import numpy as np
import pytest

import fjmodels_amd as fj
from fjmodels_amd import sweep as S


def test_draw_matches_the_stated_distributions():
    ms = fj.draw_models(256, seed=0)
    assert len(ms) == 256
    nr = np.array([len(m["ar"]) for m in ms])
    nt = np.array([m["ntheta"] for m in ms])
    lmax = np.array([2 * (m["npoly"] - 1) for m in ms])
    assert nr.min() >= 24 and nr.max() <= 96 and nt.min() >= 12 and nt.max() <= 32
    assert set(lmax) <= {8, 10, 12} and len(set(lmax)) == 3
    kinds = collections.Counter(m["model"] for m in ms)
    assert set(kinds) == {"Hernquist", "Isochrone", "NFW"} and min(kinds.values()) > 50
    for m in ms:
        d = m["df"]
        assert 0.3 <= d.dphi_h <= 1.2 and 0.3 <= d.dz_h <= 1.2 and 0.0 <= d.chi <= 2.0 and d.rot_form == 1
        assert np.all(np.diff(m["ar"]) > 0)
    pts = int(np.sum(nr * nt))
    # about 3.4e5 grid points and 3.6e9 velocity nodes per iteration
    assert 3.0e5 < pts < 3.9e5 and 3.2e9 < pts * 22 ** 3 < 4.1e9
    # the draw is a function of the seed
    again = fj.draw_models(256, seed=0)
    assert all(len(a["ar"]) == len(b["ar"]) and a["df"].chi == b["df"].chi for a, b in zip(ms, again))


def test_partition_respects_the_node_limit_and_the_order():
    ms = fj.draw_models(256, seed=0)
    groups = S.partition(ms)
    assert [k for g in groups for k in g] == list(range(256))
    for g in groups:
        assert sum(S.model_nodes(ms[k]) for k in g) <= S.BATCH_NODE_LIMIT
    # greedy: the next model would not have fitted
    for g, h in zip(groups, groups[1:]):
        assert sum(S.model_nodes(ms[k]) for k in g) + S.model_nodes(ms[h[0]]) > S.BATCH_NODE_LIMIT
    small = S.partition(ms[:5], node_limit=max(S.model_nodes(m) for m in ms[:5]))
    assert len(small) >= 2
    with pytest.raises(ValueError):
        S.partition(ms[:1], node_limit=S.model_nodes(ms[0]) - 1)
