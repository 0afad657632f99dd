"""Backbone decoding known answers: test/c++/test_block_sparse_backbone_eval.cpp:41-135 (TEST(Backbone,
indexing)), and topology enumeration: python/triqs_xca/diag.py, impurity.cpp:53-57,186-190."""
import numpy as np

import pyoracle as po
from xca_amd.diag import all_connected_pairings
from xca_amd.dlr import DLRGrid


def test_backbone_indexing_known_answer():
    topology = np.array([[0, 2], [1, 4], [3, 5]])
    n = 4
    grid = DLRGrid(2.0, 200.0, 1e-10)
    rf = grid.rf
    r = len(rf)
    assert abs(r - 36) <= 1  # the reference's datum r = 36 at Lambda = 200, eps = 1e-10
    assert rf[3] < 0 and rf[19] > 0  # same pole signs as the reference's indices (3, 19)
    fb_ix = 1 + 2 * 1
    p_ix = 3 + r * 19
    o_ix = 3 + n * 2
    f_ix = o_ix + p_ix * n * n + fb_ix * n * n * r * r
    B = po.backbone(topology, n, rf, f_ix)
    assert list(B["fb"]) == [1, 1, 0]
    assert list(B["pole_inds"]) == [3, 19]
    assert list(B["orb"][[1, 3, 4, 5]]) == [3, 2, 3, 2]
    assert list(B["bar"]) == [0, 0, 0, 0, 1, 1]
    assert list(B["dag"]) == [0, 0, 1, 1, 1, 0]
    assert [B["hyb_ind"][v] for v in (1, 3, 4, 5)] == [0, 1, 0, 1]
    assert [B["Ksign"][v] for v in (1, 3, 4, 5)] == [-1, 1, 1, -1]
    assert not B["edges"].any()
    assert list(B["pref_Ksign"]) == [-1, 1]
    assert list(B["pref_Kexp"]) == [1, 1]


def test_backbone_rule_4b_edges():
    """Forward line with a positive pole: K^+ on every edge it spans, prefactor (K^+(0))^-(span-1)
    (backbone.cpp:98-106); backward line with a negative pole: K^- on the edges, sign flip."""
    topology = np.array([[0, 2], [1, 4], [3, 5]])
    poles = np.array([-2.0, 3.0])
    # fb = (1, 1, 0), l = 1 (positive), l' = 0 (negative)
    n = 1
    f_ix = 0 + 1 * (1 + 2 * 0) * 1 + 0
    f_ix = (1 + 2 * 0) + 4 * (1 + 2 * 1)  # p_ix = 1 + p*0, fb_ix = 3, o_ix = 0
    B = po.backbone(topology, n, poles, f_ix)
    assert list(B["fb"]) == [1, 1, 0]
    assert list(B["pole_inds"]) == [1, 0]
    e = B["edges"]
    assert list(e[:, 0]) == [0, 1, 1, 1, 0]
    assert list(e[:, 1]) == [0, 0, 0, -1, -1]
    assert list(B["pref_Ksign"]) == [1, -1]
    assert list(B["pref_Kexp"]) == [2, 1]
    assert B["prefactor_sign"] == -1
    assert not B["Ksign"].any()


def test_topology_enumeration():
    counts = [sum(1 for _ in all_connected_pairings(k)) for k in (1, 2, 3, 4)]
    assert counts == [1, 1, 4, 27]
    assert list(all_connected_pairings(1)) == [(1, [(0, 1)])]
    assert list(all_connected_pairings(2)) == [(-1, [(0, 2), (1, 3)])]
    got = {tuple(t): s for s, t in all_connected_pairings(3)}
    assert got == {((0, 2), (1, 4), (3, 5)): 1, ((0, 3), (1, 4), (2, 5)): -1,
                   ((0, 3), (1, 5), (2, 4)): 1, ((0, 4), (1, 3), (2, 5)): 1}
