"""CPU: merging scenes into one block-diagonal batch scene, and the oracle's
per-scene contact filter for merged scenes (pinned against per-scene
brute-force detection)."""
import numpy as np
import pytest

from helpers import drape_on_obstacle

PARTS = [dict(side=12, radius=0.10, drop=0.05), dict(side=20, radius=0.15, drop=0.02), dict(side=9, radius=0.20, drop=0.1)]


def _parts(pa):
    return [pa.Scene.generate(4, **p) for p in PARTS]


def test_merge_layout_and_data(pa):
    parts = _parts(pa)
    merged = pa.Scene.merge(parts)
    seg = merged.segments()
    info = merged.info
    assert seg["count"] == len(parts)
    nd = [p.info["n_dyn"] for p in parts]
    ns = [p.info["n_all"] - p.info["n_dyn"] for p in parts]
    assert info["n_dyn"] == sum(nd) and info["n_all"] == sum(nd) + sum(ns)
    np.testing.assert_array_equal(seg["dyn_off"], np.concatenate([[0], np.cumsum(nd)]))
    np.testing.assert_array_equal(seg["stat_off"], sum(nd) + np.concatenate([[0], np.cumsum(ns)]))
    for k in ("m_rows", "n_tri", "n_tri_c", "n_bend_c", "n_surf_verts", "n_surf_tris", "n_surf_edges"):
        assert info[k] == sum(p.info[k] for p in parts), k
    mx = merged.mesh()["x"]
    mm, mc = merged.mass(), merged.constraints()
    for s, p in enumerate(parts):
        x = p.mesh()["x"]
        d0, s0 = seg["dyn_off"][s], seg["stat_off"][s]
        n = p.info["n_dyn"]
        np.testing.assert_array_equal(mx[d0:d0 + n], x[:n])
        np.testing.assert_array_equal(mx[s0:s0 + len(x) - n], x[n:])
        assert np.all(seg["seg"][d0:d0 + n] == s) and np.all(seg["seg"][s0:s0 + len(x) - n] == s)
        np.testing.assert_array_equal(mm[d0:d0 + n], p.mass()[:n])
    # elements keep scene order: triangle weights are the parts' concatenated
    np.testing.assert_array_equal(mc["tri_w"], np.concatenate([p.constraints()["tri_w"] for p in parts]))
    np.testing.assert_array_equal(np.sort(mc["bend_w"]), np.sort(np.concatenate([p.constraints()["bend_w"] for p in parts])))


def test_oracle_segment_filter_equals_per_scene_detection(pa):
    """Parts overlap in space (every sphere sits at the origin): without the
    filter the merged scene has cross-scene pairs; with it, its contacts are
    exactly the union of each part's own contacts."""
    from oracle.pyoracle import Oracle
    from parallel_admm_amd.configs import config_for

    cfg = config_for("cloth32")
    parts = _parts(pa)
    xs = [drape_on_obstacle(p, cfg.thickness, frac=0.3) for p in parts]
    merged = pa.Scene.merge(parts)
    seg = merged.segments()
    xm = np.zeros((merged.info["n_all"], 3))
    want = set()
    for s, (p, x) in enumerate(zip(parts, xs)):
        n = p.info["n_dyn"]
        d0, s0 = seg["dyn_off"][s], seg["stat_off"][s]
        xm[d0:d0 + n] = x[:n]
        xm[s0:s0 + len(x) - n] = x[n:]
        o = Oracle(p.mesh(), cfg)
        o.detect(x, brute_force=True)
        ints, _ = o.contacts()
        vmap = np.where(np.arange(len(x)) < n, d0 + np.arange(len(x)), s0 + np.arange(len(x)) - n)
        for row in ints:
            want.add((int(row[0]),) + tuple(int(vmap[v]) for v in row[3:7]))
        assert len(ints) > 0
    om = Oracle(merged.mesh(), cfg)
    om.detect(xm, brute_force=True)
    all_pairs = {(int(r[0]),) + tuple(int(v) for v in r[3:7]) for r in om.contacts()[0]}
    assert len(all_pairs) > len(want)  # cross-scene pairs exist without the filter
    om.set_segments(seg["seg"])
    for brute in (True, False):
        om.detect(xm, brute_force=brute)
        got = {(int(r[0]),) + tuple(int(v) for v in r[3:7]) for r in om.contacts()[0]}
        assert got == want
