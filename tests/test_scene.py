"""Scene setup (host code of the product library) against the sizes quoted for
the benchmark configurations, closed-form rest-state identities, and the
oracle's independent topology/constraint/mass construction.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp


def _product_D(scene):
    """D (m x n) rebuilt in numpy from the product's flattened constraints:
    tri rows G = S Binv, tet rows G = S Binv (S = [-1 -1 ..; I]), bend rows from
    the ring coefficients, pin rows = 1."""
    c = scene.constraints()
    n = scene.info["n_dyn"]
    rows, cols, vals = [], [], []
    for k in range(c["n_tri"]):
        B = c["tri_binv"][k]
        G = np.vstack([-(B[0] + B[1]), B[0], B[1]])  # 3 x 2
        for r in range(2):
            rows += [c["tri_row0"] + 2 * k + r] * 3
            cols += list(c["tri_inds"][k])
            vals += list(G[:, r])
    for k in range(c["n_tet"]):
        B = c["tet_binv"][k]
        G = np.vstack([-(B[0] + B[1] + B[2]), B[0], B[1], B[2]])
        for r in range(3):
            rows += [c["tet_row0"] + 3 * k + r] * 4
            cols += list(c["tet_inds"][k])
            vals += list(G[:, r])
    for k in range(c["n_bend"]):
        a, b = c["bend_ptr"][k], c["bend_ptr"][k + 1]
        rows += [c["bend_row0"] + k] * (b - a)
        cols += list(c["bend_inds"][a:b])
        vals += list(c["bend_coef"][a:b])
    for k in range(c["n_pin"]):
        rows.append(c["pin_row0"] + k)
        cols.append(c["pin_inds"][k])
        vals.append(1.0)
    m = scene.info["m_rows"]
    return sp.csr_matrix((vals, (rows, cols)), shape=(m, n)), c


def test_cloth32_sizes(pa):
    s = pa.Scene.generate(1)
    i = s.info
    assert (i["n_dyn"], i["n_all"]) == (1024, 1024 + 642)
    assert (i["n_tri_c"], i["n_bend_c"], i["n_pin_c"], i["m_rows"]) == (1922, 1022, 2, 4868)
    assert i["nnz_d"] == 18442
    assert i["n_tri"] - i["n_tri_c"] == 1280  # level-3 icosphere
    m = s.mesh()
    np.testing.assert_allclose(np.linalg.norm(m["x"][1024:], axis=1), 0.2, rtol=1e-12)
    assert s.mass()[:1024].sum() == pytest.approx(0.1, rel=1e-12)
    cloth_edges = {tuple(e) for e in s.topology()["surf_edges"] if e[1] < 1024}
    assert len(cloth_edges) == 2945
    np.testing.assert_array_equal(m["pins"], [0, 31 * 32])


def test_tetbox_sizes(pa):
    s = pa.Scene.generate(2)
    i = s.info
    assert (i["n_dyn"], i["n_tet_c"], i["m_rows"], i["nnz_d"]) == (9261, 40000, 120000, 480000)
    assert i["n_tet_boundary"] == 4800
    c = s.constraints()
    vol = np.array([abs(np.linalg.det(np.linalg.inv(B))) / 6 for B in c["tet_binv"]])
    assert vol.sum() == pytest.approx(0.2 ** 3, rel=1e-10)
    assert s.mass()[:9261].sum() == pytest.approx(1000 * 0.2 ** 3, rel=1e-10)


def test_cloth256_sizes(pa):
    s = pa.Scene.generate(3)
    i = s.info
    assert (i["n_dyn"], i["n_tri_c"], i["n_bend_c"], i["m_rows"], i["nnz_d"]) == (65536, 130050, 65534, 325634, 1237000)
    cloth_edges = [e for e in s.topology()["surf_edges"] if e[1] < 65536]
    assert len(cloth_edges) == 195585


@pytest.mark.parametrize("kind", [1, 2])
def test_rest_state_deformation_is_orthonormal(pa, kind):
    """At rest D x gives F with orthonormal columns (tri: the in-plane basis,
    tet: the identity), so every elastic prox is a fixed point."""
    s = pa.Scene.generate(kind)
    D, c = _product_D(s)
    x = s.mesh()["x"][: s.info["n_dyn"]]
    F = D @ x
    for k in range(0, c["n_tri"], 37):
        Fk = F[c["tri_row0"] + 2 * k: c["tri_row0"] + 2 * k + 2].T
        np.testing.assert_allclose(Fk.T @ Fk, np.eye(2), atol=1e-12)
    for k in range(0, c["n_tet"], 97):
        Fk = F[c["tet_row0"] + 3 * k: c["tet_row0"] + 3 * k + 3].T
        np.testing.assert_allclose(Fk, np.eye(3), atol=1e-12)
    # every elastic row (all but the pins) annihilates rigid translations
    t = D @ np.ones(D.shape[1])
    np.testing.assert_allclose(t[: c["pin_row0"]], 0.0, atol=1e-9 * np.abs(D).max())
    np.testing.assert_allclose(t[c["pin_row0"]:], 1.0)


@pytest.mark.parametrize("kind", [1, 2])
def test_scene_matches_oracle_construction(pa, kind):
    """The oracle builds topology, constraints and mass on its own from the raw mesh."""
    from oracle.pyoracle import Oracle
    from parallel_admm_amd.configs import config_for

    s = pa.Scene.generate(kind)
    o = Oracle(s.mesh(), config_for("cloth32" if kind == 1 else "tetbox"))
    tp, to = s.topology(), o.topology()
    for k in ("surf_verts", "surf_tris", "surf_edges"):
        np.testing.assert_array_equal(tp[k], to[k])
    np.testing.assert_allclose(s.mass(), o.mass(), rtol=1e-14)
    Dp, c = _product_D(s)
    Do = o.matrix(0)
    assert (Dp - Do).count_nonzero() == 0 or abs(Dp - Do).max() < 1e-9 * abs(Do).max()
    w = np.concatenate([np.repeat(c["tri_w"], 2), np.repeat(c["tet_w"], 3), c["bend_w"], c["pin_w"]])
    np.testing.assert_allclose(w, o.row_weights(), rtol=1e-14)


def test_scene_create_rejects_bad_input(pa):
    x = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 1, 0]], float)
    dens = np.ones(4)
    # collinear triangle -> geometry error
    with pytest.raises(pa.PAError) as e:
        pa.Scene.from_mesh(x, 4, dens, tris=[[0, 1, 2]], tri_k=[1.0], bend_k=np.zeros(4))
    assert e.value.code == pa.PA_ERR_GEOMETRY
    # index out of range -> invalid
    with pytest.raises(pa.PAError) as e:
        pa.Scene.from_mesh(x, 4, dens, tris=[[0, 1, 7]], tri_k=[1.0], bend_k=np.zeros(4))
    assert e.value.code == pa.PA_ERR_INVALID
    # a dynamic vertex that no element gives mass to is rejected
    with pytest.raises(pa.PAError) as e:
        pa.Scene.from_mesh(x, 4, dens, tris=[[0, 1, 3]], tri_k=[1.0], bend_k=np.zeros(4))
    assert e.value.code == pa.PA_ERR_GEOMETRY
    # a valid single triangle works: 2 rows, no bending (ring < 3)
    s = pa.Scene.from_mesh(x[[0, 1, 3]], 3, dens[:3], tris=[[0, 1, 2]], tri_k=[1.0], bend_k=np.zeros(3))
    assert s.info["m_rows"] == 2 and s.info["n_bend_c"] == 0
