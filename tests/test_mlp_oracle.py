"""CPU checks that pin the warm-start oracle (oracle/mlp.py) before it judges
the device kernels: finite differences of the loss for the backpropagated
gradient, the closed form of Adam's first step, BF16 rounding against torch's
conversion, and one contact's features worked by hand."""
import numpy as np
import pytest

from oracle import mlp as om


def _net(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    w = np.zeros(om.NPARAM)
    for lo, hi, fan in ((om.OW1, om.OW2, om.IN), (om.OW2, om.OW3, om.H), (om.OW3, om.OW4, om.H), (om.OW4, om.NPARAM, om.H)):
        w[lo:hi] = rng.uniform(-1, 1, hi - lo) * scale / np.sqrt(fan)
    return w


def test_layout_size():
    assert om.NPARAM == 34690
    w = np.arange(om.NPARAM, dtype=np.float64)
    np.testing.assert_array_equal(om.pack(om.unpack(w)), w)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    w = _net(2)
    x = rng.standard_normal((7, om.IN))
    t = rng.standard_normal((7, om.OUT))
    loss, g = om.loss_and_grad(w, x, t)
    # indices in every weight block and every bias block
    idx = np.concatenate([rng.integers(lo, hi, 6) for lo, hi in
                          ((om.OW1, om.OB1), (om.OB1, om.OW2), (om.OW2, om.OB2), (om.OB2, om.OW3), (om.OW3, om.OB3),
                           (om.OB3, om.OW4), (om.OW4, om.OB4), (om.OB4, om.NPARAM))])
    eps = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        fd = (om.loss_and_grad(wp, x, t)[0] - om.loss_and_grad(wm, x, t)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(fd), (i, fd, g[i])
    # a planted mistake (the SiLU derivative taken as sigma) must fail the same check
    saved = om.dsilu
    try:
        om.dsilu = lambda a: 1.0 / (1.0 + np.exp(-a))
        _, g_bad = om.loss_and_grad(w, x, t)
    finally:
        om.dsilu = saved
    assert np.max(np.abs(g_bad[idx] - g[idx])) > 1e-4


def test_adam_first_step_closed_form():
    """With m_hat = g and v_hat = g^2 after one step, w1 = w0 - lr g / (|g| + eps)."""
    rng = np.random.default_rng(3)
    w0 = rng.standard_normal(50)
    g = rng.standard_normal(50) * 10.0 ** rng.uniform(-6, 1, 50)
    opt = om.Adam(w0, lr=1e-3)
    w1 = opt.step(g)
    np.testing.assert_allclose(w1, w0 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-15)
    # second step against a hand-expanded recurrence
    g2 = rng.standard_normal(50)
    m = 0.9 * 0.1 * g + 0.1 * g2
    v = 0.999 * 0.001 * g * g + 0.001 * g2 * g2
    w2 = opt.step(g2)
    np.testing.assert_allclose(w2, w1 - 1e-3 * (m / (1 - 0.81)) / (np.sqrt(v / (1 - 0.999 ** 2)) + 1e-8), rtol=1e-12)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    a = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.uniform(-8, 8, 20000),
                        [0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 65504.0, 1e-40]]).astype(np.float32)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(om.bf16(a), ref)
    assert om.bf16(np.float32(1.0 + 2 ** -8)) == 1.0             # tie to even
    assert om.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6


def test_forward_bf16_emulation_is_close_to_fp64():
    rng = np.random.default_rng(5)
    w = _net(6)
    x = rng.standard_normal((64, om.IN))
    y64, _ = om.forward(w, x)
    y16, _ = om.forward(w, x, emulate_bf16=True)
    assert np.max(np.abs(y16 - y64)) < 0.03 * (np.max(np.abs(y64)) + 0.1)


def test_features_of_one_contact_by_hand():
    """A point-triangle contact: dynamic vertex 0 above a static triangle (1, 2, 3)
    that moves with velocity (0.5, 0, 0); normal +y."""
    h, dhat, mu = 0.01, 1e-3, 0.3
    n_dyn = 1
    x0 = np.zeros((4, 3))
    x = np.array([[0.002, -0.003, 0.0], [0, 0, 0], [0, 0, 0], [0, 0, 0]], float)
    Uc = np.array([[0.0, 0.001, 0.0], [0, 0, 0], [0, 0, 0], [0, 0, 0]], float)
    p_static = np.array([[0, 0, 0], [0.5, 0, 0], [0.5, 0, 0], [0.5, 0, 0]], float)
    vel = np.array([[0.1, -0.2, 0.0], [0.5, 0, 0], [0.5, 0, 0], [0.5, 0, 0]], float)
    ints = np.array([[0, 0, 7, 0, 1, 2, 3]])
    reals = np.array([[1.0, -0.2, -0.3, -0.5, 0.0, 1.0, 0.0, -0.0004, 0.0006]])
    Kc = np.array([[0.0, -0.02, 0.0]])
    # u0 = p0_0 - (p_1 + p_2 + p_3)/... with weights (1, -0.2, -0.3, -0.5):
    #   p0_0 = (x - x0 + Uc)/h = (0.2, -0.2, 0), static sum = -1.0 * (0.5, 0, 0)
    #   u0 = (0.2 - 0.5, -0.2 - 0.02, 0) = (-0.3, -0.22, 0)
    # s = h/dhat = 10: f1 = 10 * -0.22, f2 = 10 * 0.3; tangent = (-1, 0, 0)
    # v_rel = (0.1 - 0.5, -0.2, 0) = (-0.4, -0.2, 0): f3 = -2, f4 = 4
    f, t, match = om.contact_features(h, dhat, mu, n_dyn, x, x0, Uc, p_static, vel, ints, reals, Kc)
    np.testing.assert_allclose(f[0], [-0.4, -2.2, 3.0, -2.0, 4.0, 0.3, 0.0, 0.0, 0.0, 0.0], atol=1e-12)
    np.testing.assert_allclose(t[0], [-1.0, 0.0, 0.0], atol=1e-15)
    assert match[0] == -1
    # the same contact in the previous window with dual (0.01, 0.02, 0): matched by (type, prim0, prim1)
    prev = (np.array([[0, 0, 5], [0, 0, 7], [1, 2, 3]]), np.array([[9, 9, 9], [0.01, 0.02, 0.0], [9, 9, 9]]))
    f, t, match = om.contact_features(h, dhat, mu, n_dyn, x, x0, Uc, p_static, vel, ints, reals, Kc, prev)
    assert match[0] == 1
    np.testing.assert_allclose(f[0, 7:], [1.0, 0.2, -0.1], atol=1e-12)
    # vanishing tangential velocity: tangent = n x (axis of n's smallest component) = (0,1,0) x (1,0,0)
    Kc2 = np.array([[0.3, -0.02, 0.0]])
    f, t, _ = om.contact_features(h, dhat, mu, n_dyn, x, x0, Uc, p_static, vel, ints, reals, Kc2)
    assert f[0, 2] == 0.0
    np.testing.assert_allclose(t[0], [0.0, 0.0, -1.0], atol=1e-15)
