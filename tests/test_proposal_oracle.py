"""Pins the CPU restatement of the proposal network (oracle/proposal.py), which
the GPU gradient test then uses as its reference."""
import numpy as np

from oracle import proposal as P


def _instance(h=8, seed=0):
    r = np.random.default_rng(seed)
    off = np.array([0, 3, 7, 8, 12])  # four polyhedron nodes, one with a single tetrahedron
    K = int(off[-1])
    params = r.normal(scale=0.5, size=P.n_params(K, h))
    x = r.uniform(-1, 1, size=(9, 3))
    labels = np.array([0, 2, 3, 6, 7, 8, 11, 5, 1])
    return params, x, labels, off, K, h


def test_gradient_matches_central_differences():
    params, x, labels, off, K, h = _instance()
    loss, g = P.loss_grad(params, x, labels, off, K, h)
    eps = 1e-6
    r = np.random.default_rng(1)
    idx = r.choice(params.size, size=60, replace=False)
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        fd = (P.loss_grad(p1, x, labels, off, K, h)[0] - P.loss_grad(p2, x, labels, off, K, h)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_closed_forms():
    params, x, labels, off, K, h = _instance()
    d = P.unpack(params, K, h)
    # zero output layer: uniform softmax over each node's candidates -> mean log(count)
    d["W3"][:] = 0
    d["b3"][:] = 0
    t0, t1 = P.candidates(off, labels)
    loss, g = P.loss_grad(P.pack(d), x, labels, off, K, h)
    assert np.isclose(loss, np.mean(np.log(t1 - t0)))
    # a point whose node has one tetrahedron contributes nothing
    only = labels == 7
    loss1, g1 = P.loss_grad(params, x[only], labels[only], off, K, h)
    assert loss1 == 0.0 and not np.any(g1)
    # parameter layout round trip and count
    assert P.n_params(K, h) == h * 3 + h + h * h + h + K * h + K
    np.testing.assert_array_equal(P.pack(P.unpack(params, K, h)), params)


def test_predict_is_argmax_within_the_node():
    params, x, labels, off, K, h = _instance()
    t0, t1 = P.candidates(off, labels)
    pred = P.predict(params, x, t0, t1, K, h)
    assert np.all((pred >= t0) & (pred < t1))
    # a large bias on one tetrahedron makes it the proposal of its node's points
    for i in range(len(labels)):
        d = P.unpack(params, K, h)
        d["b3"][labels[i]] += 1e6
        assert P.predict(P.pack(d), x[i:i + 1], t0[i:i + 1], t1[i:i + 1], K, h)[0] == labels[i]


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.0 ** -130, 65280.0, 3.0], np.float32)
    # ties go to even; 1 + 3*2^-8 rounds up to 1 + 2^-6; 65280 (1.9921875 * 2^15) is exact
    np.testing.assert_array_equal(P.to_bf16(x), np.array([1.0, 1.0, 1.0 + 2.0 ** -6, -2.0 ** -130, 65280.0, 3.0],
                                                          np.float32))


def test_bf16_gradient_emulation():
    """loss_grad_bf16 rounds h1, W2 and da2 to bf16 where the matrix cores
    take them.  With h1 and W2 already bf16-exact the forward pass, the loss
    and the gradients that do not pass through a rounded da2 (W3, b3, b2)
    equal loss_grad's; W2, W1 and b1 differ by at most bf16 resolution."""
    h = 8
    r = np.random.default_rng(5)
    off = np.array([0, 3, 7, 8, 12])
    K = int(off[-1])
    d = P.unpack(np.zeros(P.n_params(K, h)), K, h)
    # small integers times powers of two: h1 = W1 x + b1 and W2 stay exact in bf16
    d["W1"][:] = r.integers(-3, 4, size=d["W1"].shape) * 0.25
    d["b1"][:] = r.integers(-2, 3, size=d["b1"].shape) * 0.5
    d["W2"][:] = r.integers(-4, 5, size=d["W2"].shape) * 0.125
    d["b2"][:] = r.normal(size=d["b2"].shape)
    d["W3"][:] = r.normal(scale=0.3, size=d["W3"].shape)
    d["b3"][:] = r.normal(size=d["b3"].shape)
    params = P.pack(d)
    x = r.integers(-4, 5, size=(9, 3)) * 0.25
    labels = np.array([0, 2, 3, 6, 7, 8, 11, 5, 1])
    l0, g0 = P.loss_grad(params, x, labels, off, K, h)
    l1, g1 = P.loss_grad_bf16(params, x, labels, off, K, h)
    assert abs(l0 - l1) <= 1e-12
    G0, G1 = P.unpack(g0, K, h), P.unpack(g1, K, h)
    for k in ("W3", "b3", "b2"):
        np.testing.assert_allclose(G1[k], G0[k], rtol=0, atol=1e-12)
    for k in ("W2", "W1", "b1"):
        scale = max(np.abs(G0[k]).max(), 1e-30)
        assert np.abs(G1[k] - G0[k]).max() <= 2.0 ** -7 * scale * 4, k
        assert np.abs(G1[k] - G0[k]).max() > 0 or k == "b1"
