"""CPU restatement of the learned tetrahedron proposal network (test
infrastructure only: tests/ use it to check the GPU forward and backward
kernels; nothing in the product imports it).

The network has no counterpart in the reference (it only proposes which
tetrahedron PolyNode::indices_weights, src/trellis_node.hpp:275-308, would
find; the answer is always verified in FP64 and falls back to the exact scan).

    h1 = relu(W1 x + b1)          x: 3 features, W1: H x 3
    h2 = relu(W2 h1 + b2)         W2: H x H
    z_t = W3[t] . h2 + b3[t]      for the tetrahedra t of the point's polyhedron node
    loss = mean over points of -log softmax(z)[label]

Flat parameter order (float32): W1 (H x 3, row-major), b1 (H), W2 (H x H,
row = output unit), b2 (H), W3 (K x H), b3 (K); K = number of tetrahedra.
"""
import numpy as np

H = 128


def n_params(K, h=H):
    return h * 3 + h + h * h + h + K * h + K


def unpack(p, K, h=H):
    p = np.array(p, dtype=np.float64)  # a copy: the views below must not alias the caller
    o = 0
    out = {}
    for name, shape in (("W1", (h, 3)), ("b1", (h,)), ("W2", (h, h)), ("b2", (h,)), ("W3", (K, h)), ("b3", (K,))):
        size = int(np.prod(shape))
        out[name] = p[o:o + size].reshape(shape)
        o += size
    return out


def pack(d):
    return np.concatenate([d[k].ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3")])


def candidates(poly_tet_offset, labels):
    """[t0, t1) of the polyhedron node holding each label tetrahedron."""
    off = np.asarray(poly_tet_offset, dtype=np.int64)
    node = np.searchsorted(off, np.asarray(labels, dtype=np.int64), side="right") - 1
    return off[node], off[node + 1]


def loss_grad(params, x, labels, poly_tet_offset, K, h=H):
    """Mean loss over the points and the gradient of every parameter (float64)."""
    P = unpack(params, K, h)
    x = np.asarray(x, dtype=np.float64)
    labels = np.asarray(labels, dtype=np.int64)
    t0, t1 = candidates(poly_tet_offset, labels)
    n = x.shape[0]
    a1 = x @ P["W1"].T + P["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ P["W2"].T + P["b2"]
    h2 = np.maximum(a2, 0.0)
    G = {k: np.zeros_like(v) for k, v in P.items()}
    dh2 = np.zeros_like(h2)
    loss = 0.0
    for i in range(n):
        c = np.arange(t0[i], t1[i])
        z = P["W3"][c] @ h2[i] + P["b3"][c]
        m = z.max()
        e = np.exp(z - m)
        s = e.sum()
        lab = labels[i] - t0[i]
        loss += -(z[lab] - m - np.log(s))
        dz = e / s
        dz[lab] -= 1.0
        dz /= n
        G["W3"][c] += np.outer(dz, h2[i])
        G["b3"][c] += dz
        dh2[i] = dz @ P["W3"][c]
    da2 = dh2 * (a2 > 0)
    G["W2"] = da2.T @ h1
    G["b2"] = da2.sum(axis=0)
    dh1 = da2 @ P["W2"]
    da1 = dh1 * (a1 > 0)
    G["W1"] = da1.T @ x
    G["b1"] = da1.sum(axis=0)
    return loss / n, pack(G)


def loss_grad_bf16(params, x, labels, poly_tet_offset, K, h=H):
    """loss_grad as the training kernels compute it: the three 128-wide
    products take bf16 operands (h1 and W2 in the forward product, da2 and W2
    in the backward one, da2 and h1 in the W2 gradient), each rounded from its
    FP32 value (nearest even); the rest as in loss_grad.  Sums in float64."""
    P = unpack(params, K, h)
    x32 = np.asarray(x, dtype=np.float32)
    labels = np.asarray(labels, dtype=np.int64)
    t0, t1 = candidates(poly_tet_offset, labels)
    n = x32.shape[0]
    W1 = P["W1"].astype(np.float32)
    a1 = (x32[:, :1] * W1[:, 0] + x32[:, 1:2] * W1[:, 1] + x32[:, 2:3] * W1[:, 2] + P["b1"].astype(np.float32))
    h1 = np.maximum(a1, 0).astype(np.float32)
    h1b = to_bf16(h1).astype(np.float64)
    W2b = to_bf16(P["W2"].astype(np.float32)).astype(np.float64)
    a2 = h1b @ W2b.T + P["b2"]
    h2 = np.maximum(a2, 0.0)
    G = {k: np.zeros_like(v) for k, v in P.items()}
    dh2 = np.zeros_like(h2)
    loss = 0.0
    for i in range(n):
        c = np.arange(t0[i], t1[i])
        z = P["W3"][c] @ h2[i] + P["b3"][c]
        m = z.max()
        e = np.exp(z - m)
        s_ = e.sum()
        lab = labels[i] - t0[i]
        loss += -(z[lab] - m - np.log(s_))
        dz = e / s_
        dz[lab] -= 1.0
        dz /= n
        G["W3"][c] += np.outer(dz, h2[i])
        G["b3"][c] += dz
        dh2[i] = dz @ P["W3"][c]
    da2 = dh2 * (a2 > 0)
    da2b = to_bf16(da2.astype(np.float32)).astype(np.float64)
    G["W2"] = da2b.T @ h1b
    G["b2"] = da2.sum(axis=0)
    dh1 = da2b @ W2b
    da1 = dh1 * (h1 > 0)
    G["W1"] = da1.T @ x32.astype(np.float64)
    G["b1"] = da1.sum(axis=0)
    return loss / n, pack(G)


def predict(params, x, t0, t1, K, h=H):
    """Proposed tetrahedron (argmax over the node's candidates; first on ties)."""
    P = unpack(params, K, h)
    h1 = np.maximum(np.asarray(x, np.float64) @ P["W1"].T + P["b1"], 0.0)
    h2 = np.maximum(h1 @ P["W2"].T + P["b2"], 0.0)
    out = np.empty(len(t0), np.int64)
    for i in range(len(t0)):
        c = np.arange(t0[i], t1[i])
        out[i] = c[np.argmax(P["W3"][c] @ h2[i] + P["b3"][c])]
    return out


def to_bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def predict_bf16(params, x, t0, t1, K, h=H):
    """The matrix-core inference: FP32 first layer rounded to bf16, bf16 W2 and
    W3 with FP32 accumulation.  Returns (proposal, logits gap to the runner-up)."""
    P = unpack(params, K, h)
    W1 = P["W1"].astype(np.float32)
    b1 = P["b1"].astype(np.float32)
    x = np.asarray(x, np.float32)
    a1 = x[:, :1] * W1[:, 0] + x[:, 1:2] * W1[:, 1] + x[:, 2:3] * W1[:, 2] + b1
    h1 = to_bf16(np.maximum(a1, 0)).astype(np.float64)
    h2 = np.maximum(h1 @ to_bf16(P["W2"].astype(np.float32)).astype(np.float64).T + P["b2"], 0.0)
    W3 = to_bf16(P["W3"].astype(np.float32)).astype(np.float64)
    out = np.empty(len(t0), np.int64)
    gap = np.empty(len(t0))
    for i in range(len(t0)):
        c = np.arange(t0[i], t1[i])
        z = W3[c] @ h2[i] + P["b3"][c]
        k = int(np.argmax(z))
        out[i] = c[k]
        zs = np.sort(z)
        gap[i] = zs[-1] - zs[-2] if len(z) > 1 else np.inf
    return out, gap
