"""CPU restatement of the option selector's network (test infrastructure
only: the tests compare the library's HIP kernels with it; no product path
imports it).

Layers in -> h -> h -> h -> out with ReLU between them, parameters packed
per layer as (W_l row-major I_l x J_l, b_l); loss = sum M (Y - T)^2 /
sum M; Adam (Kingma & Ba, "Adam: A Method for Stochastic Optimization",
2015) with beta1 0.9, beta2 0.999, eps 1e-8.  Computed in float64 (the
kernels run in float32, so comparisons carry a float32 tolerance).

Pinned by a central finite-difference check of the gradient (tests), which
a dropped term, a wrong sign or a transposed operand fails.
"""
import numpy as np


def layout(n_in, n_hidden, n_out):
    dims = [n_in, n_hidden, n_hidden, n_hidden, n_out]
    out, off = [], 0
    for l in range(4):
        I, J = dims[l], dims[l + 1]
        out.append((off, I, J, off + I * J))
        off += I * J + J
    return out, off


def unpack(p, n_in, n_hidden, n_out):
    lay, _ = layout(n_in, n_hidden, n_out)
    return [(p[w:w + I * J].reshape(I, J), p[b:b + J]) for (w, I, J, b) in lay]


def forward(p, X, n_hidden, n_out):
    Ws = unpack(np.asarray(p, dtype=np.float64), X.shape[1], n_hidden, n_out)
    a = np.asarray(X, dtype=np.float64)
    zs, acts = [], [a]
    for l, (W, b) in enumerate(Ws):
        z = a @ W + b
        a = np.maximum(z, 0.0) if l < 3 else z
        zs.append(z)
        acts.append(a)
    return a, zs, acts


def loss_and_grad(p, X, T, M, n_hidden, n_out):
    Y, zs, acts = forward(p, X, n_hidden, n_out)
    M = np.asarray(M, dtype=np.float64)
    cnt = M.sum()
    inv = 1.0 / cnt if cnt > 0 else 0.0
    d = M * (Y - np.asarray(T, dtype=np.float64))
    loss = float((d * (Y - T)).sum() * inv)
    dA = 2.0 * d * inv
    lay, total = layout(X.shape[1], n_hidden, n_out)
    Ws = unpack(np.asarray(p, dtype=np.float64), X.shape[1], n_hidden, n_out)
    g = np.zeros(total)
    for l in range(3, -1, -1):
        dZ = dA if l == 3 else dA * (zs[l] > 0)
        w, I, J, b = lay[l]
        g[w:w + I * J] = (acts[l].T @ dZ).reshape(-1)
        g[b:b + J] = dZ.sum(axis=0)
        dA = dZ @ Ws[l][0].T
    return loss, g


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
