"""FP64 numpy oracle of the learned sweep cost model -- TEST INFRASTRUCTURE ONLY.

The reference has no learned component (its planner is q1tsim's per-gate
loop, src/circuit.rs:643-762), so this restates the model's own mathematics:
a 20 -> 64 -> 64 -> 1 multilayer perceptron with tanh hidden layers on
standardised features, the mean squared error against targets, its gradient
by backpropagation, and one Adam step (Kingma & Ba 2015, Algorithm 1, with
bias-corrected moments).  Parameter layout (one flat vector): W1[64][20],
b1[64], W2[64][64], b2[64], w3[64], b3.

The gradient is pinned in tests/test_oracle_mlp.py by central finite
differences of ``loss`` and by closed forms (all-zero weights, a linear
network); Adam by its first step, whose update is -lr*sign(g) up to eps.

Only ``tests/`` may import this module, and only as the checker.
"""
import numpy as np

F, H = 20, 64
W1 = 0
B1 = W1 + H * F
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H
NPARAMS = B3 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1].reshape(H, F), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3], p[B3])


def forward(p, x, mu, sigma):
    """Outputs f(x) for rows x (n x 20), with the hidden activations."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    xn = (np.asarray(x, dtype=np.float64) - mu) / sigma
    h1 = np.tanh(xn @ w1.T + b1)
    h2 = np.tanh(h1 @ w2.T + b2)
    y = h2 @ w3 + b3
    return y, (xn, h1, h2)


def loss(p, x, t, mu, sigma):
    y, _ = forward(p, x, mu, sigma)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def grad(p, x, t, mu, sigma):
    """(loss, dloss/dp) by backpropagation, FP64."""
    w1, b1, w2, b2, w3, b3 = unpack(p)
    y, (xn, h1, h2) = forward(p, x, mu, sigma)
    n = len(y)
    e = y - np.asarray(t, dtype=np.float64)
    g = 2.0 * e / n                       # dL/dy
    d2 = np.outer(g, w3) * (1.0 - h2 ** 2)    # dL/d(pre-activation 2)
    d1 = (d2 @ w2) * (1.0 - h1 ** 2)          # dL/d(pre-activation 1)
    out = np.zeros(NPARAMS)
    out[W1:B1] = (d1.T @ xn).ravel()
    out[B1:W2] = d1.sum(0)
    out[W2:B2] = (d2.T @ h1).ravel()
    out[B2:W3] = d2.sum(0)
    out[W3:B3] = h2.T @ g
    out[B3] = g.sum()
    return float(np.mean(e ** 2)), out


def adam_step(p, m, v, g, step, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update at step number `step` (1-based); returns (p, m, v)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
