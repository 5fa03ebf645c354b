"""FP64 numpy restatement of the two learned networks (test infrastructure).

Kepler warm-start MLP: 4 -> 32 -> 32 -> 1, tanh, linear output, loss
0.5 (y - t)^2.  Encounter scorer: 6 -> 16 -> 1, tanh, logit output, loss
BCE-with-logits.  Flat parameter layouts match the library's (weights of each
layer row-major [out][in], then biases)."""
import numpy as np

KM = dict(IN=4, H=32)
SC = dict(IN=6, H=16)


def _split_kepler(w):
    w = np.asarray(w, np.float64)
    i, H, I = 0, KM["H"], KM["IN"]
    W1 = w[i:i + H * I].reshape(H, I); i += H * I
    b1 = w[i:i + H]; i += H
    W2 = w[i:i + H * H].reshape(H, H); i += H * H
    b2 = w[i:i + H]; i += H
    W3 = w[i:i + H]; i += H
    b3 = w[i]
    return W1, b1, W2, b2, W3, b3


def _split_scorer(w):
    w = np.asarray(w, np.float64)
    i, H, I = 0, SC["H"], SC["IN"]
    W1 = w[i:i + H * I].reshape(H, I); i += H * I
    b1 = w[i:i + H]; i += H
    W2 = w[i:i + H]; i += H
    b2 = w[i]
    return W1, b1, W2, b2


def nparams(which):
    return 1249 if which == 0 else 129


def forward(which, w, X):
    X = np.asarray(X, np.float64)
    if which == 0:
        W1, b1, W2, b2, W3, b3 = _split_kepler(w)
        h1 = np.tanh(X @ W1.T + b1)
        h2 = np.tanh(h1 @ W2.T + b2)
        return h2 @ W3 + b3
    W1, b1, W2, b2 = _split_scorer(w)
    h = np.tanh(X @ W1.T + b1)
    return h @ W2 + b2


def loss(which, w, X, t):
    y = forward(which, w, X)
    t = np.asarray(t, np.float64)
    if which == 0:
        return float(np.mean(0.5 * (y - t) ** 2))
    return float(np.mean(np.maximum(y, 0) - y * t + np.log1p(np.exp(-np.abs(y)))))


def grad(which, w, X, t):
    """Analytic gradient of the mean loss (backpropagation in FP64)."""
    X = np.asarray(X, np.float64)
    t = np.asarray(t, np.float64)
    n = len(t)
    if which == 0:
        W1, b1, W2, b2, W3, b3 = _split_kepler(w)
        h1 = np.tanh(X @ W1.T + b1)
        h2 = np.tanh(h1 @ W2.T + b2)
        e = (h2 @ W3 + b3 - t) / n
        gW3 = h2.T @ e
        gb3 = e.sum()
        d2 = np.outer(e, W3) * (1 - h2 ** 2)
        gW2 = d2.T @ h1
        gb2 = d2.sum(0)
        d1 = (d2 @ W2) * (1 - h1 ** 2)
        gW1 = d1.T @ X
        gb1 = d1.sum(0)
        return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    W1, b1, W2, b2 = _split_scorer(w)
    h = np.tanh(X @ W1.T + b1)
    y = h @ W2 + b2
    e = (1 / (1 + np.exp(-y)) - t) / n
    gW2 = h.T @ e
    gb2 = e.sum()
    d = np.outer(e, W2) * (1 - h ** 2)
    return np.concatenate([(d.T @ X).ravel(), d.sum(0), gW2, [gb2]])


def fd_grad(which, w, X, t, idx, eps=1e-6):
    """Central finite differences of the mean loss for parameters idx."""
    w = np.asarray(w, np.float64).copy()
    out = []
    for i in idx:
        w0 = w[i]
        w[i] = w0 + eps
        lp = loss(which, w, X, t)
        w[i] = w0 - eps
        lm = loss(which, w, X, t)
        w[i] = w0
        out.append((lp - lm) / (2 * eps))
    return np.array(out)
