"""CPU reference of the learned preconditioner's network and loss (test
infrastructure; the library's kernels are in csrc/precond_mlp.hip).

Network: x[4] -> tanh(x W1 + b1)[32] -> tanh(h1 W2 + b2)[32] -> y = h2 w3 + b3;
K = exp(min(y, 12)); loss L = sum A K^2 - 2 B K.  Parameters in the order
W1[4][32], b1[32], W2[32][32] (input row, output column), b2[32], w3[32], b3.
Everything in float64 here."""
import numpy as np

NF, NH = 4, 32
NP = NF * NH + NH + NH * NH + NH + NH + 1
YMAX = 12.0


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + NF * NH].reshape(NF, NH); o += NF * NH
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    w3 = p[o:o + NH]; o += NH
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(X, p):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64).reshape(-1, NF)
    h1 = np.tanh(X @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    y = h2 @ w3 + b3
    return y, h1, h2


def loss(X, A, B, p):
    y, _, _ = forward(X, p)
    K = np.exp(np.minimum(y, YMAX))
    return float(np.sum(A * K * K - 2.0 * B * K))


def loss_grad(X, A, B, p):
    """analytic gradient by backpropagation"""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64).reshape(-1, NF)
    y, h1, h2 = forward(X, p)
    K = np.exp(np.minimum(y, YMAX))
    dy = np.where(y < YMAX, 2.0 * K * (A * K - B), 0.0)
    z2 = dy[:, None] * w3[None, :] * (1.0 - h2 * h2)
    z1 = (z2 @ W2.T) * (1.0 - h1 * h1)
    g = np.concatenate([(X.T @ z1).reshape(-1), z1.sum(0), (h1.T @ z2).reshape(-1), z2.sum(0),
                        h2.T @ dy, [dy.sum()]])
    return float(np.sum(A * K * K - 2.0 * B * K)), g


def init_params(seed):
    """Glorot-uniform weights, zero biases; b3 starts at 0 (K = 1, the
    reference's step)"""
    rng = np.random.default_rng(seed)
    def glorot(a, b):
        lim = np.sqrt(6.0 / (a + b))
        return rng.uniform(-lim, lim, size=(a, b))
    return np.concatenate([glorot(NF, NH).reshape(-1), np.zeros(NH), glorot(NH, NH).reshape(-1),
                           np.zeros(NH), glorot(NH, 1).reshape(-1), [0.0]]).astype(np.float32)
