"""CPU oracle of the learned closure (TEST INFRASTRUCTURE ONLY): forward pass, loss and backward pass of the
6 -> 32 -> 32 -> 1 ReLU MLP in FP64 numpy, optionally rounding the matrix-multiply operands to BF16 exactly
where the GPU kernels do (mlp.h / mlp_train.hip), so that GPU gradients can be compared tightly.
Pinned by finite differences of the unrounded loss (tests/test_mlp_oracle.py)."""
import numpy as np

H, IN = 32, 6
NPARAM = H * IN + H + H * H + H + H + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, w3, [b3]])


def bf16(x):
    """Round to bfloat16 (round to nearest even) and return as float64."""
    a = np.asarray(x, dtype=np.float32).copy()
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def init_params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((H, IN)) * scale
    b1 = rng.standard_normal(H) * 0.1
    W2 = rng.standard_normal((H, H)) * scale / np.sqrt(H / IN)
    b2 = rng.standard_normal(H) * 0.1
    w3 = rng.standard_normal(H) * scale
    return pack(W1, b1, W2, b2, w3, 0.0)


def softplus(y):
    return np.where(y > 20, y, np.log1p(np.exp(np.minimum(y, 20))))


def forward(p, X, bf=True):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    x = np.asarray(X, dtype=np.float64)[:, :IN]
    z1 = r(x) @ r(W1).T + b1
    h1 = np.maximum(z1, 0)
    z2 = r(h1) @ r(W2).T + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ w3 + b3
    return y, (x, z1, h1, z2, h2)


C0 = np.log(np.e - 1.0)  # softplus(C0) = 1: the zero network gives the Smagorinsky viscosity


def factor(y):
    """Closure factor on the Smagorinsky eddy viscosity for network output y."""
    return softplus(np.asarray(y, dtype=np.float64) + C0)


def sample_weights(X):
    """Loss weight of each sample: feature column 7 (never read by the first layer); 0 stands for 1."""
    w = np.asarray(X, dtype=np.float64)[:, 7]
    return np.where(w > 0, w, 1.0)


def loss_and_grad(p, X, T, bf=True):
    """Mean of w (factor(y) - t)^2 and its gradient."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    r = bf16 if bf else (lambda v: np.asarray(v, dtype=np.float64))
    y, (x, z1, h1, z2, h2) = forward(p, X, bf)
    n = len(y)
    w = sample_weights(X)
    d = factor(y) - np.asarray(T, dtype=np.float64)
    L = np.mean(w * d * d)
    sig = 1.0 / (1.0 + np.exp(-(y + C0)))
    g = 2.0 * w * d * sig / n
    dw3 = h2.T @ g
    db3 = g.sum()
    dz2 = np.outer(g, w3) * (z2 > 0)
    db2 = dz2.sum(0)
    dW2 = r(dz2).T @ r(h1)
    dh1 = r(dz2) @ r(W2)
    dz1 = dh1 * (z1 > 0)
    db1 = dz1.sum(0)
    dW1 = r(dz1).T @ r(x)
    return L, pack(dW1, db1, dW2, db2, dw3, db3)


def predict_factor(p, X, bf=True):
    """nu_t / nu_S predicted by the network."""
    y, _ = forward(p, X, bf)
    return factor(y)
