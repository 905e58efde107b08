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
    # with BF16 operands the first bias enters the product as a pair of BF16 weights on inputs of 1:
    # bf16(b1) + bf16(b1 - bf16(b1)), 16 significant bits
    b1e = bf16(b1) + bf16(b1 - bf16(b1)) if bf else b1
    z1 = r(x) @ r(W1).T + b1e
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


def closure_features(P, nu, tau0, smag):
    """The closure's 6 features (the kernels' mlp_features) of stresses P (n x 6: xx yy zz xy xz yz), in FP64."""
    P = np.asarray(P, dtype=np.float64)
    pxx, pyy, pzz, pxy, pxz, pyz = (P[:, c] for c in range(6))
    n2 = pxx ** 2 + pyy ** 2 + pzz ** 2 + 2 * (pxy ** 2 + pxz ** 2 + pyz ** 2)
    pn = np.sqrt(n2)
    tr = pxx + pyy + pzz
    m = tr / 3
    dxx, dyy, dzz = pxx - m, pyy - m, pzz - m
    d2 = dxx ** 2 + dyy ** 2 + dzz ** 2 + 2 * (pxy ** 2 + pxz ** 2 + pyz ** 2)
    det = dxx * (dyy * dzz - pyz ** 2) - pxy * (pxy * dzz - pyz * pxz) + pxz * (pxy * pyz - dyy * pxz)
    tau_s = 0.5 * (tau0 + np.sqrt(tau0 ** 2 + smag * pn))
    return np.stack([np.log(pn + 1e-12) * 0.1 + 1, tr / (pn + 1e-20), d2 / (n2 + 1e-30),
                     np.clip(np.sqrt(6) * 3 * det / (d2 * np.sqrt(d2) + 1e-30), -1, 1),
                     np.full_like(pn, np.log(nu) * 0.1 + 1), np.log1p(np.maximum(0, (tau_s - tau0) / 3) / nu) / 4], 1)


def sgs_samples(ux, uy, uz, flags, fine_dims, stride, seed, nu, cs):
    """Training samples the solver draws from a fine-grid flow (aero_mlp_sample_fine), restated in FP64 numpy.

    The fine velocity is averaged over 2 x 2 x 2 blocks onto the coarse grid (blocks of plain fluid cells only);
    on a coarse sampling lattice of the given stride (offset from the seed) the strain S comes from central
    differences of the block means, the sub-filter stress tau = <(u - <u>)(u - <u>)> (deviatoric part) gives the
    resolved-scale viscosity nu_r = -tau : S / (2 S : S), and with the fine grid's own Smagorinsky viscosity nu_f =
    (smag / 72) |S| the target is nu_t = nu_r + nu_f. The features are those of the coarse non-equilibrium stress
    P = -(2/3) tau_t S of the relaxation time tau_t = clamp(3 (nu + max(0, nu_t)) + 1/2), the target the factor
    nu_t / nu_S on the Smagorinsky viscosity for P (clipped to [-8, 16]), and feature column 7 the loss weight
    (nu_S |S| / nu)^2 (clipped to [1e-12, 1e6]). Returns (X n x 8, T n, details) for the valid lattice points in
    lattice order, details holding their coarse indices and the intermediates (block mean, sub-filter stress, S,
    nu_S)."""
    fx, fy, fz = fine_dims
    cx, cy, cz = fx // 2, fy // 2, fz // 2
    u = [np.asarray(a, dtype=np.float64).reshape(fz, fy, fx) for a in (ux, uy, uz)]
    fl = np.asarray(flags).reshape(fz, fy, fx)
    blk = lambda a: a[:2 * cz, :2 * cy, :2 * cx].reshape(cz, 2, cy, 2, cx, 2)
    fluid = (blk(fl) == 0).all(axis=(1, 3, 5))
    mean = [blk(a).mean(axis=(1, 3, 5)) for a in u]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    # central second moments of each block (the kernel forms them from the deviations from the block mean)
    dev = [blk(a) - m[:, None, :, None, :, None] for a, m in zip(u, mean)]
    tau6 = [(dev[a] * dev[b]).mean(axis=(1, 3, 5)) for a, b in pairs]
    tau0, smag = 3 * nu + 0.5, 18 * np.sqrt(2) * cs * cs
    ox, oy, oz = 1 + seed % stride, 1 + (seed // stride) % stride, 1 + (seed // (stride * stride)) % stride
    X, T, det = [], [], []
    for k in range(oz, cz - 1, stride):
        for j in range(oy, cy - 1, stride):
            for i in range(ox, cx - 1, stride):
                if not (fluid[k, j, i] and fluid[k, j, i + 1] and fluid[k, j, i - 1] and fluid[k, j + 1, i] and
                        fluid[k, j - 1, i] and fluid[k + 1, j, i] and fluid[k - 1, j, i]):
                    continue
                # g[a][b] = d u_a / d x_b, one coarse cell per unit
                g = np.array([[0.5 * (mean[a][k, j, i + 1] - mean[a][k, j, i - 1]),
                               0.5 * (mean[a][k, j + 1, i] - mean[a][k, j - 1, i]),
                               0.5 * (mean[a][k + 1, j, i] - mean[a][k - 1, j, i])] for a in range(3)])
                S = 0.5 * (g + g.T)
                SS = float(np.sum(S * S))
                if not SS > 1e-14:
                    continue
                u0 = np.array([mean[a][k, j, i] for a in range(3)])
                t = np.empty((3, 3))
                for (a, b), v in zip(pairs, tau6):
                    t[a, b] = t[b, a] = v[k, j, i]
                tsub = t.copy()
                t -= np.eye(3) * np.trace(t) / 3
                nu_r = -float(np.sum(t * S)) / (2 * SS)
                nut = nu_r + 0.25 * (smag / 18) * np.sqrt(SS)
                tau = min(1.95, max(0.5015, 3 * (nu + max(0.0, nut)) + 0.5))
                Pm = -(2 / 3) * tau * S
                P6 = [Pm[0, 0], Pm[1, 1], Pm[2, 2], Pm[0, 1], Pm[0, 2], Pm[1, 2]]
                pn = np.sqrt(Pm[0, 0] ** 2 + Pm[1, 1] ** 2 + Pm[2, 2] ** 2 + 2 * (Pm[0, 1] ** 2 + Pm[0, 2] ** 2 + Pm[1, 2] ** 2))
                nu_s = (0.5 * (tau0 + np.sqrt(tau0 ** 2 + smag * pn)) - tau0) / 3
                if not nu_s > 1e-12:
                    continue
                x = np.zeros(8)
                x[:6] = closure_features(np.array([P6]), nu, tau0, smag)[0]
                x[7] = min(1e6, max(1e-12, (nu_s * np.sqrt(SS) / nu) ** 2))
                X.append(x)
                T.append(min(16.0, max(-8.0, nut / nu_s)))
                det.append({"ijk": (i, j, k), "S": S, "SS": SS, "tau": tsub, "u0": u0, "nu_s": nu_s, "nut": nut})
    return np.array(X).reshape(-1, 8), np.array(T), det
