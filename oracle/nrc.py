"""CPU (numpy, float64) restatement of the neural radiance cache.

TEST INFRASTRUCTURE: the checker for kiraray_amd's BF16 matrix-core network
(kiraray_amd/csrc/nrc.hip, include/krr_nrc.h).  Only tests/, the smoke
check and the benchmark's reporting may import it; the product never does.

The cache has no counterpart in KiRaRay.  It follows Mueller et al.,
"Real-time Neural Radiance Caching for Path Tracing" (SIGGRAPH 2021): a
5 x 64 ReLU MLP without biases on an encoded path vertex, relative L2 loss,
Adam.  Pinned by finite differences (gradients), the closed form of one Adam
step and the encoding's definition; the GPU network is then checked against
this code with BF16 rounding emulated (tight) and without it (loose).
"""
import numpy as np

WIDTH = 64
HIDDEN = 5
OUTPUTS = 3
PARAMS = HIDDEN * WIDTH * WIDTH + WIDTH * OUTPUTS


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def _one_blob(x, bins=4):
    c = (np.arange(bins) + 0.5) / bins
    d = bins * (x[:, None] - c[None, :])
    return np.exp(-0.5 * d * d)


def encode(v):
    """(N, 16) vertex records -> (N, 64) encoding, feature k = 8s + 4g + j
    holding slot q = 4s + j of lane half g (see nrc.hip)."""
    v = np.asarray(v, np.float64)
    n = len(v)
    half0 = np.zeros((n, 32))
    v32 = np.asarray(v, np.float32)
    for c in range(3):
        for f in range(5):
            # the argument is formed in binary32 as the kernel does: (2^f * pi) * x
            a = (np.float32(2.0 ** f) * np.float32(np.pi)) * v32[:, c]
            half0[:, 10 * c + 2 * f] = np.sin(a.astype(np.float64))
            half0[:, 10 * c + 2 * f + 1] = np.cos(a.astype(np.float64))
    half0[:, 30:32] = 1.0
    half1 = np.zeros((n, 32))
    half1[:, 0:3] = v[:, 0:3]
    half1[:, 3:7] = _one_blob(v[:, 3])
    half1[:, 7:11] = _one_blob(v[:, 4])
    half1[:, 11:15] = _one_blob(v[:, 5])
    half1[:, 15:19] = _one_blob(v[:, 6])
    half1[:, 19:23] = _one_blob(v[:, 7])
    half1[:, 23:29] = v[:, 8:14]
    enc = np.zeros((n, 64))
    for s in range(8):
        for j in range(4):
            enc[:, 8 * s + j] = half0[:, 4 * s + j]
            enc[:, 8 * s + 4 + j] = half1[:, 4 * s + j]
    return enc


def unpack(w):
    """Flat weights (PARAMS) -> [W_0 .. W_4 (64x64), W_5 (64x3)], [in][out]."""
    w = np.asarray(w, np.float64)
    mats = [w[l * 4096:(l + 1) * 4096].reshape(64, 64) for l in range(HIDDEN)]
    mats.append(w[HIDDEN * 4096:].reshape(64, OUTPUTS))
    return mats


def pack(mats):
    return np.concatenate([m.reshape(-1) for m in mats])


def forward(w, x0, emulate_bf16=False):
    """-> (y (N,3), inputs [x_0 .. x_5]).  emulate_bf16 rounds weights and
    every layer input to bfloat16 as the matrix cores see them."""
    mats = unpack(w)
    r = bf16 if emulate_bf16 else (lambda a: a)
    xs = [r(x0)]
    for l in range(HIDDEN):
        z = xs[-1] @ r(mats[l])
        xs.append(r(np.maximum(z, 0.0)))
    y = xs[-1] @ r(mats[HIDDEN])
    return y, xs


def loss_and_grad(w, x0, targets, loss_eps=0.01, emulate_bf16=False):
    """Mean relative L2 loss over the 3N outputs (denominator y^2 + eps held
    constant) and its gradient with respect to the flat weights."""
    mats = unpack(w)
    r = bf16 if emulate_bf16 else (lambda a: a)
    y, xs = forward(w, x0, emulate_bf16)
    t = np.asarray(targets, np.float64)
    den = y * y + loss_eps
    n = y.size
    loss = np.sum((y - t) ** 2 / den) / n
    dz = r(2.0 * (y - t) / den / n)
    grads = [None] * (HIDDEN + 1)
    grads[HIDDEN] = xs[HIDDEN].T @ dz
    dx = dz @ r(mats[HIDDEN]).T
    for l in range(HIDDEN - 1, -1, -1):
        dz = r(dx * (xs[l + 1] > 0))
        grads[l] = xs[l].T @ dz
        if l > 0:
            dx = dz @ r(mats[l]).T
    return loss, pack(grads)


def adam_step(w, m, v, g, t, lr=1e-3, beta1=0.9, beta2=0.99, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) at step t (1-based)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def random_vertices(n, rng):
    """Plausible vertex records for tests: positions in [0,1]^3, angles in
    [0,1], roughness and albedos in [0,1]."""
    v = np.zeros((n, 16), np.float32)
    v[:, 0:14] = rng.random((n, 14))
    return v
