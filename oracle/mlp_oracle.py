"""CPU oracle of the learned Schur-preconditioner scaling (test infrastructure only).

Only tests/, the smoke check and the benchmark's CPU leg use this module, as the checker of the
HIP kernels; the product path never imports it.

What it restates (numpy, float64 unless stated):
  * features: log10 of the element viscosity over the 3^dim neighbourhood of each pressure cell,
    relative to the cell itself, with cells outside the box read as WALL;
  * the 27 -> 64 -> 64 -> 1 ReLU MLP with BF16 weights (round to nearest even) and FP32 biases;
  * the loss L = 0.5/n sum (out - target)^2, its gradients, and one Adam step on FP32 master weights;
  * the preconditioner diagonal pc = approxS * 10^clip(out, -4, 4).

The scaling replaces approxS, the "uw" Schur preconditioner diagonal g^T diag(K_e)^-1 g of
UzawaPreconditionerTerm.c:195-212 (applied as Q22 in preconditioner.c:138-166), so this file has
no counterpart in the reference beyond that diagonal.

Pinned by tests/test_mlp_oracle.py: the analytic gradients against central finite differences, the
BF16 rounding against torch's bfloat16 conversion and hand-worked ties, the Adam step against its
closed form on a scalar, and the features against a brute-force loop.
"""
import numpy as np

HIDDEN = 64
WALL = 3.0


def bf16_round(x):
    """FP32 -> BF16 -> FP32 with round-to-nearest-even (NaN stays NaN)."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    quiet = np.where((u & 0xFFFF) != 0, np.uint64(0x40), np.uint64(0))
    r = np.where(nan, ((u >> np.uint64(16)) | quiet) << np.uint64(16), r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


def param_counts(nin, hidden=HIDDEN):
    return hidden * nin + hidden * hidden + hidden, 2 * hidden + 1


def unpack(w, b, nin, hidden=HIDDEN):
    """Flat parameter vectors -> (W1 [H][nin], W2 [H][H], W3 [H], b1, b2, b3)."""
    w = np.asarray(w, np.float64)
    b = np.asarray(b, np.float64)
    o1 = hidden * nin
    o2 = o1 + hidden * hidden
    return (w[:o1].reshape(hidden, nin), w[o1:o2].reshape(hidden, hidden), w[o2:o2 + hidden],
            b[:hidden], b[hidden:2 * hidden], b[2 * hidden])


def forward(x, w, b, quantize=True):
    """out, (h1, h2) for samples x [n][nin]; weights rounded to BF16 when quantize."""
    x = np.asarray(x, np.float64)
    nin = x.shape[1]
    hidden = (len(b) - 1) // 2
    if quantize:
        w = bf16_round(w)
    W1, W2, W3, b1, b2, b3 = unpack(w, b, nin, hidden)
    h1 = np.maximum(x @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    return h2 @ W3 + b3, (h1, h2)


def loss_and_grads(x, t, w, b, quantize=True):
    """L = 0.5/n sum (out - t)^2 and dL/dw, dL/db (the weights seen by the forward are the BF16 ones)."""
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    n, nin = x.shape
    hidden = (len(b) - 1) // 2
    wq = bf16_round(w) if quantize else np.asarray(w, np.float64)
    W1, W2, W3, b1, b2, b3 = unpack(wq, b, nin, hidden)
    out, (h1, h2) = forward(x, wq, b, quantize=False)
    d = (out - t) / n
    loss = 0.5 * np.sum((out - t) ** 2) / n
    gW3 = h2.T @ d
    gb3 = d.sum()
    dh2 = (h2 > 0) * (d[:, None] * W3[None, :])
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (h1 > 0) * (dh2 @ W2)
    gW1 = dh1.T @ x
    gb1 = dh1.sum(0)
    gw = np.concatenate([gW1.ravel(), gW2.ravel(), gW3])
    gb = np.concatenate([gb1, gb2, [gb3]])
    return loss, gw, gb


def adam_step(w, b, gw, gb, state, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step on FP32 master weights; state = dict(mw, vw, mb, vb, step).  Returns (w, b, state)."""
    step = state["step"] + 1
    c1 = 1.0 - beta1 ** step
    c2 = 1.0 - beta2 ** step
    out = {"step": step}
    new = []
    for p, g, mk, vk in ((w, gw, "mw", "vw"), (b, gb, "mb", "vb")):
        m = beta1 * state[mk] + (1 - beta1) * g
        v = beta2 * state[vk] + (1 - beta2) * g * g
        out[mk], out[vk] = m, v
        new.append(p - lr * (m / c1) / (np.sqrt(v / c2) + eps))
    return new[0], new[1], out


def adam_state(nw, nb):
    return {"mw": np.zeros(nw), "vw": np.zeros(nw), "mb": np.zeros(nb), "vb": np.zeros(nb), "step": 0}


def features(log_eta, shape):
    """log_eta per element (lexicographic, x fastest), shape = (nx, ny[, nz]) -> [nel][3^dim]."""
    dim = len(shape)
    nx, ny = shape[0], shape[1]
    nz = shape[2] if dim == 3 else 1
    le = np.asarray(log_eta, np.float64).reshape(nz, ny, nx)
    nel = nx * ny * nz
    nin = 27 if dim == 3 else 9
    e = np.arange(nel)
    i, j, k = e % nx, (e // nx) % ny, e // (nx * ny)
    F = np.empty((nel, nin))
    for t in range(nin):
        di, dj, dk = t % 3 - 1, (t // 3) % 3 - 1, (t // 9 - 1) if dim == 3 else 0
        ii, jj, kk = i + di, j + dj, k + dk
        inside = (ii >= 0) & (ii < nx) & (jj >= 0) & (jj < ny) & (kk >= 0) & (kk < nz)
        v = le[np.clip(kk, 0, nz - 1), np.clip(jj, 0, ny - 1), np.clip(ii, 0, nx - 1)] - le[k, j, i]
        F[:, t] = np.where(inside, v, WALL)
    return F


def scaling(approxS, out):
    return np.asarray(approxS) * 10.0 ** np.clip(np.asarray(out, np.float64), -4.0, 4.0)
