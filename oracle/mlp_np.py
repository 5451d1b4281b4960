"""CPU restatement of the learned warm start's network, features and Adam
(numpy) -- TEST INFRASTRUCTURE ONLY: the checker of the device kernels in
dnds_amd/csrc/warmstart.hip, never imported by the product.

The network and its inputs have no counterpart in the reference; this module
restates them from their definition:
  features: for 12 stencil cells (breadth-first over faces, cell2face order,
  missing cells repeat the cell) the differences u_j - u_i divided per
  variable by s_v = max(max_j |u_j - u_i|, 1e-3 max_v |u_i,v|) (1 if zero);
  the 9 base moments
  baseMoments[1..9] (DNDS_FV_VR.cpp:173-178); 0.5 log(volume).
  targets: rec coefficients [r + 9 v] / s_v.
  network: 58 -> 128 -> 128 -> 36, GELU (exact, erf), MSE loss.
  GEMM operands may be rounded to BF16 (round to nearest even) as the
  device's matrix cores see them, products accumulated in float64 here.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

N_IN, N_HID, N_OUT, N_STEN = 58, 128, 36, 12
SHAPES = [("W1", (N_IN, N_HID)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
          ("W3", (N_HID, N_OUT)), ("b3", (N_OUT,))]
N_PARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16(a: np.ndarray) -> np.ndarray:
    """float32 values rounded to the nearest BF16 (ties to even)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def _mm(a, b, round_bf16):
    if round_bf16:
        a, b = bf16(a), bf16(b)
    return a @ b


def forward(p: dict, X: np.ndarray, round_bf16=False):
    Z1 = _mm(X, p["W1"], round_bf16) + p["b1"]
    H1 = gelu(Z1)
    Z2 = _mm(H1, p["W2"], round_bf16) + p["b2"]
    H2 = gelu(Z2)
    Y = _mm(H2, p["W3"], round_bf16) + p["b3"]
    return Z1, H1, Z2, H2, Y


def loss_grad(params: np.ndarray, X: np.ndarray, T: np.ndarray, round_bf16=False):
    """Mean squared error over all outputs and its gradient (packed)."""
    p = unpack(params)
    X = np.asarray(X, dtype=np.float64)
    Z1, H1, Z2, H2, Y = forward(p, X, round_bf16)
    D = Y - T
    loss = float(np.mean(D * D))
    dY = 2.0 * D / D.size
    g = {"W3": _mm(H2.T, dY, round_bf16), "b3": dY.sum(0)}
    dZ2 = _mm(dY, p["W3"].T, round_bf16) * gelu_grad(Z2)
    g["W2"], g["b2"] = _mm(H1.T, dZ2, round_bf16), dZ2.sum(0)
    dZ1 = _mm(dZ2, p["W2"].T, round_bf16) * gelu_grad(Z1)
    g["W1"], g["b1"] = _mm(X.T, dZ1, round_bf16), dZ1.sum(0)
    return loss, pack(g)


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Textbook Adam (Kingma & Ba 2015, Algorithm 1), step t >= 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def stencil(mesh: dict) -> np.ndarray:
    """[n_cell][12] stencil cells."""
    cp, c2f, f2c = mesh["cell_ptr"], mesh["cell2face"], mesh["face2cell"]
    nc = len(cp) - 1
    out = np.empty((nc, N_STEN), dtype=np.int64)
    for i in range(nc):
        seen, ring, got = {i}, [i], []
        while len(got) < N_STEN and ring:
            nxt = []
            for a in ring:
                for f in c2f[cp[a]:cp[a + 1]]:
                    b = f2c[f, 1] if f2c[f, 0] == a else f2c[f, 0]
                    if b < 0 or b in seen:
                        continue
                    seen.add(b)
                    nxt.append(b)
                    if len(got) < N_STEN:
                        got.append(b)
            ring = nxt
        out[i] = got + [i] * (N_STEN - len(got))
    return out


def features(mesh: dict, setup: dict, u: np.ndarray):
    """(X [n][58], S [n][4])."""
    st = stencil(mesh)
    d = u[st] - u[:, None, :]                       # [n][12][4]
    s = np.maximum(np.abs(d).max(1), 1e-3 * np.abs(u).max(1)[:, None])
    s = np.where(s > 0, s, 1.0)
    X = np.empty((u.shape[0], N_IN))
    X[:, :48] = (d / s[:, None, :]).reshape(-1, 48)
    X[:, 48:57] = setup["base_moment"].reshape(-1, 10)[:, 1:]
    X[:, 57] = 0.5 * np.log(setup["volume"])
    return X, s
