"""The warm start's CPU restatement (oracle/mlp_np.py) against what fixes it
independently: central finite differences of the loss, the textbook Adam
recursion unrolled by hand, BF16 rounding on known values, and the
invariances of the features.  CPU only."""
import numpy as np
import pytest

from oracle import mlp_np as M


def _params(seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(M.N_PARAM) * 0.2


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(0)
    X = rng.standard_normal((7, M.N_IN))
    T = rng.standard_normal((7, M.N_OUT))
    loss, g = M.loss_grad(p, X, T)
    idx = rng.choice(M.N_PARAM, 60, replace=False)
    # every block of parameters is probed
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.SHAPES])
    idx = np.concatenate([idx, offs[:-1], offs[1:] - 1])
    eps = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = eps
        fd = (M.loss_grad(p + e, X, T)[0] - M.loss_grad(p - e, X, T)[0]) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-9, k


def test_gelu_matches_its_definition_and_derivative():
    z = np.linspace(-5, 5, 101)
    from math import erf, sqrt
    ref = np.array([0.5 * x * (1 + erf(x / sqrt(2))) for x in z])
    assert np.abs(M.gelu(z) - ref).max() < 1e-15
    h = 1e-6
    assert np.abs(M.gelu_grad(z) - (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)).max() < 1e-8


def test_adam_unrolled():
    p, g1, g2 = np.array([1.0, -2.0]), np.array([0.5, 0.1]), np.array([-0.3, 0.2])
    m = v = np.zeros(2)
    p1, m, v = M.adam_step(p, g1, m, v, 1, 0.01)
    # step 1: m_hat = g, v_hat = g^2, update = lr g / (|g| + eps)
    assert np.allclose(p1, p - 0.01 * g1 / (np.abs(g1) + 1e-8), rtol=0, atol=1e-15)
    p2, m, v = M.adam_step(p1, g2, m, v, 2, 0.01)
    mh = (0.9 * 0.1 * g1 + 0.1 * g2) / (1 - 0.81)
    vh = (0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2) / (1 - 0.999 ** 2)
    assert np.allclose(p2, p1 - 0.01 * mh / (np.sqrt(vh) + 1e-8), rtol=0, atol=1e-14)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -3.140625, 1e-3], dtype=np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0            # tie to even (down)
    assert r[2] == 1.0 + 2 ** -7                  # representable
    assert r[3] == 1.0 + 2 ** -6                  # tie to even (up)
    assert r[4] == -3.140625
    assert abs(r[5] - 1e-3) <= 2 ** -9 * 1e-3


def _small_mesh():
    from conftest import mesh_and_setup
    mesh, vfv = mesh_and_setup(kind=2, nx=6, ny=5, periodic=False, jitter=0.2, seed=3)
    return mesh.arrays(), vfv.arrays()


def test_stencil_rings_are_face_neighbours_first():
    m, _ = _small_mesh()
    st = M.stencil(m)
    f2c, cp, c2f = m["face2cell"], m["cell_ptr"], m["cell2face"]
    for i in range(m["n_cell"]):
        nb = [f2c[f, 1] if f2c[f, 0] == i else f2c[f, 0] for f in c2f[cp[i]:cp[i + 1]]]
        nb = [j for j in nb if j >= 0]
        assert list(st[i, :len(nb)]) == nb
        real = [j for j in st[i] if j != i]
        assert len(set(real)) == len(real)


def test_features_are_scale_invariant_and_floored():
    m, s = _small_mesh()
    rng = np.random.default_rng(2)
    u = rng.uniform(0.5, 2.0, (m["n_cell"], 4))
    X, S = M.features(m, s, u)
    X2, S2 = M.features(m, s, 3.0 * u)
    assert np.allclose(X, X2, rtol=0, atol=1e-12)
    assert np.allclose(S2, 3.0 * S, rtol=1e-13)
    assert np.abs(X[:, :48]).max() <= 1.0 + 1e-15
    # a uniform state: differences vanish, the scale is the floor
    U = np.tile([1.0, 0.5, -0.2, 2.5], (m["n_cell"], 1))
    X3, S3 = M.features(m, s, U)
    assert np.all(X3[:, :48] == 0.0) and np.allclose(S3, 2.5e-3)
