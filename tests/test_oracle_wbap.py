"""The oracle's WBAP limiters (V2: DNDS_FV_VR.hpp:1749-2047, V3:
DNDS_FV_VR.hpp:2053-2329) against what the mathematics fixes.  CPU only.

A global linear field is exact in every cell's basis: mapped through
matrixSecondary, each neighbour's coefficients equal the cell's own, so every
biway and multiway sees equal arguments and both procedures return the field
unchanged.  Scaling one cell's coefficients by alpha then gives closed forms
that tell the procedures apart (which argument is limited against which, with
which n, and which entry the multiway is centred on):

  FWBAP_L2_Biway(a, b, n)  = a (n + t^3) / (n + t^4),   t = a / b
  FWBAP_L2_Multiway(c, o_1..o_k, n1) = c (n1 + sum t_k^3) / (n1 + sum t_k^4),
                                      t_k = c / o_k

V2: each neighbour gives Biway(own, neighbour, nStd); the multiway of those
results is centred on the first one, with n1 = 1.
V3: each neighbour gives Biway(own, neighbour, 1); the multiway is centred on
the cell's own coefficients, with n1 = nStd.
"""
from math import factorial

import numpy as np
import pytest

import oracle
from conftest import mesh_and_setup, rel_err

ALPHA = 3.0
NSTD = 10.0
G = np.array([[0.7, -0.4], [1.3, 0.9], [-0.8, 1.1], [2.1, -1.7]])  # d(var)/dx, d(var)/dy


def setup(kind):
    mesh, vfv = mesh_and_setup(kind=kind, nx=8, ny=6, periodic=False, jitter=0.2, seed=4)
    st = vfv.setting
    m, s = mesh.arrays(), vfv.arrays()
    orc = oracle.Oracle(m, s, riemann_solver=1, tang_weight=st.tang_weight,
                        tang_weight_direction=st.tang_weight_direction, wbap_n_std=NSTD)
    return m, s, orc


def linear_rec(s, n):
    """Coefficients of the global linear field with gradient G: the basis
    functions 1 and 2 are (x - xc)/sx and (y - yc)/sy less their means."""
    sc = s["cell_scale"].reshape(-1, 2)
    urec = np.zeros((n, 36))
    for v in range(4):
        urec[:, 9 * v] = G[v, 0] * sc[:, 0]
        urec[:, 9 * v + 1] = G[v, 1] * sc[:, 1]
    return urec


DIFF = [(1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3)]  # rows 0..8
A3 = np.array([[0.7, -0.4, 0.3, 0.5, -0.2, 0.6, -0.3, 0.4, 0.2],
               [1.3, 0.9, -0.5, 0.2, 0.4, -0.7, 0.5, 0.3, -0.6],
               [-0.8, 1.1, 0.6, -0.3, 0.2, 0.4, 0.8, -0.5, 0.3],
               [2.1, -1.7, 0.4, 0.6, -0.5, 0.3, -0.4, 0.7, 0.5]])  # [var][monomial DIFF[k] about (0.5, 0.5)]
A3[:, 5:] *= 1e3  # third-order rows O(0.1) on these cells, well above the biway's 1e-12 regulariser


def cubic_rec(s, n):
    """Coefficients of the global cubic sum_k A3[v, k] (x-0.5)^p (y-0.5)^q:
    row (a, b) holds d^(a+b)P/dx^a dy^b at the cell's centre times
    sx^a sy^b / (a! b!)."""
    sc = s["cell_scale"].reshape(-1, 2)
    xc = s["cell_bary"].reshape(-1, 2) - 0.5
    urec = np.zeros((n, 36))
    for v in range(4):
        for k, (a, b) in enumerate(DIFF):
            d = np.zeros(n)
            for kk, (p, q) in enumerate(DIFF):
                if p >= a and q >= b:
                    d += (A3[v, kk] * factorial(p) / factorial(p - a) * factorial(q) / factorial(q - b) *
                          xc[:, 0] ** (p - a) * xc[:, 1] ** (q - b))
            urec[:, 9 * v + k] = d * sc[:, 0] ** a * sc[:, 1] ** b / (factorial(a) * factorial(b))
    return urec


def neighbours(m, i):
    cp, c2f, f2c = m["cell_ptr"], m["cell2face"], m["face2cell"].reshape(-1, 2)
    out = []
    for f in c2f[cp[i]:cp[i + 1]]:
        j = f2c[f, 1] if f2c[f, 0] == i else f2c[f, 0]
        if j >= 0:
            out.append(int(j))
    return out


def state(n):
    rho, u, v, p = 1.0, 0.3, 0.2, 1.0
    return np.tile([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)], (n, 1))


@pytest.mark.parametrize("kind", [0, 1])
def test_matrix_secondary_maps_a_linear_field_exactly(kind):
    """The premise of the closed forms below."""
    m, s, orc = setup(kind)
    n = int(m["n_cell"])
    ms = s["matrix_secondary"].reshape(-1, 2, 9, 9)
    f2c = m["face2cell"].reshape(-1, 2)
    for urec in (linear_rec(s, n).reshape(n, 4, 9), cubic_rec(s, n).reshape(n, 4, 9)):
        for f in np.nonzero(f2c[:, 1] >= 0)[0]:
            L, R = f2c[f]
            assert rel_err(urec[R] @ ms[f, 0].T, urec[L]) < 1e-12
            assert rel_err(urec[L] @ ms[f, 1].T, urec[R]) < 1e-12


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("proc", ["wbap_v2", "wbap_v3"])
@pytest.mark.parametrize("field", [linear_rec, cubic_rec])
def test_wbap_keeps_a_global_polynomial(kind, proc, field):
    m, s, orc = setup(kind)
    n = int(m["n_cell"])
    urec = field(s, n)
    lim, _ = getattr(orc, proc)(state(n), urec, if_all=True)
    assert np.abs(lim - urec).max() <= 1e-11 * np.abs(urec).max()


def biway(t, n):
    return (n + t ** 3) / (n + t ** 4)


def expected_v2(nb_count, first_is_scaled, own_scaled):
    """The factor of the linear coefficients under V2."""
    if own_scaled:
        return ALPHA * biway(ALPHA, NSTD)  # every neighbour gives the same, the multiway keeps it
    d = biway(1 / ALPHA, NSTD)  # the face with the scaled cell; all others give 1
    if first_is_scaled:
        k = nb_count - 1
        return d * (1 + k * d ** 3) / (1 + k * d ** 4)
    k = nb_count - 2
    return (1 + k + d ** -3) / (1 + k + d ** -4)


def expected_v3(nb_count, own_scaled):
    if own_scaled:
        g = ALPHA * biway(ALPHA, 1.0)
        t = ALPHA / g
        return ALPHA * (NSTD + nb_count * t ** 3) / (NSTD + nb_count * t ** 4)
    e = biway(1 / ALPHA, 1.0)
    return (NSTD + nb_count - 1 + e ** -3) / (NSTD + nb_count - 1 + e ** -4)


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("where", ["interior", "boundary"])
def test_wbap_closed_forms_around_one_scaled_cell(kind, where):
    m, s, orc = setup(kind)
    n = int(m["n_cell"])
    base = linear_rec(s, n)
    counts = np.array([len(neighbours(m, i)) for i in range(n)])
    full = counts == (m["cell_ptr"][1:] - m["cell_ptr"][:-1])
    bary = s["cell_bary"].reshape(-1, 2)
    pick = full if where == "interior" else ~full
    i = int(np.nonzero(pick)[0][np.argmin(np.linalg.norm(bary[pick] - 0.5, axis=1))])
    urec = base.copy()
    urec[i] *= ALPHA
    u = state(n)
    lin = [9 * v + r for v in range(4) for r in (0, 1)]
    lim2, _ = orc.wbap_v2(u, urec, if_all=True)
    lim3, _ = orc.wbap_v3(u, urec, if_all=True)
    nbs = neighbours(m, i)
    assert rel_err(lim2[i, lin], expected_v2(counts[i], True, True) * base[i, lin]) < 1e-10
    assert rel_err(lim3[i, lin], expected_v3(counts[i], True) * base[i, lin]) < 1e-10
    for j in nbs:
        first = neighbours(m, j)[0] == i
        assert rel_err(lim2[j, lin], expected_v2(counts[j], first, False) * base[j, lin]) < 1e-10
        assert rel_err(lim3[j, lin], expected_v3(counts[j], False) * base[j, lin]) < 1e-10
    others = np.setdiff1d(np.arange(n), [i] + nbs)
    for lim in (lim2, lim3):
        assert np.abs(lim[others] - base[others]).max() <= 1e-11 * np.abs(base).max()
        assert np.all(lim[:, [9 * v + r for v in range(4) for r in range(2, 9)]] == 0.0)


def test_v2_and_v3_share_the_indicator_and_leave_smooth_cells_alone():
    mesh, vfv = mesh_and_setup(kind=0, nx=24, ny=24, periodic=True, jitter=0.0, seed=2)
    st = vfv.setting
    orc = oracle.Oracle(mesh.arrays(), vfv.arrays(), riemann_solver=1, tang_weight=st.tang_weight,
                        tang_weight_direction=st.tang_weight_direction,
                        wbap_smooth_indicator_scale=st.wbap_smooth_indicator_scale)
    bary = vfv.arrays(["cell_bary"])["cell_bary"].reshape(-1, 2)
    u = state(mesh.n_cell)
    u[:, 0] *= 1 + 0.01 * np.sin(2 * np.pi * bary[:, 0])
    u[bary[:, 0] > 0.5, 3] *= 0.3
    urec = np.zeros((mesh.n_cell, 36))
    for _ in range(30):
        urec = orc.jacobi_step(u, urec)
    lim2, ind2 = orc.wbap_v2(u, urec)
    lim3, ind3 = orc.wbap_v3(u, urec)
    assert np.array_equal(ind2, ind3)
    smooth = ind2 < st.wbap_smooth_indicator_scale
    assert smooth.sum() > 0 and (~smooth).sum() > 0
    assert np.array_equal(lim2[smooth], urec[smooth]) and np.array_equal(lim3[smooth], urec[smooth])
    assert not np.array_equal(lim2[~smooth], lim3[~smooth])
    assert np.abs(lim2).max() <= np.abs(urec).max() * (1 + 1e-12)


@pytest.mark.parametrize("kind", [0, 1])
def test_wbap_order3_closed_forms_and_the_partly_limited_neighbours(kind):
    """A global cubic with one cell's third-order rows scaled: the third-order
    rows follow the closed forms; the cell's second-order rows, equal in every
    neighbour's mapped polynomial as given, must change, because the second
    order reads the neighbours' third-order rows as already limited (V2:
    uRecNewBuf, DNDS_FV_VR.hpp:1935-1941) or the cell's own limited face values
    (V3: uFaces, DNDS_FV_VR.hpp:2248-2252)."""
    m, s, orc = setup(kind)
    n = int(m["n_cell"])
    base = cubic_rec(s, n)
    counts = np.array([len(neighbours(m, i)) for i in range(n)])
    full = counts == (m["cell_ptr"][1:] - m["cell_ptr"][:-1])
    bary = s["cell_bary"].reshape(-1, 2)
    i = int(np.nonzero(full)[0][np.argmin(np.linalg.norm(bary[full] - 0.5, axis=1))])
    o3 = [9 * v + r for v in range(4) for r in range(5, 9)]
    o2 = [9 * v + r for v in range(4) for r in range(2, 5)]
    urec = base.copy()
    urec[i, o3] *= ALPHA
    u = state(n)
    lim2, _ = orc.wbap_v2(u, urec, if_all=True)
    lim3, _ = orc.wbap_v3(u, urec, if_all=True)
    assert rel_err(lim2[i, o3], expected_v2(counts[i], True, True) * base[i, o3]) < 1e-10
    assert rel_err(lim3[i, o3], expected_v3(counts[i], True) * base[i, o3]) < 1e-10
    for j in neighbours(m, i):
        first = neighbours(m, j)[0] == i
        assert rel_err(lim2[j, o3], expected_v2(counts[j], first, False) * base[j, o3]) < 1e-10
        assert rel_err(lim3[j, o3], expected_v3(counts[j], False) * base[j, o3]) < 1e-10
    for lim in (lim2, lim3):
        assert rel_err(lim[i, o2], base[i, o2]) > 1e-7  # the unchanged rows would give ~1e-12
