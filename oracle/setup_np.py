"""numpy restatement of the geometry and VR setup (DNDS_FV_VR.hpp:29-102,
DNDS_FV_VR.cpp:82-1154) for small meshes, global-coordinate basis or the
local inertia coordinates (setting "use_local_coord", DNDS_FV_VR.hpp:358-1018),
p = 3 or p = 1 (setting "rec_order"; p = 1 matrices in the top-left 2 x 2
corner of the 9 x 9 blocks, the rest zero).

TEST INFRASTRUCTURE ONLY: an independent second implementation that the
tests hold the product's host setup against.  It is written from the
reference's formulas with dense numpy algebra (matrix products, pinv) rather
than the product's per-entry loops, so agreement is to round-off of the
conditioning of A_ii, not bit for bit.
"""
from __future__ import annotations

import math

import numpy as np

NO_CELL = -(2 ** 63)
DIFF = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3)]  # diffOperatorOrderList2D
NCOMB = [1, 1, 1, 1, 2, 1, 1, 3, 3, 1]
DIRW3 = [1, .5295, .2117, .2117]

G3 = [-0.7745966692414833, 0.0, 0.7745966692414833]
W3 = [5. / 9., 8. / 9., 5. / 9.]
G5 = [-0.906179845938664, -0.538469310105683, 0, 0.538469310105683, 0.906179845938664]
W5 = [0.236926885056188, 0.478628670499366, 0.568888888888889, 0.478628670499366, 0.236926885056190]
TRI4 = [(1 / 3, 1 / 3, -27 / 96), (0.6, 0.2, 25 / 96), (0.2, 0.6, 25 / 96), (0.2, 0.2, 25 / 96)]
_a1, _b1, _w1 = 0.059715871789770, 0.470142064105115, 0.132394152788506 * 0.5
_a2, _b2, _w2 = 0.797426985353087, 0.101286507323456, 0.125939180544827 * 0.5
TRI7 = [(1 / 3, 1 / 3, 9 / 80), (_a1, _b1, _w1), (_b1, _a1, _w1), (_b1, _b1, _w1), (_a2, _b2, _w2), (_b2, _a2, _w2),
        (_b2, _b2, _w2)]
QUAD9 = [(G3[i], G3[j], W3[i] * W3[j]) for j in range(3) for i in range(3)]

BC_WALL, BC_FAR, BC_WALL_EULER, BC_NOSLIP = 1, 2, 3, 4
FAR_KINDS = {2, -2, -3, -4, -5}


def shape(nv, xi, eta):
    if nv == 3:
        return np.array([1 - xi - eta, xi, eta]), np.array([-1.0, 1.0, 0.0]), np.array([-1.0, 0.0, 1.0])
    N = np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)]) / 4
    Nx = np.array([-(1 - eta), 1 - eta, 1 + eta, -(1 + eta)]) / 4
    Ny = np.array([-(1 - xi), -(1 + xi), 1 + xi, 1 - xi]) / 4
    return N, Nx, Ny


def point_det(X, xi, eta):
    N, Nx, Ny = shape(len(X), xi, eta)
    J = np.array([Nx @ X, Ny @ X])  # rows d/dxi, d/deta; columns x, y
    return N @ X, np.linalg.det(J)


def face_param(nv, lf, inv, xi):
    """Parametric point of face coordinate xi on local face lf (FaceSpace2VolSpace)."""
    if nv == 3:
        t = 0.5 * (xi + 1) if not inv else 0.5 * (1 - xi)
        return [(t, 0.0), (1 - t, t), (0.0, 1 - t)][lf]
    s = xi if not inv else -xi
    return [(s, -1.0), (1.0, s), (-s, 1.0), (-1.0, -s)][lf]


def basis(p, c, scale, moment):
    """10 x 10 derivatives of the scaled monomials (FDiffBaseValue, global coords)."""
    X, Y = (p - c) / scale
    D = np.zeros((10, 10))
    for i, (dx, dy) in enumerate(DIFF):
        for j, (px, py) in enumerate(DIFF):
            if px >= dx and py >= dy:
                coef = math.factorial(px) // math.factorial(px - dx) * math.factorial(py) // math.factorial(py - dy)
                D[i, j] = coef * X ** (px - dx) * Y ** (py - dy) / (scale[0] ** dx * scale[1] ** dy)
    D[0] -= moment
    return D


def _poly_mul(a, b):
    """product of two polynomials in (dx, dy) held as 4 x 4 coefficient grids"""
    out = np.zeros((4, 4))
    for i in range(4):
        for j in range(4):
            if a[i, j]:
                out[i:, j:] += a[i, j] * b[:4 - i, :4 - j]
    return out


def basis_local(p, c, A, moment):
    """10 x 10 derivatives of the monomials of xi = A (p - c) (FDiffBaseValue,
    local coordinates), by a route of its own: each xi-monomial is expanded
    into monomials of (dx, dy) = p - c, whose derivatives in x and y are
    those of the unscaled global basis."""
    e0 = np.zeros((4, 4))
    e0[1, 0], e0[0, 1] = A[0, 0], A[0, 1]
    e1 = np.zeros((4, 4))
    e1[1, 0], e1[0, 1] = A[1, 0], A[1, 1]
    one = np.zeros((4, 4))
    one[0, 0] = 1.0
    pw0, pw1 = [one], [one]
    for _ in range(3):
        pw0.append(_poly_mul(pw0[-1], e0))
        pw1.append(_poly_mul(pw1[-1], e1))
    Dg = basis(p, c, np.ones(2), np.zeros(10))  # derivatives of dx^i dy^j at p
    D = np.zeros((10, 10))
    for j, (a, b) in enumerate(DIFF):
        coef = _poly_mul(pw0[a], pw1[b])
        for k, (i2, j2) in enumerate(DIFF):
            if coef[i2, j2]:
                D[:, j] += coef[i2, j2] * Dg[:, k]
    D[0] -= moment
    return D


def local_frame(X, bary, inertia, anisotropic):
    """invPJacobi of the local coordinates: the principal axes of the cell's
    second-moment tensor (largest first, oriented so that the first axis's
    larger component is positive; the x axis when the tensor is isotropic to
    1e-10 of its trace; the second axis the first turned by +90 degrees),
    scaled by the cell's half extents along them (or both by their
    diagonal)."""
    w, V = np.linalg.eigh(inertia)
    if w[1] - w[0] <= 1e-10 * np.trace(inertia):
        a0 = np.array([1.0, 0.0])
    else:
        a0 = V[:, np.argmax(w)]
        if a0[np.argmax(np.abs(a0))] < 0:
            a0 = -a0
    P = np.stack([a0, np.array([-a0[1], a0[0]])], axis=1)
    nc = np.linalg.solve(P, (X - bary).T)
    half = (nc.max(1) - nc.min(1)) / 2
    s = half if anisotropic else np.full(2, np.linalg.norm(half))
    return np.linalg.inv(P * s[None, :])


def functional_rows(W, n, tang_weight, tang_norm):
    """Rows P (10 x 10) and diagonal weights q with Conj(DI, DJ) = (P DI)^T diag(q) (P DJ)
    (FFaceFunctional with USE_NORM_FUNCTIONAL, tangWeightModMin = 1)."""
    if abs(W[1]) >= 1e-50 or abs(W[2]) >= 1e-50:
        big1 = abs(W[1]) >= abs(W[2])
        w2r = W[3] / W[1] ** 2 if big1 else W[5] / W[2] ** 2
        w3r = W[6] / W[1] ** 3 if big1 else W[9] / W[2] ** 3
        length = math.hypot(W[1], W[2])
    else:
        w2r = w3r = length = 0.0
    nn = (n / np.abs(n).max()) if tang_norm else None
    fn = (nn / np.linalg.norm(nn) * length) if tang_norm else np.array([W[1], W[2]])
    n0, n1 = fn
    t0, t1 = -n1 * tang_weight, n0 * tang_weight
    P = np.zeros((10, 10))
    P[0, 0] = 1
    P[1, 1:3] = [n0, n1]
    P[2, 1:3] = [t0, t1]
    P[3, 3:6] = np.array([n0 * n0, 2 * n0 * n1, n1 * n1]) * w2r
    P[4, 3:6] = np.array([n0 * t0, n0 * t1 + n1 * t0, n1 * t1]) * w2r
    P[5, 3:6] = np.array([t0 * t0, 2 * t0 * t1, t1 * t1]) * w2r
    P[6, 6:10] = np.array([n0 ** 3, 3 * n0 * n0 * n1, 3 * n0 * n1 * n1, n1 ** 3]) * w3r
    P[7, 6:10] = np.array([n0 * n0 * t0, n0 * n0 * t1 + 2 * n0 * t0 * n1, 2 * n0 * n1 * t1 + t0 * n1 * n1,
                           n1 * n1 * t1]) * w3r
    P[8, 6:10] = np.array([n0 * t0 * t0, 2 * t0 * n0 * t1 + t0 * t0 * n1, 2 * t0 * n1 * t1 + n0 * t1 * t1,
                           n1 * t1 * t1]) * w3r
    P[9, 6:10] = np.array([t0 ** 3, 3 * t0 * t0 * t1, 3 * t0 * t1 * t1, t1 ** 3]) * w3r
    q = np.array([W[0] ** 2, 1, 1, 1, 2, 1, 1, 3, 3, 1], dtype=float)
    return P, q


def pinv(A):
    return np.linalg.pinv(A, rcond=A.shape[0] * np.finfo(float).eps)


def build(mesh: dict, setting: dict) -> dict:
    """Setup arrays (keys as the product's VR view) for a small mesh.

    setting keys: wall_weight, far_weight, tang_weight, tang_weight_direction
    (0 Norm / 1 Bary), anisotropic_lengths, weight_scheme_geom (0 None, 1 D,
    2 S, 3 SDHQM), weight_scheme_dir (0 None / 1 optHQM), weight_len_scale."""
    xy, cp, c2n, c2f = mesh["node_xy"], mesh["cell_ptr"], mesh["cell2node"], mesh["cell2face"]
    f2c, f2n, fref, bc, shift = (mesh["face2cell"], mesh["face2node"], mesh["face2cell_ref"], mesh["face_bc"],
                                 mesh["face_shift"])
    nc, nf = mesh["n_cell"], mesh["n_face"]
    coords = [xy[c2n[cp[i]:cp[i + 1]]] for i in range(nc)]
    vol, bary, scale, mom = np.zeros(nc), np.zeros((nc, 2)), np.zeros((nc, 2)), np.zeros((nc, 10))
    local = bool(setting.get("use_local_coord", 0))
    frame = np.zeros((nc, 2, 2))

    def cell_basis(i, p, c, moment):
        return basis_local(p, c, frame[i], moment) if local else basis(p, c, scale[i], moment)

    for i, X in enumerate(coords):
        rule_v, rule_b = (TRI7, TRI4) if len(X) == 3 else (QUAD9, QUAD9)
        vol[i] = sum(point_det(X, a, b)[1] * w for a, b, w in rule_v)
        bary[i] = sum(point_det(X, a, b)[0] * point_det(X, a, b)[1] * w for a, b, w in rule_b) / vol[i]
        s = (X.max(0) - X.min(0)) / 2
        scale[i] = s if setting["anisotropic_lengths"] else np.full(2, np.linalg.norm(s))
        if local:
            inertia = sum(np.outer(point_det(X, a, b)[0] - bary[i], point_det(X, a, b)[0] - bary[i])
                          * point_det(X, a, b)[1] * w for a, b, w in rule_b) / vol[i]
            frame[i] = local_frame(X, bary[i], inertia, setting["anisotropic_lengths"])
        mom[i] = sum(cell_basis(i, point_det(X, a, b)[0], bary[i], np.zeros(10))[0] * point_det(X, a, b)[1] * w
                     for a, b, w in rule_b) / vol[i]
        mom[i, 0] = 1.0

    def side_point(f, side, xi):
        """(point, base centre) of face point xi seen from side `side`"""
        i = f2c[f, side]
        sh = shift[f] if side == 1 else np.zeros(2)
        X = coords[i] + sh
        nv = len(X)
        lf = fref[f, side]
        start = X[lf]
        fx0 = xy[f2n[f, 0]]
        inv = not np.allclose(start, fx0, atol=1e-12 * (1 + np.abs(fx0).max()))
        a, b = face_param(nv, lf, inv, xi)
        p, _ = point_det(X, a, b)
        return p, bary[i] + sh

    def side_basis(f, side, xi):
        p, c = side_point(f, side, xi)
        i = f2c[f, side]
        return cell_basis(i, p, c, mom[i])

    fn = np.zeros((nf, 2))
    area = np.zeros(nf)
    fc = np.zeros((nf, 2))
    W = np.zeros((nf, 10))
    for f in range(nf):
        a, b = xy[f2n[f, 0]], xy[f2n[f, 1]]
        d = 0.5 * (b - a)
        fn[f] = [d[1], -d[0]]
        area[f] = sum(np.linalg.norm(fn[f]) * w for w in W5)
        fc[f] = 0.5 * (a + b)
        L, R = f2c[f]
        if R != NO_CELL:
            delta = bary[R] + shift[f] - bary[L]
            Wf = np.ones(10)
        else:
            delta = fc[f] - bary[L]
            Wf = np.zeros(10)
            k = bc[f]
            Wf[0] = setting["far_weight"] if k in FAR_KINDS else setting["wall_weight"]
            if k == BC_NOSLIP:
                u = fn[f] / np.linalg.norm(fn[f])
                delta = u * (u @ delta)
        D = np.linalg.norm(delta)
        GW = {0: 1.0, 1: (1 / D) ** 0.5, 2: (1 / area[f]) ** 0.5, 3: (area[f] / D) ** 0.25}[setting["weight_scheme_geom"]]
        if setting["tang_weight_direction"] == 0:
            delta = fn[f] / np.linalg.norm(fn[f]) * np.linalg.norm(delta)
        delta = delta * setting["weight_len_scale"]
        for k, (dx, dy) in enumerate(DIFF):
            if setting["weight_scheme_dir"] == 0:
                Wf[k] *= GW * delta[0] ** dx * delta[1] ** dy * NCOMB[k] / math.factorial(dx + dy)
            else:
                Wf[k] *= GW * delta[0] ** dx * delta[1] ** dy * DIRW3[dx + dy]
        W[f] = Wf

    d0 = np.zeros((nf, 3, 2, 9))
    gxy = np.zeros((nf, 3, 2, 2))
    dg0 = np.zeros((nf, 2, 10, 9))
    ms = np.zeros((nf, 2, 9, 9))
    for f in range(nf):
        for side in (0, 1) if f2c[f, 1] != NO_CELL else (0,):
            for g in range(3):
                Db = side_basis(f, side, G3[g])
                d0[f, g, side] = Db[0, 1:]
                p, c = side_point(f, side, G3[g])
                gxy[f, g, side] = frame[f2c[f, side]] @ (p - c) if local else (p - c) / scale[f2c[f, side]]
                if g == 0:
                    dg0[f, side] = Db[:, 1:]
        if f2c[f, 1] != NO_CELL:
            AL, AR = side_basis(f, 0, 0.0)[1:, 1:], side_basis(f, 1, 0.0)[1:, 1:]
            ms[f, 0] = pinv(AL) @ AR
            ms[f, 1] = pinv(AR) @ AL

    tn = setting["tang_weight_direction"] == 0
    p1 = setting.get("rec_order", 3) == 1
    nr = 2 if p1 else 9
    ainv = np.zeros((nc, 9, 9))
    ainv_b = np.zeros((cp[-1], 9, 9))
    ainv_bvec = np.zeros((cp[-1], 9))
    for i in range(nc):
        A = np.zeros((9, 9))
        faces = c2f[cp[i]:cp[i + 1]]
        blocks = []
        for f in faces:
            side = 0 if f2c[f, 0] == i else 1
            nl = np.linalg.norm(fn[f])
            if p1:
                # FFaceFunctional with 3 weights: DiffI^T diag(W)^2 DiffJ over
                # the rows 0, d/dx, d/dy (DNDS_FV_VR.hpp:1283-1286)
                P = np.zeros((10, 10))
                P[0, 0] = P[1, 1] = P[2, 2] = 1.0
                q = np.zeros(10)
                q[:3] = W[f, :3] ** 2
            else:
                P, q = functional_rows(W[f], fn[f], setting["tang_weight"], tn)
            other = f2c[f, 1 - side]
            B = np.zeros((9, 9))
            row = np.zeros(9)
            for g in range(3):
                DI = side_basis(f, side, G3[g])
                PI = P @ DI
                A += (PI.T @ (q[:, None] * PI))[1:, 1:] * nl * W3[g]
                if other != NO_CELL:
                    PJ = P @ side_basis(f, 1 - side, G3[g])
                    B += (PI.T @ (q[:, None] * PJ))[1:, 1:] * nl * W3[g]
                elif bc[f] in FAR_KINDS:
                    B += (PI.T @ (q[:, None] * PI))[1:, 1:] * nl * W3[g]
                row += W[f, 0] ** 2 * DI[0, 1:] * nl * W3[g]
            blocks.append((B, row))
        Ai = np.zeros((9, 9))
        Ai[:nr, :nr] = pinv(A[:nr, :nr])
        ainv[i] = Ai
        for k, (B, row) in enumerate(blocks):
            B[nr:, :] = 0.0
            B[:, nr:] = 0.0
            row[nr:] = 0.0
            ainv_b[cp[i] + k] = Ai @ B
            ainv_bvec[cp[i] + k] = Ai @ row
    return dict(volume=vol, face_area=area, cell_bary=bary.ravel(), base_moment=mom.ravel(), face_center=fc.ravel(),
                face_norm_center=fn.ravel(), face_weight=W.ravel(), face_d0=d0.ravel(), face_dibj_g0=dg0.ravel(),
                matrix_secondary=ms.ravel(), ainv=ainv.ravel(), ainv_b=ainv_b.ravel(), ainv_bvec=ainv_bvec.ravel(),
                face_gauss_xy=gxy.ravel(), cell_scale=np.asarray(scale, dtype=float).ravel(),
                cell_frame=frame.ravel() if local else None)
