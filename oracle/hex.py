"""FP64 restatement of the reference's solid-shell hexahedron, vectorised over elements.

TEST INFRASTRUCTURE ONLY: the checker for the library's hex element and hex solve.
Imported by tests/, never by the product.

Follows engine/sim/solid_shell.cpp (solidshell::computeForms, computeBlocks, the
condensation of elementStiffness, elementMass, rotationFrom/polarRotation and
internalForce), restated in numpy over a batch of elements [E, 8, 3]:
  - natural coordinates of the nodes, solid_shell.hpp:124-125; node ordering 0-3 on
    zeta = -1 and 4-7 on zeta = +1 (solid_shell.hpp:115-122);
  - 2x2x2 Gauss at +-1/sqrt(3), point bits (g & 1 -> xi, g & 2 -> eta, g & 4 -> zeta),
    weight det J (solid_shell.cpp:237-260);
  - covariant strain rows [E11, E22, E33, 2E12, 2E23, 2E31] (solid_shell.cpp:140-158);
  - ANS transverse shear sampled at (xi, eta) = (-1, 0), (+1, 0) for 2E23 and
    (0, -1), (0, +1) for 2E31, linear in xi resp. eta; ANS thickness strain from the
    four in-plane corners, bilinear; both at the point's zeta (solid_shell.cpp:181-235);
  - pushed to Cartesian with A = J^-1 (solid_shell.cpp:116-138);
  - seven EAS modes on the covariant rows (xi on E11, eta on E22, zeta, xi zeta,
    eta zeta on E33, xi and eta on 2E12), transformed with the centre's A, scaled by
    det J0 / det J and normalised per mode in the weighted L2 norm over the points
    (solid_shell.cpp:160-179, 313-365);
  - K = Kuu - Kua Kaa^-1 Kua^T (solid_shell.cpp:384-458, 500-526);
  - mass rho sum_g w_g N_a(g) (solid_shell.cpp:655-663);
  - element size and the regularised failure strain (elementSize solid_shell.cpp:770-806,
    initialisePlasticState :808-815, regularisedFailureStrain plasticity.cpp:158-169):
    in-plane size sqrt(mid-surface area as two triangles), thickness volume / area;
  - R the polar factor of F = J(x, 0) J(X, 0)^-1 by R <- (R + R^-T)/2 to convergence,
    f = -R K (R^T x - X) (solid_shell.cpp:680-766);
  - dt = safety * 2 / sqrt(lambda_max(M^-1/2 K M^-1/2)) (solid_shell.cpp:1088-1134).
Parity pins: tests/test_hex_cpu.py (golden vectors of the compiled element; rigid-body
null space, patch test, rotation invariance, closed-form mass).
"""
import numpy as np

XI = np.array([-1, 1, 1, -1, -1, 1, 1, -1], np.float64)
ETA = np.array([-1, -1, 1, 1, -1, -1, 1, 1], np.float64)
ZETA = np.array([-1, -1, -1, -1, 1, 1, 1, 1], np.float64)
Q = 1.0 / np.sqrt(3.0)
GAUSS = np.array([[Q if g & 1 else -Q, Q if g & 2 else -Q, Q if g & 4 else -Q]
                  for g in range(8)])


def shape(xi, eta, zeta):
    """N_a at one natural point: [8]."""
    return 0.125 * (1 + xi * XI) * (1 + eta * ETA) * (1 + zeta * ZETA)


def dshape(xi, eta, zeta):
    """d N_a / d(xi, eta, zeta) at one natural point: [8, 3]."""
    return 0.125 * np.stack([XI * (1 + eta * ETA) * (1 + zeta * ZETA),
                             ETA * (1 + xi * XI) * (1 + zeta * ZETA),
                             ZETA * (1 + xi * XI) * (1 + eta * ETA)], axis=1)


def jacobian(x, dN):
    """J[e, i, k] = d x_i / d xi_k for nodes x [E, 8, 3]."""
    return np.einsum("eai,ak->eik", x, dN)


def covariant_rows(J, dN):
    """[E, 6, 24]: rows E11, E22, E33, 2E12, 2E23, 2E31 per DOF 3a + i, with
    E_kl = (g_k . du/dxi_l + g_l . du/dxi_k) / 2, g_k = J[:, :, k]."""
    E = J.shape[0]
    B = np.zeros((E, 6, 8, 3))
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]
    for r, (k, l) in enumerate(pairs):
        if k == l:
            B[:, r] = dN[None, :, k, None] * J[:, None, :, k]
        else:
            B[:, r] = dN[None, :, l, None] * J[:, None, :, k] + dN[None, :, k, None] * J[:, None, :, l]
    return B.reshape(E, 6, 24)


def voigt_push(A):
    """[E, 6, 6] Voigt map of eps = A^T E A (covariant to Cartesian strain, engineering
    shears in and out)."""
    E = A.shape[0]
    T = np.zeros((E, 6, 6))
    idx = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]
    for j, (k, l) in enumerate(idx):
        e = np.zeros((3, 3))
        v = 1.0 if k == l else 0.5
        e[k, l] = e[l, k] = v
        m = np.einsum("eki,kl,elj->eij", A, e, A)
        for r, (p, q) in enumerate(idx):
            T[:, r, j] = m[:, p, q] * (1.0 if p == q else 2.0)
    return T


def forms(X):
    """Rest forms of elements X [E, 8, 3]: B [E, 8, 6, 24], G [E, 8, 6, 7], w [E, 8].
    Raises on a non-positive Jacobian at the centre or a Gauss point."""
    X = np.asarray(X, np.float64).reshape(-1, 8, 3)
    E = X.shape[0]
    J0 = jacobian(X, dshape(0, 0, 0))
    det0 = np.linalg.det(J0)
    if np.any(det0 <= 0):
        raise ValueError("hex with a non-positive centre Jacobian")
    T0 = voigt_push(np.linalg.inv(J0))
    B = np.zeros((E, 8, 6, 24))
    G = np.zeros((E, 8, 6, 7))
    w = np.zeros((E, 8))
    for g, (xi, eta, zeta) in enumerate(GAUSS):
        dN = dshape(xi, eta, zeta)
        J = jacobian(X, dN)
        det = np.linalg.det(J)
        if np.any(det <= 0):
            raise ValueError("hex with a non-positive Jacobian at a Gauss point")
        w[:, g] = det
        Bc = covariant_rows(J, dN)
        # ANS transverse shear at the point's zeta
        s = {}
        for key, (a, b) in {"xm": (-1, 0), "xp": (1, 0), "ym": (0, -1), "yp": (0, 1)}.items():
            d = dshape(a, b, zeta)
            s[key] = covariant_rows(jacobian(X, d), d)
        Bc[:, 4] = 0.5 * (1 - xi) * s["xm"][:, 4] + 0.5 * (1 + xi) * s["xp"][:, 4]
        Bc[:, 5] = 0.5 * (1 - eta) * s["ym"][:, 5] + 0.5 * (1 + eta) * s["yp"][:, 5]
        # ANS thickness strain from the in-plane corners
        e33 = 0.0
        for a, b in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
            d = dshape(a, b, zeta)
            e33 = e33 + 0.25 * (1 + a * xi) * (1 + b * eta) * covariant_rows(jacobian(X, d), d)[:, 2]
        Bc[:, 2] = e33
        B[:, g] = np.einsum("eij,ejc->eic", voigt_push(np.linalg.inv(J)), Bc)
        M = np.zeros((6, 7))
        M[0, 0], M[1, 1], M[2, 2] = xi, eta, zeta
        M[2, 3], M[2, 4] = xi * zeta, eta * zeta
        M[3, 5], M[3, 6] = xi, eta
        G[:, g] = (det0 / det)[:, None, None] * np.einsum("eij,jk->eik", T0, M)
    norm = np.sqrt(np.einsum("eg,egik,egik->ek", w, G, G))
    G = G / np.where(norm > 0, norm, 1.0)[:, None, None, :]
    return B, G, w


def elasticity(E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[np.arange(3), np.arange(3)] += 2 * mu
    C[np.arange(3, 6), np.arange(3, 6)] = mu
    return C


def stiffness(X, E, nu):
    """Condensed element stiffness [E, 24, 24] of elements X [E, 8, 3]."""
    B, G, w = forms(X)
    C = elasticity(E, nu)
    Kuu = np.einsum("eg,egia,ij,egjb->eab", w, B, C, B)
    Kua = np.einsum("eg,egia,ij,egjk->eak", w, B, C, G)
    Kaa = np.einsum("eg,egik,ij,egjl->ekl", w, G, C, G)
    return Kuu - Kua @ np.linalg.solve(Kaa, np.swapaxes(Kua, 1, 2))


def mass(X, rho):
    """Lumped nodal mass [E, 8]: rho sum_g w_g N_a(g)."""
    X = np.asarray(X, np.float64).reshape(-1, 8, 3)
    out = np.zeros((X.shape[0], 8))
    for xi, eta, zeta in GAUSS:
        det = np.linalg.det(jacobian(X, dshape(xi, eta, zeta)))
        out += rho * det[:, None] * shape(xi, eta, zeta)[None, :]
    return out


def element_size(X):
    """(in-plane size, thickness) [E] each: sqrt of the mid-surface quad's area over its two
    triangles (0-1-2, 0-2-3), and the Gauss volume over that area; zeros where the area
    is not positive."""
    X = np.asarray(X, np.float64).reshape(-1, 8, 3)
    mid = 0.5 * (X[:, :4] + X[:, 4:])
    area = np.zeros(X.shape[0])
    for b, c in ((1, 2), (2, 3)):
        area += 0.5 * np.linalg.norm(np.cross(mid[:, b] - mid[:, 0], mid[:, c] - mid[:, 0]), axis=1)
    vol = mass(X, 1.0).sum(axis=1)
    ok = area > 0
    length = np.where(ok, np.sqrt(np.where(ok, area, 1.0)), 0.0)
    thick = np.where(ok, vol / np.where(ok, area, 1.0), 0.0)
    return length, thick


def failure_strain(X, uniform, fracture):
    """[E] the failure strain each element resolves: uniform + (fracture - uniform) *
    min(1, t / l), fracture when the size is degenerate."""
    length, thick = element_size(X)
    share = np.where((length > 0) & (thick > 0), np.minimum(1.0, thick / np.where(length > 0, length, 1.0)), 1.0)
    return uniform + (fracture - uniform) * share


def rotation(X, x, tol=1e-16, iterations=64):
    """Polar factor of the centre deformation gradient [E, 3, 3] (row-major R[i, j])."""
    X = np.asarray(X, np.float64).reshape(-1, 8, 3)
    x = np.asarray(x, np.float64).reshape(-1, 8, 3)
    d0 = dshape(0, 0, 0)
    R = jacobian(x, d0) @ np.linalg.inv(jacobian(X, d0))
    for _ in range(iterations):
        nxt = 0.5 * (R + np.swapaxes(np.linalg.inv(R), 1, 2))
        change = np.abs(nxt - R).max()
        R = nxt
        if change < tol:
            break
    return R


def force(K, X, x):
    """f = -R K (R^T x - X) per element: [E, 24]."""
    X = np.asarray(X, np.float64).reshape(-1, 8, 3)
    x = np.asarray(x, np.float64).reshape(-1, 8, 3)
    R = rotation(X, x)
    d = (np.einsum("eji,eaj->eai", R, x) - X).reshape(-1, 24)
    kd = np.einsum("eab,eb->ea", K, d).reshape(-1, 8, 3)
    return -np.einsum("eij,eaj->eai", R, kd).reshape(-1, 24)


def critical_timestep(K, m, safety=0.9):
    """safety * 2 / omega_max per element, omega_max^2 the largest eigenvalue of
    M^-1/2 K M^-1/2 (taken here by a symmetric eigensolver rather than power
    iteration: the same number to the power iteration's tolerance)."""
    s = 1.0 / np.sqrt(np.repeat(m, 3, axis=1))
    A = s[:, :, None] * K * s[:, None, :]
    A = 0.5 * (A + np.swapaxes(A, 1, 2))
    lam = np.linalg.eigvalsh(A)[:, -1]
    return safety * 2.0 / np.sqrt(lam)


def plate_mesh(lx, ly, thickness, nx, ny, nz=1):
    """solidshell::makePlateMesh (solid_shell.cpp:1152-1185): nodes with the thickness
    index fastest, then y, then x; mid-surface at z = 0."""
    sx, sy, sz = nx + 1, ny + 1, nz + 1
    i, j, k = np.meshgrid(np.arange(sx), np.arange(sy), np.arange(sz), indexing="ij")
    X = np.stack([lx * i / nx, ly * j / ny, -0.5 * thickness + thickness * k / nz], -1).reshape(-1, 3)
    at = lambda i, j, k: (i * sy + j) * sz + k  # noqa: E731
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    index = np.stack([at(i, j, k), at(i + 1, j, k), at(i + 1, j + 1, k), at(i, j + 1, k),
                      at(i, j, k + 1), at(i + 1, j, k + 1), at(i + 1, j + 1, k + 1),
                      at(i, j + 1, k + 1)], 1).astype(np.uint32)
    return X, index
