"""Synthetic P1 finite-element test systems in closed form (numpy, vectorised).

All matrices are 0-based CSR with ascending column indices per row, unknowns
ordered lexicographically (x fastest), Dirichlet nodes eliminated.

* ``poisson2d_p1(m)``: -Δu on the unit square, P1 on the uniform right-triangle
  mesh with m x m cells.  Interior stiffness is the 5-point stencil
  [4, -1, -1, -1, -1]; the couplings along the triangle diagonals are exactly
  zero and are not stored (m = 128 gives n = 16,129, nnz = 80,137).
* ``poisson3d_kuhn(m)``: -Δu on the unit cube, P1 on m^3 cubes each split into
  the 6 Kuhn tetrahedra along the (1,1,1) diagonal.  Stiffness is
  h * (6 on the diagonal, -1 on the 6 axis neighbours, 0 on the 8 neighbours
  along the face and body diagonals of the Kuhn split); the zeros ARE stored
  (15-point rows), as an assembled FE matrix holds them (m = 128 gives
  n = 2,048,383, nnz = 30,340,171).
* ``poisson3d_kuhn_coef(m, nblocks, seed)``: -div(kappa grad u) on the same
  mesh with kappa piecewise constant on nblocks^3 blocks, log-uniform in
  [1e-3, 1e3] (same pattern as poisson3d_kuhn).
* ``aniso2d_p1(m, theta, eps)`` / ``aniso3d_kuhn(m, theta, eps)``: rotated
  anisotropic diffusion (tensor diag(1, eps) rotated by theta in the x-y plane)
  on the same meshes.
* ``uniform_rhs(n, seed)``: f_i in [-1, 1) from splitmix64(seed XOR i), the
  same bits in every language.
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """Steele, Lea & Flood (2014) SplitMix64 finaliser applied to x + golden gamma."""
    with np.errstate(over="ignore"):
        z = (x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform_rhs(n: int, seed: int = 1) -> np.ndarray:
    """Counter-based U[-1, 1): f_i = 2 * (splitmix64(seed ^ i) >> 11) * 2^-53 - 1."""
    i = np.arange(n, dtype=np.uint64)
    z = splitmix64(np.uint64(seed) ^ i)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def _stencil_csr(shape, offsets, values):
    """CSR of a stencil on the interior grid `shape` (z, y, x order), offsets
    given as (dz, dy, dx) tuples with a constant or per-row value array each;
    entries outside the grid are dropped.
    Offsets are sorted by their linear index offset, so columns ascend."""
    dims = len(shape)
    nz, ny, nx = (1,) * (3 - dims) + tuple(shape)
    n = nz * ny * nx
    lin = [dz * ny * nx + dy * nx + dx for (dz, dy, dx) in offsets]
    order = np.argsort(lin, kind="stable")
    offsets = [offsets[k] for k in order]
    values = [values[k] for k in order]
    idx = np.arange(n, dtype=np.int64)
    x = idx % nx
    y = (idx // nx) % ny
    z = idx // (nx * ny)
    cols = np.empty((len(offsets), n), dtype=np.int64)
    valid = np.empty((len(offsets), n), dtype=bool)
    vals = np.empty((len(offsets), n), dtype=np.float64)   # a value per offset, or per row
    for k, (dz, dy, dx) in enumerate(offsets):
        xx, yy, zz = x + dx, y + dy, z + dz
        valid[k] = (xx >= 0) & (xx < nx) & (yy >= 0) & (yy < ny) & (zz >= 0) & (zz < nz)
        cols[k] = zz * ny * nx + yy * nx + xx
        vals[k] = values[k]
    counts = valid.sum(axis=0)
    ia = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=ia[1:])
    vt = valid.T.reshape(-1)
    ja = cols.T.reshape(-1)[vt].astype(np.int32)
    va = vals.T.reshape(-1)[vt]
    return ia.astype(np.int32), ja, va


def poisson2d_p1(m: int = 128):
    """5-point P1 Laplacian on (m-1)^2 interior nodes."""
    k = m - 1
    offs = [(0, 0, 0), (0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0)]
    vals = [4.0, -1.0, -1.0, -1.0, -1.0]
    return _stencil_csr((k, k), offs, vals)


KUHN_DIAGONALS = [(0, 1, 1), (1, 0, 1), (1, 1, 0), (1, 1, 1)]


def poisson3d_kuhn(m: int = 128):
    """15-point P1 Kuhn-split Laplacian on (m-1)^3 interior nodes (zeros stored).
    Offsets are (dz, dy, dx); the Kuhn diagonals (dx,dy,dz) in {(1,1,0),(1,0,1),
    (0,1,1),(1,1,1)} and their negatives carry exact zeros."""
    k = m - 1
    h = 1.0 / m
    offs = [(0, 0, 0)]
    vals = [6.0 * h]
    for ax in [(0, 0, 1), (0, 1, 0), (1, 0, 0)]:
        for s in (1, -1):
            offs.append(tuple(s * a for a in ax))
            vals.append(-h)
    for d in KUHN_DIAGONALS:
        for s in (1, -1):
            offs.append(tuple(s * a for a in d))
            vals.append(0.0)
    return _stencil_csr((k, k, k), offs, vals)


# 6 x the P1 stiffness of one unit cube split into the 6 Kuhn tetrahedra
# (vertex v = x + 2y + 4z); kuhn_element_stiffness_check() derives it from the
# tetrahedra and the tests compare the two.
KUHN_CUBE_K6 = np.array([
    [6, -2, -2, 0, -2, 0, 0, 0],
    [-2, 4, 0, -1, 0, -1, 0, 0],
    [-2, 0, 4, -1, 0, 0, -1, 0],
    [0, -1, -1, 4, 0, 0, 0, -2],
    [-2, 0, 0, 0, 4, -1, -1, 0],
    [0, -1, 0, 0, -1, 4, 0, -2],
    [0, 0, -1, 0, -1, 0, 4, -2],
    [0, 0, 0, -2, 0, -2, -2, 6]], dtype=np.float64)


def block_coefficients(nblocks: int = 8, seed: int = 1, log10_lo: float = -3.0,
                       log10_hi: float = 3.0) -> np.ndarray:
    """kappa on nblocks^3 blocks, log-uniform in [10^lo, 10^hi]:
    kappa_b = 10^(lo + (hi - lo) u_b), u_b from splitmix64(seed ^ b), b = bx +
    nb (by + nb bz).  Returned as an array indexed [bz, by, bx]."""
    nb = nblocks
    b = np.arange(nb ** 3, dtype=np.uint64)
    u = (splitmix64(np.uint64(seed) ^ b) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return (10.0 ** (log10_lo + (log10_hi - log10_lo) * u)).reshape(nb, nb, nb)


def poisson3d_kuhn_coef(m: int = 128, nblocks: int = 8, seed: int = 1, log10_lo: float = -3.0,
                        log10_hi: float = 3.0):
    """-div(kappa grad u) on the unit cube, P1 on the m^3 Kuhn-split cubes, kappa
    piecewise constant on nblocks^3 blocks of cubes (block_coefficients).  Same
    15-point pattern as poisson3d_kuhn (the Kuhn-diagonal couplings stay exact
    zeros: they vanish cube by cube); entry (p, p+d) = h/6 sum over the cubes
    holding both nodes of kappa_cube * K6[corner(p), corner(p+d)].  With kappa = 1
    it equals poisson3d_kuhn(m).  Returns (ia, ja, val, kappa_blocks)."""
    k = m - 1
    h = 1.0 / m
    kb = block_coefficients(nblocks, seed, log10_lo, log10_hi)
    cell_block = (np.arange(m) * nblocks) // m
    kc = kb[np.ix_(cell_block, cell_block, cell_block)]      # [Z, Y, X] per cube
    offs = [(0, 0, 0)]
    for ax in [(0, 0, 1), (0, 1, 0), (1, 0, 0)]:
        for sgn in (1, -1):
            offs.append(tuple(sgn * a for a in ax))
    for d in KUHN_DIAGONALS:
        for sgn in (1, -1):
            offs.append(tuple(sgn * a for a in d))
    vals = []
    for (dz, dy, dx) in offs:
        acc = np.zeros((k, k, k))
        for cz in (0, 1):
            for cy in (0, 1):
                for cx in (0, 1):
                    ex, ey, ez = cx + dx, cy + dy, cz + dz
                    if not (0 <= ex <= 1 and 0 <= ey <= 1 and 0 <= ez <= 1):
                        continue
                    w = KUHN_CUBE_K6[cx + 2 * cy + 4 * cz, ex + 2 * ey + 4 * ez]
                    if w == 0.0:
                        continue
                    acc += w * kc[1 - cz:m - cz, 1 - cy:m - cy, 1 - cx:m - cx]
        vals.append((h * acc.reshape(-1)) / 6.0)
    ia, ja, va = _stencil_csr((k, k, k), offs, vals)
    return ia, ja, va, kb


def _p1_cell_matrix(dim: int, K: np.ndarray) -> np.ndarray:
    """P1 stiffness (unit cell) with constant tensor K: the square split along
    (1,1) into two triangles (2-D, vertex x + 2y) or the cube split into the 6
    Kuhn tetrahedra (3-D, vertex x + 2y + 4z)."""
    import itertools
    nv = 2 ** dim
    verts = np.array([[(v >> a) & 1 for a in range(dim)] for v in range(nv)], float)
    M = np.zeros((nv, nv))
    for perm in itertools.permutations(range(dim)):
        p = np.zeros(dim)
        simplex = [0]
        for ax in perm:
            p = p.copy()
            p[ax] = 1
            simplex.append(int(sum(int(p[a]) << a for a in range(dim))))
        X = verts[simplex]
        T = np.hstack([np.ones((dim + 1, 1)), X])
        G = np.linalg.inv(T)[1:, :]
        vol = abs(np.linalg.det(T)) / (2.0 if dim == 2 else 6.0)
        M[np.ix_(simplex, simplex)] += vol * G.T @ K @ G
    return M


def anisotropic_tensor(dim: int, theta: float, eps: float) -> np.ndarray:
    """R(theta) diag(1, eps[, eps]) R(theta)^T, the rotation in the x-y plane."""
    c, s = np.cos(theta), np.sin(theta)
    R = np.eye(dim)
    R[:2, :2] = [[c, -s], [s, c]]
    return R @ np.diag([1.0] + [eps] * (dim - 1)) @ R.T


def _constant_cell_stencil(dim: int, m: int, Mc: np.ndarray, scale: float):
    """CSR of the assembled operator on the (m-1)^dim interior nodes: entry
    (p, p+d) = scale * sum over cells holding both nodes of Mc[corner(p), corner(p+d)]."""
    k = m - 1
    if dim == 2:
        offs3 = [(0, 0, 0), (0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (0, 1, 1), (0, -1, -1)]
    else:
        offs3 = [(0, 0, 0)]
        for ax in [(0, 0, 1), (0, 1, 0), (1, 0, 0)]:
            for sg in (1, -1):
                offs3.append(tuple(sg * a for a in ax))
        for d in KUHN_DIAGONALS:
            for sg in (1, -1):
                offs3.append(tuple(sg * a for a in d))
    vals = []
    for (dz, dy, dx) in offs3:
        v = 0.0
        for c in range(2 ** dim):
            cx, cy, cz = c & 1, (c >> 1) & 1, (c >> 2) & 1
            ex, ey, ez = cx + dx, cy + dy, cz + dz
            if not (0 <= ex <= 1 and 0 <= ey <= 1 and 0 <= ez <= 1) or (dim == 2 and ez != 0):
                continue
            v += Mc[c, ex + 2 * ey + 4 * ez]
        vals.append(scale * v)
    shape = (k, k) if dim == 2 else (k, k, k)
    return _stencil_csr(shape, offs3, vals)


def aniso2d_p1(m: int, theta: float, eps: float):
    """-div(K grad u), K = R(theta) diag(1, eps) R(theta)^T, on the unit square
    with m x m cells each split along (1,1) into two P1 triangles; 7-point rows
    on the (m-1)^2 interior nodes (scale-free in 2-D)."""
    Mc = _p1_cell_matrix(2, anisotropic_tensor(2, theta, eps))
    return _constant_cell_stencil(2, m, Mc, 1.0)


def aniso3d_kuhn(m: int, theta: float, eps: float):
    """-div(K grad u), K = R_z(theta) diag(1, eps, eps) R_z(theta)^T, on the unit
    cube with m^3 Kuhn-split cubes; 15-point rows (h-scaled)."""
    Mc = _p1_cell_matrix(3, anisotropic_tensor(3, theta, eps))
    return _constant_cell_stencil(3, m, Mc, 1.0 / m)


def kuhn_element_stiffness_check():
    """Assembles the P1 stiffness of one unit cube split into the 6 Kuhn
    tetrahedra and returns the 8x8 matrix (vertex v = x + 2y + 4z).  Used by the
    tests to confirm the stencil of poisson3d_kuhn from first principles."""
    import itertools
    verts = np.array([[(v >> 0) & 1, (v >> 1) & 1, (v >> 2) & 1] for v in range(8)], float)
    K = np.zeros((8, 8))
    for perm in itertools.permutations(range(3)):
        # path 0 -> e_p0 -> e_p0+e_p1 -> (1,1,1)
        p = np.zeros(3)
        tet = [0]
        for ax in perm:
            p = p.copy()
            p[ax] = 1
            tet.append(int(p[0] + 2 * p[1] + 4 * p[2]))
        X = verts[tet]
        T = np.hstack([np.ones((4, 1)), X])
        G = np.linalg.inv(T)[1:, :]          # gradients of barycentric coords (3x4)
        vol = abs(np.linalg.det(T)) / 6.0
        K[np.ix_(tet, tet)] += vol * G.T @ G
    return K


def batch_anisotropic_2d(count: int = 256, seed: int = 4, side_range=(33, 513)):
    """The batched workload: `count` 2-D P1 anisotropic diffusion systems with
    grid sides (nodes per side, boundary included) log-uniform in side_range,
    theta ~ U[0, pi), eps = 10^U[-4, 0].  Returns a list of dicts with the
    parameters; the matrix of entry d is aniso2d_p1(d["side"] - 1, d["theta"], d["eps"])."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(side_range[0]), np.log(side_range[1])
    out = []
    for k in range(count):
        side = int(np.clip(np.rint(np.exp(rng.uniform(lo, hi))), side_range[0], side_range[1]))
        out.append(dict(side=side, theta=float(rng.uniform(0.0, np.pi)), eps=float(10.0 ** rng.uniform(-4.0, 0.0)),
                        index=k))
    return out


def batch_rhs(n: int, system: int, nrhs: int = 8, seed: int = 4):
    """nrhs seeded U[-1, 1) right-hand sides of system `system`, shape (nrhs, n)."""
    return np.stack([uniform_rhs(n, seed=(seed * 1000003 + system) * 16 + q) for q in range(nrhs)])
