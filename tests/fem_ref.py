"""Independent finite-element model used to pin the oracle.

Nothing here calls the oracle.  Element matrices come from a 2-point tensor
rule whose 1-D moments (orders 0..2) equal those of the reference's n-point
midpoint rule, so for the brick element (integrand of degree <= 2 in each
variable) both integrate identically; global operators are assembled with
scipy element by element; transfer operators are built from node
coordinates.  Conventions are the reference's: halo node layout
i*wrapy*wrapz + k*wrapy + j (stencil_utility.h:8-52), local node order
stencil_utility.h:19-26, y-displacement measured along -y (the sign that
local_matrix.c:298 produces).
"""
import itertools

import numpy as np
import scipy.sparse as sp

# local node positions (x, y, z) in an element, stencil_utility.h:19-26
LNODE = np.array([[0, 1, 0], [1, 1, 0], [1, 0, 0], [0, 0, 0],
                  [0, 1, 1], [1, 1, 1], [1, 0, 1], [0, 0, 1]])
E0, EMIN, NU = 1.0, 1e-6, 0.3
YFLIP = np.diag(np.tile([1.0, -1.0, 1.0], 8))


def cmat(nu=NU):
    t1 = (1 - nu) / ((1 + nu) * (1 - 2 * nu))
    t2 = nu / ((1 + nu) * (1 - 2 * nu))
    t3 = 1 / (2 * (1 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = t2
    np.fill_diagonal(C[:3, :3], t1)
    C[3, 3] = C[4, 4] = C[5, 5] = t3
    return C


def element_matrix(H, box, npts=10, nu=NU):
    """Stiffness of a cube of side H restricted to the sub-box `box` (unit
    coordinates), grid frame, y-displacement along -y."""
    t = np.sqrt(1.0 / 3.0 - 1.0 / (3.0 * npts ** 2))
    C = cmat(nu)
    K = np.zeros((24, 24))
    pts = []
    for a, b in box:
        c, hw = (a + b) / 2, (b - a) / 2
        pts.append([c - hw * t, c + hw * t])
    vol = np.prod([b - a for a, b in box]) * H ** 3
    for q in itertools.product(*pts):
        B = np.zeros((6, 24))
        for n in range(8):
            p = LNODE[n]
            f = [q[d] if p[d] else 1 - q[d] for d in range(3)]
            df = [1.0 if p[d] else -1.0 for d in range(3)]
            dx, dy, dz = (df[0] * f[1] * f[2] / H, f[0] * df[1] * f[2] / H,
                          f[0] * f[1] * df[2] / H)
            B[0, 3 * n], B[1, 3 * n + 1], B[2, 3 * n + 2] = dx, dy, dz
            B[3, 3 * n], B[3, 3 * n + 1] = dy, dx
            B[4, 3 * n + 1], B[4, 3 * n + 2] = dz, dy
            B[5, 3 * n], B[5, 3 * n + 2] = dz, dx
        K += vol / 8 * B.T @ C @ B
    return YFLIP @ K @ YFLIP


def subcell_matrices(l, npts=10):
    """KE_l[cell], cell = ncell^2*ii + ncell*kk + jj (ii along x, kk along z,
    jj along y), each the part of the level-l element over one sub-cell."""
    n = 1 << l
    out = np.zeros((n ** 3, 24, 24))
    for ii, kk, jj in itertools.product(range(n), repeat=3):
        box = [(ii / n, (ii + 1) / n), (jj / n, (jj + 1) / n), (kk / n, (kk + 1) / n)]
        out[n * n * ii + n * kk + jj] = element_matrix(float(n), box, npts)
    return out


class Layout:
    """Halo layout of level l of a fine grid (nelx, nely, nelz)."""

    def __init__(self, nelx, nely, nelz, l, yalign=1):
        n = 1 << l
        self.l, self.ncell = l, n
        self.nelx, self.nely, self.nelz = nelx // n, nely // n, nelz // n
        pady = (yalign - ((self.nely + 1) % yalign)) % yalign
        self.wx, self.wy, self.wz = self.nelx + 3, self.nely + pady + 3, self.nelz + 3
        self.ndof = 3 * self.wx * self.wy * self.wz

    def node(self, i, j, k):
        return (i * self.wy * self.wz + k * self.wy + j)

    def interior_nodes(self):
        return [(i, j, k) for i in range(1, self.nelx + 2)
                for k in range(1, self.nelz + 2) for j in range(1, self.nely + 2)]

    def interior_dofs(self):
        idx = np.array([self.node(*t) for t in self.interior_nodes()])
        return (3 * idx[:, None] + np.arange(3)).ravel()

    def fixed_dofs(self):
        """New-cantilever support (stencil_grid_utility.c:34-59)."""
        lim = self.nely // 4 + 1
        ys = list(range(1, lim + 1)) + list(range(self.nely + 2 - lim, self.nely + 2))
        out = [3 * self.node(1, j, k) + c for k in range(1, self.nelz + 2)
               for j in ys for c in range(3)]
        return np.array(sorted(out))


def fine_elem_index(lay0, i, j, k):
    """Element (i,j,k) (halo, 1-based) of the fine grid."""
    return i * (lay0.wy - 1) * (lay0.wz - 1) + k * (lay0.wy - 1) + j


def young(x):
    return EMIN + x ** 3 * (E0 - EMIN)


def assemble(nelx, nely, nelz, l, x, yalign=1, npts=10):
    """Pre-BC global K_l (padded halo numbering) from sub-cell matrices and
    fine densities x (fine padded element array)."""
    L0 = Layout(nelx, nely, nelz, 0, yalign)
    L = Layout(nelx, nely, nelz, l, yalign)
    KE = subcell_matrices(l, npts)
    n = L.ncell
    rows, cols, vals = [], [], []
    for i in range(1, L.nelx + 1):
        for k in range(1, L.nelz + 1):
            for j in range(1, L.nely + 1):
                ke = np.zeros((24, 24))
                for ii, kk, jj in itertools.product(range(n), repeat=3):
                    e = fine_elem_index(L0, (i - 1) * n + ii + 1, (j - 1) * n + jj + 1,
                                        (k - 1) * n + kk + 1)
                    ke += young(x[e]) * KE[n * n * ii + n * kk + jj]
                dofs = np.array([3 * L.node(i + p[0], j + p[1], k + p[2]) + c
                                 for p in LNODE for c in range(3)])
                rows.append(np.repeat(dofs, 24))
                cols.append(np.tile(dofs, 24))
                vals.append(ke.ravel())
    K = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(L.ndof, L.ndof))
    return K


def bc_identity_rows(K, fixed):
    """Matrix-free BC convention: fixed rows replaced by identity, columns
    kept (stencil_methods.c:424-427)."""
    K = K.tolil()
    for f in fixed:
        K.rows[f] = [f]
        K.data[f] = [1.0]
    return K.tocsr()


def prolongation(nelx, nely, nelz, l, yalign=1):
    """Trilinear interpolation from level l+1 nodes to level l interior nodes
    built from node coordinates (coarse node ic at fine coordinate 2(ic-1))."""
    F = Layout(nelx, nely, nelz, l, yalign)
    C = Layout(nelx, nely, nelz, l + 1, yalign)
    rows, cols, vals = [], [], []
    for (i, j, k) in F.interior_nodes():
        w1 = []
        for t in (i - 1, j - 1, k - 1):
            lo = t // 2
            if t % 2 == 0:
                w1.append([(lo + 1, 1.0)])
            else:
                w1.append([(lo + 1, 0.5), (lo + 2, 0.5)])
        for (ci, wi), (cj, wj), (ck, wk) in itertools.product(*w1):
            for c in range(3):
                rows.append(3 * F.node(i, j, k) + c)
                cols.append(3 * C.node(ci, cj, ck) + c)
                vals.append(wi * wj * wk)
    return sp.csr_matrix((vals, (rows, cols)), shape=(F.ndof, C.ndof))


def filter_matrix(nelx, nely, nelz, rmin, yalign=1):
    """Density filter H over the padded element array, neighbourhood of
    stencil_optimization.c:250-275 (weights max(0, rmin - dist))."""
    L0 = Layout(nelx, nely, nelz, 0, yalign)
    nelem = (L0.wx - 1) * (L0.wy - 1) * (L0.wz - 1)
    R = int(np.ceil(rmin))
    rows, cols, vals = [], [], []
    for i1 in range(1, nelx + 1):
        for k1 in range(1, nelz + 1):
            for j1 in range(1, nely + 1):
                e1 = fine_elem_index(L0, i1, j1, k1)
                for i2 in range(max(i1 - R, 1), min(i1 + R, nelx) + 1):
                    for k2 in range(max(k1 - R, 1), min(k1 + R, nelz) + 1):
                        for j2 in range(max(j1 - R, 1), min(j1 + R, nely) + 1):
                            w = max(0.0, rmin - np.sqrt((i1 - i2) ** 2 + (j1 - j2) ** 2 +
                                                        (k1 - k2) ** 2))
                            if w > 0:
                                rows.append(e1)
                                cols.append(fine_elem_index(L0, i2, j2, k2))
                                vals.append(w)
    return sp.csr_matrix((vals, (rows, cols)), shape=(nelem, nelem))


def splitmix64_densities(n, seed=20241202, lo=0.1, hi=1.0):
    """Portable seeded densities: rho = lo + (hi-lo) * (u >> 11) * 2^-53."""
    mask = (1 << 64) - 1
    s = seed & mask
    out = np.empty(n)
    for i in range(n):
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z = z ^ (z >> 31)
        out[i] = lo + (hi - lo) * (z >> 11) * 2.0 ** -53
    return out


def interior_density(nelx, nely, nelz, values, yalign=1, halo=0.0):
    """Padded density array with `values` on interior elements (i,k,j order)."""
    L0 = Layout(nelx, nely, nelz, 0, yalign)
    x = np.full((L0.wx - 1) * (L0.wy - 1) * (L0.wz - 1), halo)
    t = 0
    for i in range(1, nelx + 1):
        for k in range(1, nelz + 1):
            for j in range(1, nely + 1):
                x[fine_elem_index(L0, i, j, k)] = values[t]
                t += 1
    return x
