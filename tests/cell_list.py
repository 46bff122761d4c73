"""Builds the reference's general cell list for a rectangle with land cells
(geo_cart, source/geospatial/geospatial.F90:2709-2930): active cells numbered
row-major, one ghost cell appended per missing neighbour (N,E,S,W order)."""
import numpy as np


def cart_cell_list(active):
    ny, nx = active.shape
    idmap = -np.ones((ny, nx), dtype=np.int64)
    n = 0
    for j in range(ny):
        for i in range(nx):
            if active[j, i]:
                idmap[j, i] = n
                n += 1
    c2c = -np.ones((n, 4), dtype=np.int64)
    for j in range(ny):
        for i in range(nx):
            c = idmap[j, i]
            if c < 0:
                continue
            for k, (di, dj) in enumerate(((0, 1), (1, 0), (0, -1), (-1, 0))):
                ii, jj = i + di, j + dj
                if 0 <= ii < nx and 0 <= jj < ny and idmap[jj, ii] >= 0:
                    c2c[c, k] = idmap[jj, ii]
    g = n
    ghosts = []
    for c in range(n):
        for k in range(4):
            if c2c[c, k] < 0:
                c2c[c, k] = g
                ghosts.append((g, c, k))
                g += 1
    full = -np.ones((g, 4), dtype=np.int32)
    full[:n] = c2c
    for gi, c, k in ghosts:
        full[gi, (k + 2) % 4] = c
    return n, g, full


def dense_system(n, acoef, c2c, sp, ss, phi, relax=1.0):
    """ap*phi - sum a*phi_nb = ss (+ relaxation) over active cells, ghosts fixed."""
    A = np.zeros((n, n))
    b = np.zeros(n)
    for i in range(n):
        asum = acoef[i].sum()
        ap = (asum - sp[i]) / relax
        A[i, i] = ap
        b[i] = ss[i] + (1 - relax) * phi[i] * ap
        for k in range(4):
            j = c2c[i, k]
            if j < n:
                A[i, j] -= acoef[i, k]
            else:
                b[i] += acoef[i, k] * phi[j]
    return A, b
