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


def tel_cell_list(refined):
    """Cell list of a two-level telescoping grid (the layout read_tel builds,
    source/geospatial/geospatial.F90:403-795): every coarse cell of the
    [ny, nx] rectangle with refined[j, i] is split into 2x2 fine cells, so a
    coarse cell facing refined ones has two neighbours on that side and up to
    nmaxfaces = 6 faces.  Cells are numbered coarse row-major, the fine cells
    of a split cell in (SW, SE, NW, NE) order; faces are listed N, E, S, W with
    the two neighbours of a split side consecutive; one ghost cell per open
    side at the domain edge.  Returns (n, ncellsD, c2c [ncellsD, 6], level
    [n], and each cell's (x, y) centre in coarse-cell units)."""
    ny, nx = refined.shape
    ids = {}                               # (level, i, j) -> id; level 1 = fine cell on a 2*nx x 2*ny lattice
    xy = []
    for j in range(ny):
        for i in range(nx):
            if refined[j, i]:
                for (di, dj) in ((0, 0), (1, 0), (0, 1), (1, 1)):
                    ids[(1, 2 * i + di, 2 * j + dj)] = len(xy)
                    xy.append((i + 0.25 + 0.5 * di, j + 0.25 + 0.5 * dj))
            else:
                ids[(0, i, j)] = len(xy)
                xy.append((i + 0.5, j + 0.5))
    n = len(xy)
    level = np.zeros(n, dtype=np.int32)
    faces = [[] for _ in range(n)]
    dirs = ((0, 1), (1, 0), (0, -1), (-1, 0))
    for (lev, i, j), c in ids.items():
        level[c] = lev
        for (di, dj) in dirs:
            if lev == 0:
                I, J = i + di, j + dj
                if not (0 <= I < nx and 0 <= J < ny):
                    faces[c].append(None)
                elif refined[J, I]:
                    # the two fine cells of (I, J) that touch this side
                    if di:
                        fi = 2 * I + (0 if di > 0 else 1)
                        faces[c] += [ids[(1, fi, 2 * J)], ids[(1, fi, 2 * J + 1)]]
                    else:
                        fj = 2 * J + (0 if dj > 0 else 1)
                        faces[c] += [ids[(1, 2 * I, fj)], ids[(1, 2 * I + 1, fj)]]
                else:
                    faces[c].append(ids[(0, I, J)])
            else:
                fi, fj = i + di, j + dj
                I, J = fi // 2, fj // 2
                if not (0 <= I < nx and 0 <= J < ny):
                    faces[c].append(None)
                elif refined[J, I]:
                    faces[c].append(ids[(1, fi, fj)])
                else:
                    faces[c].append(ids[(0, I, J)])
    nf = max(len(f) for f in faces)
    assert nf <= 6
    g = n
    ghosts = []
    for c in range(n):
        for k, nb in enumerate(faces[c]):
            if nb is None:
                faces[c][k] = g
                ghosts.append((g, c))
                g += 1
    c2c = -np.ones((g, 6), dtype=np.int32)
    for c in range(n):
        c2c[c, :len(faces[c])] = faces[c]
    for gi, c in ghosts:
        c2c[gi, 0] = c
    return n, g, c2c, level, np.array(xy)


def dense_system(n, acoef, c2c, sp, ss, phi, relax=1.0):
    """ap*phi - sum a*phi_nb = ss (+ relaxation) over active cells, ghosts fixed;
    faces with c2c = -1 do not count."""
    A = np.zeros((n, n))
    b = np.zeros(n)
    for i in range(n):
        has = c2c[i] >= 0
        asum = acoef[i][has].sum()
        ap = (asum - sp[i]) / relax
        A[i, i] = ap
        b[i] = ss[i] + (1 - relax) * phi[i] * ap
        for k in range(c2c.shape[1]):
            j = c2c[i, k]
            if j < 0:
                continue
            if j < n:
                A[i, j] -= acoef[i, k]
            else:
                b[i] += acoef[i, k] * phi[j]
    return A, b
