"""The narrow-channel rule set of flow_wetdry (flow.F90:1342-1455) dries a cell
only when fewer of its neighbours are wet, and never re-wets one, so every
update order reaches the same fixed point.  The HIP path relies on this to
replace the reference's sequential sweep by parallel passes; this test checks
the claim on random masks by applying the rules in many random orders."""
import numpy as np


def rules_dry(w, h, i, j, hdry1, hdry2):
    ny, nx = w.shape
    g = lambda a, b: w[b, a] if 0 <= a < nx and 0 <= b < ny else 0
    n, e, s, wv = g(i, j + 1), g(i + 1, j), g(i, j - 1), g(i - 1, j)
    if n + e + s + wv <= 1:
        return True
    if ((n + s) == 0 or (e + wv) == 0) and h[j, i] < hdry1:
        return True
    ss2 = g(i, j - 2) if j > 0 else 0
    ww2 = g(i - 2, j) if i > 0 else 0
    ee2 = g(i + 2, j) if i < nx - 1 else 0
    if n + ss2 == 0 and h[j, i] < hdry2:
        return True
    if e + ww2 == 0 and h[j, i] < hdry2:
        return True
    if wv + ee2 == 0 and h[j, i] < hdry2:
        return True
    return False


def fixed_point(w0, h, order, hdry1=0.25, hdry2=0.1):
    w = w0.copy()
    ny, nx = w.shape
    while True:
        changed = False
        for t in order:
            j, i = divmod(t, nx)
            if w[j, i] and rules_dry(w, h, i, j, hdry1, hdry2):
                w[j, i] = 0
                changed = True
        if not changed:
            return w


def test_order_independent():
    rng = np.random.default_rng(11)
    for trial in range(6):
        ny, nx = 14, 17
        w0 = (rng.random((ny, nx)) > 0.3).astype(np.int8)
        h = rng.uniform(0.05, 0.4, (ny, nx))
        ref = fixed_point(w0, h, list(range(nx * ny)))
        for _ in range(4):
            assert np.array_equal(ref, fixed_point(w0, h, list(rng.permutation(nx * ny))))
