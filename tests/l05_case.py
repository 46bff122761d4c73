"""The reference's own implicit test case L05_Implicit_Test01
(testing/tests_flow/L05_Implicit_Test01: L05_Implicit_Test01.cmcards, its
.tel grid and XMDF inputs, and the Base/*.h5 solution files of the
reference's run), as a Case of this package.  Data files copied to
tests/golden/l05/; the card values are restated here:

* grid: the .tel file's 660 cells of 0.1 m x 0.1 m, 220 along x by 3 along y,
  numbered row-major from the south-west, so a uniform Cartesian 220 x 3 grid;
  depth from the file's last-but-one column (read in single precision);
* Manning n: MANNINGS_N_DATASET "Datasets/ManningsN" of _datasets.h5;
* boundaries: Boundary_#2 (cells 220, 440, 660: the east column) a WSE_CURVE
  from _mp.h5 (Times, WaterLevel; OFFSET_CONSTANT 0); Boundary_#1 (cells 1,
  221, 441: the west column) a FLUX_CURVE (Times, Flow; 'm^3/s/boundary',
  CONVEYANCE 0.667);
* HYDRO_TIMESTEP 60 s, DURATION_RAMP 0.1 h, DRYING_DEPTH 0.05, wall friction
  on, SUBGRID turbulence, the default exponential scheme, no waves;
* CALC_SEDIMENT_TRANSPORT ON with MORPH_UPDATE_INTERVAL 1800 s: the bed
  changes from the first half hour on, which this hydrodynamic step does not
  model (so the baseline is matched up to the first bed update)."""
import os

import numpy as np

from cms2d_amd import cases
from cms2d_amd.h5read import H5File

DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "l05")
NAME = "L05_Implicit_Test01"


def l05_case():
    rows = [ln.split() for ln in open(os.path.join(DIR, NAME + ".tel")).read().splitlines()[2:] if ln.strip()]
    geom = np.array([[float(r[1]), float(r[2]), float(r[3]), float(r[4])] for r in rows])
    depth = np.array([float(r[13]) for r in rows], dtype=np.float32)
    nx, ny = 220, 3
    assert len(rows) == nx * ny and np.allclose(geom[:, 2:], 0.1)
    j, i = np.divmod(np.arange(nx * ny), nx)
    assert np.allclose(geom[:, 0], 0.05 + 0.1 * i) and np.allclose(geom[:, 1], 0.05 + 0.1 * j)
    ds = H5File(os.path.join(DIR, NAME + "_datasets.h5"))
    mann = ds.read("Datasets/ManningsN/Values")[0].astype(np.float64)
    mp = H5File(os.path.join(DIR, NAME + "_mp.h5"))
    b1 = "PROPERTIES/Model Params/Boundary_#1/"
    b2 = "PROPERTIES/Model Params/Boundary_#2/"
    assert list(mp.read(b1 + "Cells")) == [1, 221, 441] and list(mp.read(b2 + "Cells")) == [220, 440, 660]
    c = cases.Case(NAME, nx, ny, 0.1, 0.1, (-depth).astype(np.float64), mann, None, None, 1, [], [], [],
                   60.0, 900, rampdur=0.1, hdry=0.05)
    c.wse_times = list(mp.read(b2 + "Times"))
    c.wse_values = list(mp.read(b2 + "WaterLevel"))
    c.q_side = 3
    c.q_times = list(mp.read(b1 + "Times"))
    c.q_values = list(mp.read(b1 + "Flow"))
    c.q_cmvel = 0.667
    return c


def l05_tel_case():
    """The same case on its own .tel grid (the reference reads it with
    read_tel): the file's neighbour slots N1 N2 E1 E2 S1 S2 W1 W2 (columns 6-13)
    and the two cellstrings as cell lists; every cell is its own base cell."""
    c = l05_case()
    rows = [ln.split() for ln in open(os.path.join(DIR, NAME + ".tel")).read().splitlines()[2:] if ln.strip()]
    geom = np.array([[float(r[1]), float(r[2]), float(r[3]), float(r[4])] for r in rows])
    nb = np.array([[int(v) for v in r[5:13]] for r in rows], dtype=np.int32)
    n = len(rows)
    tel = cases.TelGrid(geom, nb, np.arange(n, dtype=np.int32), np.zeros(n, dtype=np.int32),
                        np.array([219, 439, 659], dtype=np.int32))
    tel.flux_cells = np.array([0, 220, 440], dtype=np.int32)
    c.tel = tel
    return c


def l05_baseline(field="wse"):
    """(times [hr], values [ntimes, 660]) of the reference's run: water level,
    or the velocity components."""
    if field == "wse":
        f = H5File(os.path.join(DIR, NAME + "_wse.h5"))
        g = "Test02/Water_Elevation/"
        return f.read(g + "Times"), f.read(g + "Values").astype(np.float64)
    f = H5File(os.path.join(DIR, NAME + "_vel.h5"))
    for path, kind, node in f.walk():
        if kind == "dataset" and path.endswith("/Values"):
            return f.read(path[:-6] + "Times"), node.read().astype(np.float64)
    raise KeyError("velocity")
