"""cms2d_amd.h5read on the reference's own HDF5 files (test case
L05_Implicit_Test01: its XMDF inputs and the Base/*.h5 solution files of the
reference's run, copied to tests/golden/l05/).  Pinned by what the files
themselves fix: the model-parameter file's boundary strings and curves agree
with the card file and the .tel grid (cells 1, 221, 441 and 220, 440, 660 of
the 220 x 3 flume), the solution files' per-time extrema (Maxs, Mins, written
by the reference alongside each dataset) equal the extrema of the decoded,
decompressed and reassembled Values, and the FP64 oracle run on the case
reproduces the reference's water level and velocity at the first output.

The reference run also moves sediment and updates the bed (CALC_SEDIMENT_
TRANSPORT ON), which the hydrodynamic step leaves out: at the first output
(0.5 h) the water level agrees within 7.3e-4 m (2 % of the surface's 0.034 m
range) and u within 0.017 m/s (3 %), most near the scour hole at x = 4-6 m,
and the difference grows with the bed change (4.1e-3 m at 3 h)."""
import os

import numpy as np
import pytest

from cms2d_amd.h5read import H5File
from l05_case import DIR, NAME, l05_baseline, l05_case

TOL_ETA_L05, TOL_VEL_L05 = 1e-3, 0.025     # at 0.5 h, sediment and bed change not modelled


def test_model_parameters_and_datasets():
    mp = H5File(os.path.join(DIR, NAME + "_mp.h5"))
    b = "PROPERTIES/Model Params/"
    assert list(mp.read(b + "Boundary_#1/Cells")) == [1, 221, 441]
    assert list(mp.read(b + "Boundary_#2/Cells")) == [220, 440, 660]
    assert list(mp.read(b + "Boundary_#1/Flow")) == [0.0594, 0.0594]
    assert list(mp.read(b + "Boundary_#1/Times")) == [0.0, 100.0]
    assert list(mp.read(b + "Boundary_#2/WaterLevel")) == [-0.0122, -0.0101]
    assert list(mp.read(b + "Boundary_#2/Times")) == [0.0, 16.0]
    assert mp.read("PROPERTIES/Model")[0] == "CMS-FLOW"
    ds = H5File(os.path.join(DIR, NAME + "_datasets.h5"))
    n = ds.read("Datasets/ManningsN/Values")
    assert n.shape == (1, 660) and n.dtype == np.float32 and np.all(n == np.float32(0.02))
    assert ds.get("Datasets").attrs["Grouptype"][0] == "MULTI DATASETS"
    assert np.all(ds.read("Datasets/D50/Values") == np.float32(0.16))
    paths = [p for p, kind, _ in mp.walk() if kind == "dataset"]
    assert paths == sorted(paths) and "/PROPERTIES/Model Params/Boundary_#1/Cells" in paths


def test_solution_extrema_match_decoded_values():
    f = H5File(os.path.join(DIR, NAME + "_wse.h5"))
    g = "Test02/Water_Elevation/"
    v = f.read(g + "Values")
    assert v.shape == (31, 660)
    assert np.array_equal(f.read(g + "Times"), np.arange(31) * 0.5)
    assert np.array_equal(f.read(g + "Maxs"), v.max(axis=1))
    assert np.array_equal(f.read(g + "Mins"), v.min(axis=1))
    assert f.get("Test02/Water_Elevation").attrs["DatasetUnits"][0] == "m"
    f = H5File(os.path.join(DIR, NAME + "_vel.h5"))
    g = "Test02/Current_Velocity/"
    v = f.read(g + "Values")
    assert v.shape == (31, 660, 2)
    mag = np.sqrt((v.astype(np.float64) ** 2).sum(axis=2))
    assert np.allclose(f.read(g + "Maxs"), mag.max(axis=1), rtol=1e-6, atol=0)
    assert np.allclose(f.read(g + "Mins"), mag.min(axis=1), rtol=1e-6, atol=0)


def test_not_hdf5_rejected(tmp_path):
    p = tmp_path / "x.h5"
    p.write_bytes(b"not an hdf5 file")
    with pytest.raises(ValueError, match="not an HDF5"):
        H5File(str(p))


def test_oracle_reproduces_reference_run_at_first_output():
    from pyoracle import Oracle
    c = l05_case()
    t, wse = l05_baseline()
    tv, vel = l05_baseline("vel")
    assert t[1] == 0.5 and np.all(wse[0] == 0.0)
    o = Oracle(c, serial=1)
    for _ in range(30):
        rc, _ = o.step(1)
        assert rc == 0
    de = np.abs(o.get("eta") - wse[1]).max()
    du = np.abs(o.get("u") - vel[1][:, 0]).max()
    dv = np.abs(o.get("v") - vel[1][:, 1]).max()
    print("L05 at 0.5 h: eta", de, "u", du, "v", dv)
    assert de <= TOL_ETA_L05 and du <= TOL_VEL_L05 and dv <= TOL_VEL_L05
