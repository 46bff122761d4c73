"""Writes a Case as the case file of cms2d_amd/fortran/flow_run."""
import numpy as np


def write(case, path):
    hdr = np.array([case.nx, case.ny, case.dx, case.dy, case.dt, case.amp[0],
                    case.speed_rad_hr[0], case.phase_rad[0], float(np.float32(case.hdry)),
                    case.noptset], dtype=np.float64)
    parts = [hdr, np.asarray(case.zb, np.float64), np.asarray(case.mann, np.float64)]
    if case.noptset >= 3:
        parts += [np.asarray(case.wavestrx, np.float64), np.asarray(case.wavestry, np.float64)]
    np.concatenate(parts).tofile(path)
