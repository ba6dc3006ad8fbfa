"""Restatement of the reference's adjoint-source tool (test infrastructure only).

``src/auxiliaries/adj_seismogram.f90`` (xadj_seismogram tstart tend station comp): cross-
correlation travel-time adjoint source of a displacement trace in a window [tstart, tend]
with a parabolic taper (window_taper_type = 1).  Only the traces the tool writes matter:
the selected component carries ft_bar, the other one zeros.
"""
import math

import numpy as np


def traveltime_adjoint_source(time, seism, tstart, tend):
    """time, seism: the trace as written in OUTPUT_FILES/*.semd (double).  Returns ft_bar."""
    time = np.asarray(time, np.float64)
    seism = np.asarray(seism, np.float64)
    nstep = len(time)
    t0 = time[0]
    deltat = (time[nstep - 1] - t0) / float(nstep - 1)
    istart = math.floor((tstart - t0) / deltat) + 1
    iend = math.ceil((tend - t0) / deltat) + 1
    istart = min(max(istart, 1), nstep - 1) if istart <= nstep else nstep - 1
    if iend < istart:
        iend = istart + 1
    iend = min(iend, nstep)
    window = np.zeros(nstep)
    for it in range(istart, iend + 1):  # 1-based
        window[it - 1] = 1.0 - (2 * (float(it) - istart) / (iend - istart) - 1.0) ** 2
    veloc = np.zeros(nstep)
    veloc[1:-1] = (seism[2:] - seism[:-2]) / (2 * deltat)
    veloc[0] = (seism[1] - seism[0]) / deltat
    veloc[-1] = (seism[-1] - seism[-2]) / deltat
    accel = np.zeros(nstep)
    accel[1:-1] = (veloc[2:] - veloc[:-2]) / (2 * deltat)
    accel[0] = (veloc[1] - veloc[0]) / deltat
    accel[-1] = (veloc[-1] - veloc[-2]) / deltat
    # Fortran sum() of the product, in order
    nnorm = 0.0
    prod = window * seism * accel
    for x in prod:
        nnorm += x
    nnorm *= deltat
    if abs(nnorm) > 1e-40:
        return veloc * window / nnorm
    return np.zeros(nstep)
