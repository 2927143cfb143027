"""G-code -> toolpath rows (reference go_melt/createPath.py).

Each row is ``x,y,z,Ljump,Ldwell,dt,P``: one L3 time step of the laser.
Ljump = 0 marks travel with the laser off (a G0 target), Ldwell = 0 marks
dwell rows.  Coordinates are written ``%.8e`` right-justified to 15 columns
(cP:199-201) so every row has the same byte length: the driver resumes a
checkpoint by seeking ``rows * line_length`` bytes into the file.
"""

from __future__ import annotations

import os
import re

import numpy as np

_MOVE = re.compile(r"(?:G(\d+)\s*X(-?\d+\.\d+|-?\d+)\s*Y(-?\d+\.\d+|-?\d+)"
                   r"(?:\s*Z(-?\d+\.\d+|-?\d+))?)")


def format_fixed(val, width=15, precision=8):
    """cP:199-201."""
    return f"{val:.{precision}e}".rjust(width)


def _row(x, y, z, jump, dwell, dt, P):
    return (f"{format_fixed(x)},{format_fixed(y)},{format_fixed(z)},"
            f"{jump:d},{dwell:d},{dt:.8e},{P:.8e}\n")


def _points(text):
    """(x, y, z, skip) per G command; z persists, G0 targets get skip = 1."""
    pts, z = [], None
    for cmd, xs, ys, zs in _MOVE.findall(text):
        if zs:
            z = float(zs)
        pts.append((float(xs), float(ys), z, 1.0 if cmd != "1" else 0.0))
    return pts


def _dwell_rows(out, x, y, z, jump, nm, dwell_t, coef, first_test_after=False):
    """Dwell rows: up to wait_time rows of dt3, then long rows of dt3*coef and
    a remainder (cP:61-104, 158-187)."""
    dt3 = nm["timestep_L3"]
    n = 0
    for k in range(int(nm["wait_time"])):
        if (k + 1) * dt3 > nm["dwell_time"]:
            break
        out.append(_row(x, y, z, jump, 0, dt3, 0))
        n += 1
    if dwell_t > 0 or first_test_after:
        rem = max(0, dwell_t - nm["wait_time"] * dt3)
        steps = int(rem / dt3 / coef)
        for _ in range(steps):
            out.append(_row(x, y, z, jump, 0, dt3 * coef, 0))
            n += 1
        small = rem - steps * dt3 * coef
        if small > 0:
            out.append(_row(x, y, z, jump, 0, small, 0))
            n += 1
    return n


def parsingGcode(Nonmesh, Properties, L2h):
    """cP:6-188: write ``Nonmesh['toolpath']`` from ``Nonmesh['gcode']`` and
    return the number of rows."""
    nm = Nonmesh
    dt3 = nm["timestep_L3"]
    coef = float(nm["dwell_time_multiplier"] * nm["subcycle_num_L2"] * nm["subcycle_num_L3"])
    dwell_t = max(nm["dwell_time"] - nm["wait_time"] * dt3, nm["dwell_time"])
    with open(nm["gcode"], "r") as fh:
        pts = _points(fh.read())
    dx = nm["laser_velocity"] * dt3
    out = []
    n = 0
    z = pts[0][2]
    x, y = pts[0][0], pts[0][1]
    jump = 1
    for a, b in zip(pts[:-1], pts[1:]):
        jump, dwell = 1, 1
        if a[2] != b[2]:
            x, y = a[0], a[1]
            n += _dwell_rows(out, x, y, z, jump, nm, dwell_t, coef)
            continue
        if b[3] == 1:
            jump = 0
        dist = float(np.linalg.norm(np.array([b[0] - a[0], b[1] - a[1]])))
        nfull = int(dist // dx)
        short = (dist % dx) / dx * dt3
        x, y, z = a[0], a[1], a[2]
        vx = nm["laser_velocity"] * (b[0] - a[0]) / dist
        vy = nm["laser_velocity"] * (b[1] - a[1]) / dist
        P = jump * Properties["laser_power"]
        for _ in range(nfull):
            x += vx * dt3
            y += vy * dt3
            out.append(_row(x, y, z, jump, dwell, dt3, P))
            n += 1
        if short > 0:
            x += vx * short
            y += vy * short
            out.append(_row(x, y, z, jump, dwell, short, P))
            n += 1
    n += _dwell_rows(out, x, y, z, jump, nm, dwell_t, coef, first_test_after=True)
    path = nm["toolpath"]
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "w") as fh:
        fh.writelines(out)
    return n


def count_lines(file_path):
    """cP:191-196."""
    with open(file_path, "r") as fh:
        return sum(1 for _ in fh)


def read_rows(file_path):
    """All toolpath rows as an (n, 7) float64 array."""
    with open(file_path, "r") as fh:
        return np.array([[float(v) for v in ln.split(",")] for ln in fh if ln.strip()],
                        dtype=np.float64).reshape(-1, 7)
