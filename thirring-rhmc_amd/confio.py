"""On-disk formats of the reference: gauge configurations (`con`), the
runtime input `midout`, and the rational-approximation files (`remez*`).

`con` holds theta(Lx,Ly,Lt,3) as default real followed by the RNG seed as
real(dp) (dwf3d_lib.F90:618-717).  Two byte layouts exist:

  serial  (Fortran sequential record, dwf3d_lib.F90:652-655, 711-714):
          [int32 N][theta: 3V float32][seed float64][int32 N],  N = 3V*4 + 8
  MPI-IO  (dwf3d_lib.F90:634-648, 688-703):
          [theta: 3V float32][4 unwritten bytes][seed float64][int32 N]

Both are 3V*4 + 16 bytes long; the serial form starts with the record
marker N, the MPI form with theta(1,1,1,1).  An MPI build that reads a
serial file therefore sees theta shifted by one element (a reference quirk);
`read_con` detects the layout instead.
"""
import os
import struct

import numpy as np


def con_nbytes(lx, ly, lt):
    return 3 * lx * ly * lt * 4 + 16


def detect_layout(buf, lx, ly, lt):
    n = 3 * lx * ly * lt * 4 + 8
    if len(buf) != n + 8:
        raise ValueError(f"con size {len(buf)} does not match {lx}x{ly}x{lt} ({n + 8} bytes)")
    head = struct.unpack("<i", buf[:4])[0]
    return "serial" if head == n else "mpi"


def read_con(path, lx, ly, lt, layout="auto"):
    """Returns (theta (lx,ly,lt,3) float32 Fortran order, seed float, layout)."""
    with open(path, "rb") as fh:
        buf = fh.read()
    if layout == "auto":
        layout = detect_layout(buf, lx, ly, lt)
    n3 = 3 * lx * ly * lt
    off = 4 if layout == "serial" else 0
    theta = np.frombuffer(buf, dtype="<f4", count=n3, offset=off)
    seed = struct.unpack("<d", buf[n3 * 4 + 4:n3 * 4 + 12])[0]
    return np.asfortranarray(theta.reshape((lx, ly, lt, 3), order="F").copy()), seed, layout


def write_con(path, theta, seed, layout="serial"):
    lx, ly, lt, _ = theta.shape
    n = 3 * lx * ly * lt * 4 + 8
    th = np.asarray(theta, dtype="<f4").reshape(-1, order="F").tobytes()
    if layout == "serial":
        data = struct.pack("<i", n) + th + struct.pack("<d", float(seed)) + struct.pack("<i", n)
    elif layout == "mpi":
        data = th + b"\0\0\0\0" + struct.pack("<d", float(seed)) + struct.pack("<i", n)
    else:
        raise ValueError(layout)
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        fh.write(data)
    os.replace(tmp, path)


def read_midout(path):
    """midout (dwf3d_lib.F90:141-144): dt beta am3 am imass iterl iter2 walltime."""
    with open(path) as fh:
        v = fh.readline().split()
    keys = ("dt", "beta", "am3", "am", "imass", "iterl", "iter2", "walltime")
    conv = (float, float, float, float, int, int, int, float)
    return {k: c(x) for k, c, x in zip(keys, conv, v)}


def read_remez(path):
    """read_remez_file (remez_common_subroutines.F90:20-35): two blocks of
    a0 then (a_i, sigma_i) lines; returns [(anum[0..n], aden[1..n]), (bnum, bden)]."""
    rows = [l.replace(",", " ").split() for l in open(path) if l.strip()]
    n = (len(rows) - 2) // 2
    if len(rows) != 2 * (n + 1):
        raise ValueError(f"{path}: malformed remez file")
    out = []
    for b in range(2):
        blk = rows[b * (n + 1):(b + 1) * (n + 1)]
        num = np.array([float(blk[0][0])] + [float(r[0]) for r in blk[1:]])
        den = np.array([float(r[1]) for r in blk[1:]])
        out.append((num, den))
    return out


DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def remez(name):
    return read_remez(os.path.join(DATA, name))
