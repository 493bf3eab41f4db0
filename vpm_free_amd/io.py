"""On-disk formats on either side of the path.

particles.bin (test.f90:24-36): Fortran unformatted sequential records
    [NVR int32] [Vref real8] then NVR records [x y z wx wy wz] (6 x real8);
the application stores QP(1:3) = -w*Vref and QP(4) = Vref after reading.

NNhist.bin (writeline, vpm.f90:789-833): per mode-2 call one record
    [NTIME int32, u(NXs:NXf), v(NXs:NXf), w(NXs:NXf)] after a 3-record header at NTIME=1.

Each sequential record is framed by 4-byte little-endian length markers.
"""
import struct

import numpy as np


def _records(buf):
    off = 0
    while off < len(buf):
        (n,) = struct.unpack_from("<i", buf, off)
        yield buf[off + 4:off + 4 + n]
        off += 8 + n


def write_particles_bin(path, xp, w, vref):
    """Write the driver input: positions xp (N,3) and vorticity w (N,3)."""
    xp = np.asarray(xp, dtype="<f8")
    w = np.asarray(w, dtype="<f8")
    with open(path, "wb") as f:
        def rec(b):
            f.write(struct.pack("<i", len(b)))
            f.write(b)
            f.write(struct.pack("<i", len(b)))
        rec(struct.pack("<i", xp.shape[0]))
        rec(struct.pack("<d", float(vref)))
        rows = np.concatenate([xp, w], axis=1)
        for r in rows:
            rec(r.tobytes())


def read_particles_bin(path):
    """Returns (xp (N,3), w (N,3), vref)."""
    buf = open(path, "rb").read()
    recs = list(_records(buf))
    n = struct.unpack("<i", recs[0])[0]
    vref = struct.unpack("<d", recs[1])[0]
    rows = np.frombuffer(b"".join(recs[2:2 + n]), dtype="<f8").reshape(n, 6)
    return rows[:, :3].copy(), rows[:, 3:].copy(), vref


def particles_to_qp(xp, w, vref):
    """The application's conversion after reading (test.f90:36-39)."""
    qp = np.zeros((xp.shape[0], 4))
    qp[:, :3] = -w * vref
    qp[:, 3] = vref
    return qp


def read_hist(path):
    """writeline output: header (origin, spacing, NXs, NXf, j, k) and the per-call lines."""
    buf = open(path, "rb").read()
    recs = list(_records(buf))
    x0 = struct.unpack("<3d", recs[0])
    dx = struct.unpack("<3d", recs[1])
    nxs, nxf, j, k = struct.unpack("<4i", recs[2])
    n = nxf - nxs + 1
    lines = []
    for r in recs[3:]:
        nt = struct.unpack_from("<i", r, 0)[0]
        v = np.frombuffer(r[4:], dtype="<f8").reshape(3, n)
        lines.append((nt, v))
    return {"origin": x0, "spacing": dx, "nxs": nxs, "nxf": nxf, "j": j, "k": k, "lines": lines}


def write_pm_inp(path, dpm, interf_iproj=4, ibctyp=2, idvpm=1, epsvol=0.0, ncoarse=8, blocks=(1, 1, 1),
                 nremesh=1, ncell_rem=1, iyntree=0, ilevmax=1, ompthreads=1, idefine=1, iynslice=0,
                 ipmwrite=0, windows=()):
    """pm.inp in the positional layout vpm.f90:132-152 reads."""
    d = [float(v) for v in np.broadcast_to(np.asarray(dpm, dtype=float), (3,))]
    lines = [f"{d[0]!r} {d[1]!r} {d[2]!r}   ! Dpm", f"{interf_iproj}   ! projection fun",
             f"{ibctyp}   ! boundary cond type", f"{idvpm}   ! Variable/Constant Volume(0,1)",
             f"{epsvol}   ! EPSVOL", f"{ncoarse}   ! ncoarse", f"{blocks[0]} {blocks[1]} {blocks[2]}   ! blocks",
             f"{nremesh} {ncell_rem}   ! remesh", f"{iyntree} {ilevmax}   ! tree", f"{ompthreads}   ! threads",
             f"{idefine}   ! box fixed", f"{iynslice}   ! iynslice", f"{ipmwrite}   ! IPMWRITE"]
    for s, n in windows:
        lines.append(f"{s} {n}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
