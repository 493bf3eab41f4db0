"""The host writers of the Fortran drop-in against a reader built from the reference's WRITE
statements (vpm.f90:662-833): record sequence, record lengths and every value.

tests/fortran/writers_probe.f90 fills the public grid state (pmgrid) with known values and
calls writeline twice (NTIME 1 and 2), writesol and writesolXavatar for both slice sets; no
GPU call is made.  The files are Fortran sequential unformatted: each record is a 4-byte
length, the payload, the same length again.
  NNhist.bin (writeline): at NTIME = 1 the file is rewound and gets XMIN/YMIN/ZMIN (3 doubles),
    DX/DY/DZ (3 doubles), NXs,NXf,j,k (4 int32); every call appends NTIME (int32) followed by
    velvrx, velvry, velvrz on nodes NXs..NXf of line (j, k) (doubles);
    j = int(f*NNJ), k = int(g*NNK) from the reference's jmat/kmat table.
  NNNNNsolution.bin (writesol): the VARIABLES string, the ZONE record ('ZONE I=', int, ' J=',
    int, ' K=', int, ' F=POINT'), then one record per node k->j->i of the NN_bl window:
    X, Y, Z, U, V, W, -RHS(1:3).
  NNNNNsolX.bin (writesolXavatar): 10 slices at NX_AVA (2, then int((POSX - off - XMIN)/DX) + 1
    for the reference's POSX tables), each a record (NJ, NK) then per node j-fastest the same
    9 doubles.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "vpm_free_amd", "lib")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
NX, NY, NZ = 110, 6, 5
XMIN, YMIN, ZMIN = -360.0, -20.0, -16.0
DX, DY, DZ = 8.0, 7.0, 6.5
NXS, NXF, NYS, NYF, NZS, NZF = 2, NX - 1, 2, NY - 1, 1, NZ


def records(path):
    out = []
    with open(path, "rb") as f:
        data = f.read()
    p = 0
    while p < len(data):
        (n,) = struct.unpack_from("<i", data, p)
        payload = data[p + 4:p + 4 + n]
        (n2,) = struct.unpack_from("<i", data, p + 4 + n)
        assert n2 == n, "record markers disagree"
        out.append(payload)
        p += 8 + n
    return out


def field(name, i, j, k):
    """The probe's values (1-based indices)."""
    return {
        "u": 1.0 * i + 1e-2 * j + 1e-4 * k,
        "v": 2.0 * i + 2e-2 * j + 2e-4 * k,
        "w": 3.0 * i + 3e-2 * j + 3e-4 * k,
        "r1": 5.0 * i - 1e-2 * j,
        "r2": 6.0 * i - 1e-2 * k,
        "r3": 7.0 * i + 1e-3 * j * k,
    }[name]


def point(i, j, k):
    return [XMIN + (i - 1) * DX, YMIN + (j - 1) * DY, ZMIN + (k - 1) * DZ, field("u", i, j, k),
            field("v", i, j, k), field("w", i, j, k), -field("r1", i, j, k), -field("r2", i, j, k),
            -field("r3", i, j, k)]


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    if not os.path.exists(FLANG) or not os.path.exists(os.path.join(LIB, "libvpm_lib.so")):
        pytest.skip("flang or the Fortran drop-in is not built")
    d = tmp_path_factory.mktemp("writers")
    exe = str(d / "writers_probe")
    src = os.path.join(ROOT, "tests", "fortran", "writers_probe.f90")
    subprocess.run([FLANG, "-O0", "-I", os.path.join(LIB, "mod"), "-o", exe, src, "-L" + LIB, "-lvpm_lib", "-lvpmx",
                    "-Wl,-rpath," + LIB], check=True, capture_output=True)
    r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return d


def test_writeline_records(written):
    nnj, nnk = NYF - NYS + 1, NZF - NZS + 1
    # jmat/kmat of vpm.f90:808-816: real products assigned to integers (truncation)
    fr = [(0.5, 0.5), (0.25, 0.25), (0.25, 0.75), (0.75, 0.25), (0.75, 0.75), (0.5, 0.25), (0.5, 0.75),
          (0.25, 0.5), (0.75, 0.5)]
    for line in range(1, 10):
        rec = records(written / ("%02dhist.bin" % line))
        assert len(rec) == 3 + 2, line  # header at NTIME 1, one record per call
        assert struct.unpack("<3d", rec[0]) == (XMIN, YMIN, ZMIN)
        assert struct.unpack("<3d", rec[1]) == (DX, DY, DZ)
        j, k = int(fr[line - 1][0] * nnj), int(fr[line - 1][1] * nnk)
        assert struct.unpack("<4i", rec[2]) == (NXS, NXF, j, k)
        n = NXF - NXS + 1
        for t, r in zip((1, 2), rec[3:]):
            assert len(r) == 4 + 3 * 8 * n
            (nt,) = struct.unpack_from("<i", r, 0)
            vals = np.frombuffer(r[4:], dtype="<f8")
            assert nt == t
            want = [field(c, i, j, k) for c in ("u", "v", "w") for i in range(NXS, NXF + 1)]
            assert np.array_equal(vals, np.array(want))


def test_writesol_records(written):
    rec = records(written / "00007solution.bin")
    assert rec[0] == b'VARIABLES = "X" "Y" "Z" "U" "V" "W" "VORTX" "VORTY" "VORTZ"'
    z = rec[1]
    assert len(z) == 7 + 4 + 3 + 4 + 3 + 4 + 8
    assert z[:7] == b"ZONE I=" and z[11:14] == b" J=" and z[18:21] == b" K=" and z[25:] == b" F=POINT"
    assert struct.unpack_from("<i", z, 7)[0] == NXF - NXS + 1
    assert struct.unpack_from("<i", z, 14)[0] == NYF - NYS + 1
    assert struct.unpack_from("<i", z, 21)[0] == NZF - NZS + 1
    pts = [point(i, j, k) for k in range(NZS, NZF + 1) for j in range(NYS, NYF + 1) for i in range(NXS, NXF + 1)]
    assert len(rec) == 2 + len(pts)
    for r, p in zip(rec[2:], pts):
        assert np.array_equal(np.frombuffer(r, dtype="<f8"), np.array(p))


@pytest.mark.parametrize("kind,ntime", [(1, 7), (0, 8)])
def test_writesolx_records(written, kind, ntime):
    if kind == 1:
        posx, off = [-350, -250, -150, -50, 50, 150, 250, 350, 450], 0.0
    else:
        posx, off = [-204, -102, -50, -4, 4, 50, 102, 204, 306], 7.1
    nxa = [2] + [int((p - off - XMIN) / DX) + 1 for p in posx]
    assert all(1 <= i <= NX for i in nxa)  # the probe's box holds every slice
    rec = records(written / ("%05dsolX.bin" % ntime))
    nj, nk = NYF - NYS + 1, NZF - NZS + 1
    assert len(rec) == 10 * (1 + nj * nk)
    q = 0
    for i in nxa:
        assert struct.unpack("<2i", rec[q]) == (nj, nk)
        q += 1
        for k in range(NZS, NZF + 1):
            for j in range(NYS, NYF + 1):
                assert np.array_equal(np.frombuffer(rec[q], dtype="<f8"), np.array(point(i, j, k))), (i, j, k)
                q += 1
