"""VTK output of nodal fields: structure of the .vtu file, exact resampling
of polynomials of degree <= N to equispaced points, and ascii / base64
binary agreement.  CPU only."""
import base64
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import bfam_amd as bx
import bfam_amd.workloads as W


def read_array(da):
    fmt = da.get("format")
    typ = da.get("type")
    if fmt == "ascii":
        return np.array(da.text.split(), dtype=float)
    raw = base64.b64decode("".join(da.text.split()))
    n = struct.unpack("<Q", raw[:8])[0]
    body = raw[8:8 + n]
    dt = {"Float64": "<f8", "Int64": "<i8", "UInt8": "u1"}[typ]
    return np.frombuffer(body, dtype=dt).astype(float)


def arrays(path):
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    out = {da.get("Name"): read_array(da) for da in piece.iter("DataArray")}
    return piece, out


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("Np_write", [0, 4])
def test_vtk_roundtrip(tmp_path, dim, Np_write):
    N = 3
    mesh = W.Mesh(dim, [3, 2, 2][:dim], N, L=[1.0, 0.6, 0.9][:dim])
    X = mesh.coords()
    f = X[0] + 2 * X[1] ** 2 + (X[2] ** 3 if dim == 3 else 0)  # degree <= N
    res = {}
    for binary in (0, 1):
        p = tmp_path / f"t{binary}.vtu"
        bx.vtk_write(p, dim, N, mesh.x0, mesh.h, mesh.r, {"f": f, "g": 2 * f}, time=1.5,
                     binary=binary, Np_write=Np_write)
        piece, a = arrays(p)
        nv = Np_write or N + 1
        npts = mesh.K * nv ** dim
        ncells = mesh.K * (nv - 1) ** dim
        assert int(piece.get("NumberOfPoints")) == npts
        assert int(piece.get("NumberOfCells")) == ncells
        xyz = a["Position"].reshape(-1, 3)
        want = xyz[:, 0] + 2 * xyz[:, 1] ** 2 + (xyz[:, 2] ** 3 if dim == 3 else 0)
        assert np.abs(a["f"] - want).max() < 1e-12
        assert np.array_equal(a["g"], 2 * a["f"])
        conn = a["connectivity"]
        assert conn.min() == 0 and conn.max() == npts - 1
        assert (a["types"] == (9 if dim == 2 else 12)).all()
        assert (a["time"] == 1.5).all()
        # every cell is an axis-aligned box with positive extent
        nvert = 4 if dim == 2 else 8
        c = conn.reshape(-1, nvert).astype(int)
        ext = xyz[c[:, 2], :dim] - xyz[c[:, 0], :dim]
        assert (ext[:, 0] > 0).all() and (ext[:, 1] > 0).all()
        res[binary] = a
    for k in res[0]:
        assert np.allclose(res[0][k], res[1][k], rtol=1e-15, atol=0)
