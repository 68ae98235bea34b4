"""Map files of RigidBody::storeDistanceMapToFile / buildDistanceMapFromFile
(rigidbody.cpp:1214-1376), parsed on the host (no device needed).

The reference writes with the default stream formatting (six significant
digits) and reads with stream extraction into binary64, skipping the rest
of each record's last line.  The parser must give the correctly rounded
binary64 of every token -- checked against Python's own correctly rounded
float() of the same text -- and follow the stream's record semantics.
"""
from __future__ import annotations

import numpy as np
import pytest

from fullc0ntact_amd import _lib
from fullc0ntact_amd.distance_map import read_map_file, write_map_file


def _lattice(cells, cs, bmin, ystep=1.0):
    cx, cy, cz = cells
    z, y, x = np.meshgrid(np.arange(cz + 1), np.arange(cy + 1), np.arange(cx + 1), indexing="ij")
    v = np.stack([bmin[0] + x * cs, bmin[1] + y * cs * ystep, bmin[2] + z * cs], -1)
    return v.reshape(-1, 3)


def _tokens(path):
    return open(path).read().split()


def test_reference_format_parses_to_correctly_rounded_doubles(tmp_path):
    rng = np.random.default_rng(5)
    cells, cs = (3, 2, 4), 0.1234567
    v = _lattice(cells, cs, (-0.7, -0.31, 0.05), ystep=4.0)  # the 4*cellSize y quirk
    V = len(v)
    cps = v + rng.normal(0, 1e-3, (V, 3))
    nrm = rng.normal(0, 1, (V, 3))
    dist = rng.normal(0, 0.2, V)
    dist[:3] = [1e-7, -0.0, 12345.678]
    fbm = rng.integers(0, 2, V)
    path = tmp_path / "body.dmap"
    write_map_file(path, cells, cs, v, cps, nrm, dist, fbm)  # six digits, as the reference
    m = read_map_file(path)
    tok = _tokens(path)
    assert m["cells"] == cells and m["dim"] == (4, 3 * 5)
    assert m["cell_size"] == float(tok[5])
    body = tok[6:]
    for k, name in enumerate(("vertex_coords", "contact_points", "normals")):
        ref = np.array([float(t) for t in body[3 * V * k:3 * V * (k + 1)]]).reshape(V, 3)
        assert np.array_equal(m[name], ref), name
    ref_d = np.array([float(t) for t in body[9 * V:10 * V]])
    assert np.array_equal(m["distance"], ref_d)
    assert np.array_equal(m["fbm"], fbm)


def test_record_semantics_of_stream_extraction(tmp_path):
    # numbers of one record may span lines; the rest of its last line is
    # skipped; leading '+' and exponents are accepted
    text = ("2 1 comment after dims\n"
            "1 1\n1\n"
            "+0.5 extra words\n"
            "0 0 0\n1 0 0\n0 1 0\n1 1 0\n0 0 1\n1 0 1\n0 1 1\n1 1 1   # coords\n"
            + "0 0 0\n" * 8 + "1 0 0\n" * 8
            + "1e-1\n-2.5E+00 ignored\n3\n4\n5\n6\n7\n8\n"
            + "1\n0\n1\n0\n1\n0\n1\n0\n")
    path = tmp_path / "tiny.dmap"
    path.write_text(text)
    m = read_map_file(path)
    assert m["cells"] == (1, 1, 1) and m["cell_size"] == 0.5 and m["dim"] == (2, 1)
    assert m["vertex_coords"][3].tolist() == [1.0, 1.0, 0.0]
    assert m["distance"].tolist() == [0.1, -2.5, 3, 4, 5, 6, 7, 8]
    assert m["fbm"].tolist() == [1, 0] * 4


def test_header_only_and_errors(tmp_path):
    import ctypes as C
    cells, cs = (2, 2, 2), 0.5
    v = _lattice(cells, cs, (0, 0, 0))
    path = tmp_path / "m.dmap"
    write_map_file(path, cells, cs, v, v, v, np.zeros(len(v)))
    c3, d = (C.c_int32 * 3)(), C.c_double()
    _lib.check(_lib.lib.fc_distmap_file_read(str(path).encode(), c3, C.byref(d), None, None,
                                             None, None, None, None))
    assert tuple(c3) == cells and d.value == 0.5
    trunc = tmp_path / "t.dmap"
    trunc.write_text("\n".join(path.read_text().split("\n")[:40]))
    with pytest.raises(_lib.FcError):
        read_map_file(trunc)
    with pytest.raises(_lib.FcError):
        read_map_file(tmp_path / "missing.dmap")
    bad = tmp_path / "b.dmap"
    bad.write_text("3 9\n2 2 x\n0.5\n")
    with pytest.raises(_lib.FcError):
        read_map_file(bad)
