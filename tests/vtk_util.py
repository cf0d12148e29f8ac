"""Reader of the legacy binary VTK files the reference's examples wrote (test
helper, shared by the oracle and GPU writer tests)."""
import numpy as np

import oracle as O

VTK_CELL = {3: O.LINE, 5: O.TRI, 9: O.QUAD, 10: O.TET, 12: O.HEX}


def read_legacy_vtk(d):
    """Minimal reader of the reference's legacy binary files: title, points,
    connectivity, cell type and scalar fields."""
    pos = 0

    def line():
        nonlocal pos
        j = d.index(b"\n", pos)
        s = d[pos:j].decode()
        pos = j + 1
        return s

    assert line() == "# vtk DataFile Version 4.1"
    title = line()
    assert line() == "BINARY" and line() == "" and line() == "DATASET UNSTRUCTURED_GRID"
    n = int(line().split()[1])
    pts = np.frombuffer(d[pos:pos + 24 * n], ">f8").reshape(n, 3).copy()
    pos += 24 * n
    assert line() == "" and line() == ""
    _, nc, sz = line().split()
    nc, sz = int(nc), int(sz)
    verts = np.frombuffer(d[pos:pos + 4 * sz], ">u4").astype(np.int64)
    pos += 4 * sz
    assert line() == "" and line() == "" and line() == f"CELL_TYPES {nc}"
    types = np.frombuffer(d[pos:pos + 4 * nc], ">i4")
    pos += 4 * nc
    assert line() == "" and line() == "" and line() == f"POINT_DATA {n}"
    fields = []
    while True:
        assert line() == ""
        s = line()
        if s.startswith("CELL_DATA"):
            break
        name = s.split()[1]
        assert line() == "LOOKUP_TABLE default"
        fields.append((name, np.frombuffer(d[pos:pos + 8 * n], ">f8").copy()))
        pos += 8 * n
        assert line() == ""
    assert line() == "" and pos == len(d)
    nn = int(verts[0])
    return title, pts, verts.reshape(nc, nn + 1)[:, 1:], VTK_CELL[int(types[0])], fields
