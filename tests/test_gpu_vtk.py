"""Legacy binary VTK files written from device meshes (write_vtk /
write_vtk_with_data, crates/tpt-fem-io-vtk/src/lib.rs:117-183): byte for byte
the reference's own files and the oracle's writer."""
import os

import numpy as np
import pytest

import oracle as O
from vtk_util import read_legacy_vtk

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vtk")


def tp():
    import tptfem

    return tptfem


def test_box_mesh_file_is_the_reference_example(ctx, tmp_path):
    """mesh_gen_box.rs: write_vtk(box_mesh([0,0,0], [1,1,1], [4,4,4])), the
    mesh generated and written on the GPU."""
    T = tp()
    m = T.Mesh.box(ctx, [0, 0, 0], [1, 1, 1], (4, 4, 4))
    out = tmp_path / "box.vtk"
    T.write_vtk(m, out)
    assert out.read_bytes() == open(os.path.join(GOLD, "mesh_gen_box_structured.vtk"), "rb").read()


@pytest.mark.parametrize("name", ["mesh_gen_box_delaunay.vtk", "modal_analysis.vtk", "elasticity_frame.vtk",
                                  "abaqus_import.vtk", "thermal_solve.vtk"])
def test_reference_files_round_trip(ctx, tmp_path, name):
    T = tp()
    d = open(os.path.join(GOLD, name), "rb").read()
    _, pts, conn, cell, fields = read_legacy_vtk(d)
    m = T.Mesh.from_arrays(ctx, cell, pts, conn)
    out = tmp_path / name
    T.write_vtk(m, out, fields)
    assert out.read_bytes() == d


def test_fields_and_cells_match_oracle(ctx, tmp_path):
    """2-D grids (points padded with z = 0), hexes and several fields."""
    T = tp()
    rng = np.random.default_rng(2)
    qc, qe = O.grid(6, 4, 0.3)
    hx = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    cases = [(O.QUAD, qc, qe), (O.HEX, hx, np.arange(8)[None, :]),
             (O.TRI, np.array([[0, 0], [1, 0], [0, 1.5]]), np.array([[0, 1, 2]]))]
    for cell, c, e in cases:
        m = T.Mesh.from_arrays(ctx, cell, c, e)
        fields = [("temperature", rng.standard_normal(len(c))), ("u_mag", np.abs(rng.standard_normal(len(c))))]
        out = tmp_path / f"c{cell}.vtk"
        T.write_vtk(m, out, fields)
        assert out.read_bytes() == O.vtk_legacy_bytes(c, cell, e, fields)
    with pytest.raises(T.TfmError, match="I/O error on VTK file"):
        T.write_vtk(m, tmp_path / "no_such_dir" / "x.vtk")
