"""Gmsh v4 ASCII reader (libdopamine_hip.so, dm_gmsh_read; dopamine_amd.mesh.read_gmsh).

Reference: src/mesh/mod_gmsh_reader.f90:45-236 (sections, entity -> physical tag, node and
element order), :618-649 (patches = 2-D groups with faces, in group order), and
src/bc/mod_bc_setup.f90:42-106 (patch types by name; unlisted patch -> zeroGradient; input name
missing from the mesh -> error). The reference's own fixture tests/unit/mesh/gmsh_open_tet.msh
(copied to tests/golden/mesh/) is a single tet with one unlabelled face, which its mesh test
expects to be rejected.
"""
import os

import numpy as np
import pytest

from dopamine_amd import mesh as M

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "mesh")

# One unit hex in the layout tests/regression/mk_box_mesh.py writes (no physical group on the
# volume, surfaces 1..6 carrying physical tags 1..6).
UNIT_CUBE = """$MeshFormat
4.1 0 8
$EndMeshFormat
$PhysicalNames
6
2 1 "x0"
2 2 "xL"
2 3 "y0"
2 4 "yL"
2 5 "z0"
2 6 "zL"
$EndPhysicalNames
$Entities
0 0 6 1
1 0 0 0 1 1 1 1 1 0
2 0 0 0 1 1 1 1 2 0
3 0 0 0 1 1 1 1 3 0
4 0 0 0 1 1 1 1 4 0
5 0 0 0 1 1 1 1 5 0
6 0 0 0 1 1 1 1 6 0
1 0 0 0 1 1 1 0 0
$EndEntities
$Nodes
1 8 1 8
3 1 0 8
11
12
13
14
15
16
17
18
0 0 0
2 0 0
2 3 0
0 3 0
0 0 5
2 0 5
2 3 5
0 3 5
$EndNodes
$Elements
7 7 1 7
2 1 3 1
1 11 14 18 15
2 2 3 1
2 12 13 17 16
2 3 3 1
3 11 12 16 15
2 4 3 1
4 14 18 17 13
2 5 3 1
5 11 12 13 14
2 6 3 1
6 15 16 17 18
3 1 5 1
7 11 12 13 14 15 16 17 18
$EndElements
"""


def _write(tmp_path, text, name="m.msh"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_reference_open_tet_fixture_is_rejected_as_open():
    e = M.read_gmsh_elements(os.path.join(GOLDEN, "gmsh_open_tet.msh"), warn=None)
    # the file parses: 4 nodes, one tet, three labelled triangles in group "a"
    assert len(e["coords"]) == 12 and list(e["vol_type"]) == [M.GMSH_TET4]
    assert list(e["bnd_type"]) == [M.GMSH_TRI3] * 3 and list(e["bnd_group"]) == [0, 0, 0]
    assert e["group_names"] == ["a"]
    np.testing.assert_array_equal(e["vol_nodes"], [0, 2, 3, 1])
    with pytest.raises(RuntimeError, match="open cells"):
        M.Mesh(e)


def test_unit_cube_known_answers(tmp_path):
    types = {"x0": "noSlip", "xL": "noSlip", "y0": "slip", "yL": "slip", "z0": "wallModel", "zL": "wallModel"}
    m = M.read_gmsh(_write(tmp_path, UNIT_CUBE), types)
    assert m.n_cells == 1 and m.n_faces == 6 and m.n_int_faces == 0
    assert m.cell_volume[0] == pytest.approx(30.0, rel=1e-14)
    np.testing.assert_allclose(m.cell_centroid[0], [1.0, 1.5, 2.5], rtol=1e-14)
    assert m.patch_names == ["x0", "xL", "y0", "yL", "z0", "zL"]
    np.testing.assert_array_equal(m.patch_bc, [M.BC_WALL_NOSLIP] * 2 + [M.BC_SLIP] * 2 + [M.BC_WALL_MODEL] * 2)
    area = {"x0": 15.0, "xL": 15.0, "y0": 10.0, "yL": 10.0, "z0": 6.0, "zL": 6.0}
    for p, name in enumerate(m.patch_names):
        f = m.patch_start[p]
        assert m.patch_nfaces[p] == 1
        assert m.face_area[f] == pytest.approx(area[name], rel=1e-14)
    # node tags are not positions: the reader maps tags 11..18 to nodes 0..7
    np.testing.assert_array_equal(m.vertex_coords[6], [2.0, 3.0, 5.0])


@pytest.mark.parametrize("maker", ["skewed_hex", "shuffled_tets"])
def test_write_read_round_trip_gives_the_same_mesh(tmp_path, maker):
    if maker == "skewed_hex":
        e = M.skewed_box_elements(4, 0.04, group_types=("noSlip",) * 6)
    else:
        e = M.tet_box_elements(3, jitter=0.2, shuffle_seed=7)
    src = M.Mesh(e)
    path = str(tmp_path / "rt.msh")
    M.write_gmsh(e, path)
    back = M.read_gmsh_elements(path, {n: "noSlip" for n in e["group_names"]}, warn=None)
    for k in ("coords", "vol_type", "vol_offset", "vol_nodes", "bnd_type", "bnd_offset", "bnd_nodes", "bnd_group",
              "group_bc"):
        np.testing.assert_array_equal(back[k], e[k], err_msg=k)
    m = M.Mesh(back)
    for name in M.Mesh.ARRAYS:
        np.testing.assert_array_equal(getattr(m, name), getattr(src, name), err_msg=name)
    assert m.patch_names == src.patch_names


def test_patch_types_follow_the_reference_bc_matching(tmp_path):
    path = _write(tmp_path, UNIT_CUBE)
    seen = []
    e = M.read_gmsh_elements(path, {"x0": "inlet", "xL": "outlet"}, warn=seen.append)
    np.testing.assert_array_equal(e["group_bc"], [M.BC_INLET, M.BC_OUTLET] + [M.BC_ZEROGRADIENT] * 4)
    assert len(seen) == 4 and all("zeroGradient" in s for s in seen)
    with pytest.raises(ValueError, match="not in the mesh: inflow"):
        M.read_gmsh_elements(path, {"inflow": "inlet"}, warn=None)


def test_empty_physical_group_gets_no_patch(tmp_path):
    text = UNIT_CUBE.replace('6\n2 1 "x0"', '7\n2 1 "x0"').replace('2 6 "zL"\n', '2 6 "zL"\n2 9 "unused"\n')
    e = M.read_gmsh_elements(_write(tmp_path, text), warn=None)
    assert e["group_names"] == ["x0", "xL", "y0", "yL", "z0", "zL", "unused"]
    m = M.Mesh(e)
    assert m.n_patches == 6 and m.patch_names == ["x0", "xL", "y0", "yL", "z0", "zL"]


@pytest.mark.parametrize("edit,msg", [
    (("4.1 0 8", "2.2 0 8"), "v4"),
    (("4.1 0 8", "4.1 1 8"), "ASCII"),
    (("3 1 5 1\n", "3 1 12 1\n"), "unsupported Gmsh element type 12"),
    (("5 0 0 0 1 1 1 1 5 0", "5 0 0 0 1 1 1 0 0"), "no 2-D physical group"),
    (("7 11 12 13 14 15 16 17 18", "7 11 12 13 14 15 16 17 99"), "node tag"),
])
def test_malformed_files_are_rejected(tmp_path, edit, msg):
    text = UNIT_CUBE.replace(*edit)
    assert text != UNIT_CUBE
    with pytest.raises(ValueError, match=msg):
        M.read_gmsh_elements(_write(tmp_path, text), warn=None)


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(ValueError, match="cannot open"):
        M.read_gmsh_elements(str(tmp_path / "none.msh"))
