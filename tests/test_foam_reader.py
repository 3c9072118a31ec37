"""OpenFOAM polyMesh reader and the face-based mesh builder (libdopamine_hip.so: dm_foam_read,
dm_mesh_build_poly; dopamine_amd.mesh.read_foam / write_foam).

Reference: src/mesh/mod_foam_reader.f90:16-97 (read, topology checks, boundary patches),
:139-415 (ASCII parsers), :617-1104 (binary forms: compact and plain face lists, label and scalar
widths, byte order), :1144-1167 (patch type -> bc id); src/mesh/mod_mesh.f90:1336-1461 (polyhedral
geometry). The reference holds no polyMesh fixture (its benchmark writes one at run time,
tests/bench/mk_foam_box.py: binary, label=32, scalar=64, compact faces, 'patch' types); these
tests pin the reader with a hand-written file of known geometry, with closed-form invariants
(divergence theorem, total volume and first moment), with the element-list builder on meshes
whose faces are planar (where both geometry routes are exact up to rounding), and with every
binary variant against the ASCII file of the same mesh.
"""
import os

import numpy as np
import pytest

from dopamine_amd import mesh as M

HDR = """/*--------------------------------*- C++ -*----------------------------------*\\
  a comment block as OpenFOAM writes it
\\*---------------------------------------------------------------------------*/
FoamFile
{{
    version     2.0;
    format      ascii;
    class       {cls};
    location    "constant/polyMesh";
    object      {obj};
}}
// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //

"""


def _box_files(d):
    """One 2x3x5 hex as a hand-written ASCII polyMesh, outward right-hand faces."""
    os.makedirs(d, exist_ok=True)
    pts = ["(0 0 0)", "(2 0 0)", "(2 3 0)", "(0 3 0)", "(0 0 5)", "(2 0 5)", "(2 3 5)", "(0 3 5)"]
    faces = ["4(0 4 7 3)", "4(1 2 6 5)", "4(0 1 5 4)", "4(3 7 6 2)", "4(0 3 2 1)", "4(4 5 6 7)"]
    files = {
        "points": ("vectorField", "8\n(\n" + "\n".join(pts) + "\n)\n"),
        "faces": ("faceList", "6\n(\n" + "\n".join(faces) + "\n)\n"),
        "owner": ("labelList", "// note: nPoints:8  nCells:1  nFaces:6  nInternalFaces:0\n6\n(\n0\n0\n0\n0\n0\n0\n)\n"),
        "neighbour": ("labelList", "0\n(\n)\n"),
        "boundary": ("polyBoundaryMesh", """6
(
    x0
    {
        type            wall;
        inGroups        List<word> 1(wall);
        nFaces          1;
        startFace       0;
    }
    xL { type wall; nFaces 1; startFace 1; }
    y0
    {
        type            symmetryPlane;
        inGroups        List<word> 1(symmetryPlane);
        nFaces          1;
        startFace       2;
    }
    yL
    {
        type            patch;
        nFaces          1;
        startFace       3;
    }
    z0
    {
        type            empty;
        nFaces          1;
        startFace       4;
    }
    zL
    {
        type            empty;
        nFaces          1;
        startFace       5;
    }
)
"""),
    }
    for name, (cls, body) in files.items():
        with open(os.path.join(d, name), "w") as f:
            f.write(HDR.format(cls=cls, obj=name) + body)
    return d


def test_hand_written_box_known_answers(tmp_path):
    m = M.read_foam(_box_files(str(tmp_path / "pm")), warn=None)
    assert (m.n_cells, m.n_faces, m.n_int_faces, m.n_patches) == (1, 6, 0, 6)
    assert m.cell_volume[0] == pytest.approx(30.0, rel=1e-15)
    np.testing.assert_allclose(m.cell_centroid[0], [1.0, 1.5, 2.5], rtol=1e-15)
    np.testing.assert_allclose(m.face_area, [15, 15, 10, 10, 6, 6], rtol=1e-15)
    np.testing.assert_allclose(m.face_normal, [[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]],
                               atol=1e-15)
    np.testing.assert_allclose(m.face_centre[1], [2.0, 1.5, 2.5], rtol=1e-15)
    assert m.patch_names == ["x0", "xL", "y0", "yL", "z0", "zL"]
    # bc_type_to_id: wall -> wallModel, symmetryPlane -> symmetry, patch -> zeroGradient, empty -> slip
    np.testing.assert_array_equal(m.patch_bc, [M.BC_WALL_MODEL] * 2 + [M.BC_SYMMETRY, M.BC_ZEROGRADIENT]
                                  + [M.BC_SLIP] * 2)
    # the first-appearance vertex order over the cell's faces
    np.testing.assert_array_equal(m.cell_vertices, [0, 4, 7, 3, 1, 2, 6, 5])


def test_patch_types_override_the_mesh_types(tmp_path):
    d = _box_files(str(tmp_path / "pm"))
    seen = []
    e = M.read_foam_input(d, {"x0": "inlet", "xL": "outlet"}, warn=seen.append)
    np.testing.assert_array_equal(e["patch_bc"], [M.BC_INLET, M.BC_OUTLET, M.BC_SYMMETRY, M.BC_ZEROGRADIENT,
                                                  M.BC_SLIP, M.BC_SLIP])
    assert len(seen) == 4
    assert e["patch_types"] == ["wall", "wall", "symmetryPlane", "patch", "empty", "empty"]
    with pytest.raises(ValueError, match="not in the mesh: inflow"):
        M.read_foam_input(d, {"inflow": "inlet"}, warn=None)


def _same_connectivity(a, b):
    for k in ("face_owner", "face_neighbour", "face_patch", "cell_face_offset", "cell_faces", "patch_bc",
              "patch_start", "patch_nfaces", "periodic_pair", "ls_offset", "face_node_offset", "face_nodes"):
        np.testing.assert_array_equal(getattr(a, k), getattr(b, k), err_msg=k)
    for c in _uncapped(a):
        s = slice(a.ls_offset[c], a.ls_offset[c + 1])
        assert sorted(a.ls_cell[s]) == sorted(b.ls_cell[s]), c


def _uncapped(m):
    """Cells whose LS stencil is below the cap max(32, 4 max_cell_faces) (mod_mesh.f90:513-560):
    a capped stencil keeps the first vertex neighbours in the cell's vertex order, which the two
    input routes list differently (element node order vs first appearance over the faces)."""
    cap = max(32, 4 * m.max_cell_faces)
    return [c for c in range(m.n_cells) if m.ls_offset[c + 1] - m.ls_offset[c] < cap]


@pytest.mark.parametrize("maker", ["channel", "tets"])
def test_planar_faces_match_the_element_builder(tmp_path, maker):
    """Tets and axis-aligned (stretched) hexes have planar faces: vertex-mean and area-weighted
    face centres coincide, and the 5-tet and face-fan volumes are both exact, so the two geometry
    routes agree to rounding and the connectivity is identical."""
    if maker == "channel":
        src = M.channel_box(6, 5, 4)
    else:
        src = M.tet_box(3, jitter=0.2, shuffle_seed=11)
    d = str(tmp_path / "pm")
    M.write_foam(src, d)
    kind = {M.BC_PERIODIC: "periodic", M.BC_WALL_MODEL: "wallModel", M.BC_WALL_NOSLIP: "noSlip"}
    m = M.read_foam(d, dict(zip(src.patch_names, [kind[int(b)] for b in src.patch_bc])), warn=None)
    _same_connectivity(m, src)
    scale = np.abs(src.vertex_coords).max()
    for k, tol in (("cell_volume", 1e-13), ("cell_centroid", 1e-14), ("face_area", 1e-14), ("face_centre", 1e-14),
                   ("face_normal", 1e-13), ("face_dcf", 1e-14), ("face_gc", 1e-12)):
        a, b = getattr(m, k), getattr(src, k)
        ref = np.abs(b).max() if k not in ("face_centre", "cell_centroid") else scale
        assert np.abs(a - b).max() <= tol * max(ref, 1.0), k
    # LS weights per (cell, stencil member): the vertex neighbours may come in another order
    wmax = np.abs(src.ls_weight).max()
    for c in _uncapped(m):
        wa, wb = _ls_map(m, c), _ls_map(src, c)
        assert wa.keys() == wb.keys()
        for key in wa:
            assert np.abs(wa[key] - wb[key]).max() <= 1e-12 * wmax, (c, key)


def _ls_map(m, c):
    out = {}
    for k in range(m.ls_offset[c], m.ls_offset[c + 1]):
        key = ("cell", int(m.ls_cell[k])) if m.ls_cell[k] >= 0 else ("face", int(m.ls_face[k]))
        assert key not in out
        out[key] = m.ls_weight[k]
    return out


def test_polyhedral_geometry_invariants_on_curved_hexes(tmp_path):
    """Non-planar faces: the face-fan geometry is exact for the triangulated polyhedra it
    describes, so each cell closes (sum of outward area vectors 0), the volumes add up to the
    unit box (its boundary stays planar) and the first moments to the box centroid."""
    src = M.skewed_box(5, 0.06, group_types=("noSlip",) * 6)
    d = str(tmp_path / "pm")
    M.write_foam(src, d)
    m = M.read_foam(d, {n: "noSlip" for n in src.patch_names}, warn=None)
    s = np.zeros((m.n_cells, 3))
    np.add.at(s, m.face_owner, m.face_sf)
    i = m.face_neighbour >= 0
    np.add.at(s, m.face_neighbour[i], -m.face_sf[i])
    assert np.abs(s).max() < 1e-15
    assert m.cell_volume.sum() == pytest.approx(1.0, abs=1e-14)
    np.testing.assert_allclose((m.cell_volume[:, None] * m.cell_centroid).sum(axis=0), [0.5] * 3, atol=1e-14)
    # same cells and faces as the element builder (whose 5-tet volumes split warped faces along
    # other diagonals, so the geometry differs by the face warp, not by rounding)
    _same_connectivity(m, src)


@pytest.mark.parametrize("variant", [
    dict(binary=True),
    dict(binary=True, compact=False),
    dict(binary=True, label_bytes=8),
    dict(binary=True, msb=True),
    dict(binary=True, msb=True, label_bytes=8, compact=False),
])
def test_binary_variants_read_as_the_ascii_file(tmp_path, variant):
    src = M.tet_box(2, jitter=0.25, shuffle_seed=5)
    M.write_foam(src, str(tmp_path / "a"))
    M.write_foam(src, str(tmp_path / "b"), **variant)
    a = M.read_foam_input(str(tmp_path / "a"), warn=None)
    b = M.read_foam_input(str(tmp_path / "b"), warn=None)
    for k in ("points", "face_offset", "face_nodes", "owner", "neighbour", "patch_start", "patch_nfaces", "patch_bc"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    assert a["patch_names"] == b["patch_names"]


def test_single_precision_points(tmp_path):
    src = M.tet_box(2, jitter=0.25)
    M.write_foam(src, str(tmp_path / "b"), binary=True, scalar_bytes=4)
    b = M.read_foam_input(str(tmp_path / "b"), warn=None)
    np.testing.assert_array_equal(b["points"], src.vertex_coords.astype(np.float32).astype(np.float64))


def _bench_box_layout(nx, ny, nz, lx, ly, lz):
    """A uniform box in the layout of the reference's benchmark polyMesh: cell i + nx(j + ny k);
    internal faces per owner cell in the order +x, +y, +z; boundary patches x0, xL, y0, yL, z0, zL
    each ordered by owner cell; binary compact faces, label=32, type 'patch'."""
    def pid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)
    pts = np.array([[i * lx / nx, j * ly / ny, k * lz / nz] for k in range(nz + 1) for j in range(ny + 1)
                    for i in range(nx + 1)])
    faces, owner, nbr = [], [], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = i + nx * (j + ny * k)
                if i < nx - 1:
                    faces.append([pid(i + 1, j, k), pid(i + 1, j + 1, k), pid(i + 1, j + 1, k + 1), pid(i + 1, j, k + 1)])
                    owner.append(c), nbr.append(c + 1)
                if j < ny - 1:
                    faces.append([pid(i, j + 1, k), pid(i, j + 1, k + 1), pid(i + 1, j + 1, k + 1), pid(i + 1, j + 1, k)])
                    owner.append(c), nbr.append(c + nx)
                if k < nz - 1:
                    faces.append([pid(i, j, k + 1), pid(i + 1, j, k + 1), pid(i + 1, j + 1, k + 1), pid(i, j + 1, k + 1)])
                    owner.append(c), nbr.append(c + nx * ny)
    cells = [(i, j, k) for k in range(nz) for j in range(ny) for i in range(nx)]
    sides = [
        ("x0", lambda i, j, k: i == 0, lambda i, j, k: [pid(i, j, k), pid(i, j, k + 1), pid(i, j + 1, k + 1), pid(i, j + 1, k)]),
        ("xL", lambda i, j, k: i == nx - 1,
         lambda i, j, k: [pid(i + 1, j, k), pid(i + 1, j + 1, k), pid(i + 1, j + 1, k + 1), pid(i + 1, j, k + 1)]),
        ("y0", lambda i, j, k: j == 0, lambda i, j, k: [pid(i, j, k), pid(i + 1, j, k), pid(i + 1, j, k + 1), pid(i, j, k + 1)]),
        ("yL", lambda i, j, k: j == ny - 1,
         lambda i, j, k: [pid(i, j + 1, k), pid(i, j + 1, k + 1), pid(i + 1, j + 1, k + 1), pid(i + 1, j + 1, k)]),
        ("z0", lambda i, j, k: k == 0, lambda i, j, k: [pid(i, j, k), pid(i, j + 1, k), pid(i + 1, j + 1, k), pid(i + 1, j, k)]),
        ("zL", lambda i, j, k: k == nz - 1,
         lambda i, j, k: [pid(i, j, k + 1), pid(i + 1, j, k + 1), pid(i + 1, j + 1, k + 1), pid(i, j + 1, k + 1)]),
    ]
    nint = len(faces)
    starts, counts = [], []
    for name, sel, quad in sides:
        starts.append(len(faces))
        for c, (i, j, k) in enumerate(cells):
            if sel(i, j, k):
                faces.append(quad(i, j, k))
                owner.append(c)
        counts.append(len(faces) - starts[-1])
    f = np.array(faces, dtype=np.int32)
    return pts, f, np.array(owner, dtype=np.int32), np.array(nbr, dtype=np.int32), nint, starts, counts


def _write_bench_layout(d, pts, f, owner, nbr, starts, counts):
    os.makedirs(d, exist_ok=True)
    hdr = ('FoamFile\n{{\n    version     2.0;\n    format      {fmt};\n    arch        "LSB;label=32;scalar=64";\n'
           '    class       {cls};\n    location    "constant/polyMesh";\n    object      {obj};\n}}\n\n')

    def lists(name, cls, arrays):
        with open(os.path.join(d, name), "wb") as fh:
            fh.write(hdr.format(fmt="binary", cls=cls, obj=name).encode())
            for a in arrays:
                fh.write(f"{a.shape[0]}\n(".encode())
                fh.write(np.ascontiguousarray(a).tobytes())
                fh.write(b")\n")
    lists("points", "vectorField", [pts.astype("<f8")])
    lists("faces", "faceCompactList", [(np.arange(len(f) + 1) * 4).astype("<i4"), f.reshape(-1).astype("<i4")])
    lists("owner", "labelList", [owner.astype("<i4")])
    lists("neighbour", "labelList", [nbr.astype("<i4")])
    with open(os.path.join(d, "boundary"), "w") as fh:
        fh.write(hdr.format(fmt="ascii", cls="polyBoundaryMesh", obj="boundary") + f"{len(starts)}\n(\n")
        for name, s, n in zip(("x0", "xL", "y0", "yL", "z0", "zL"), starts, counts):
            fh.write(f"    {name}\n    {{\n        type            patch;\n        nFaces          {n};\n"
                     f"        startFace       {s};\n    }}\n")
        fh.write(")\n")


def test_benchmark_layout_box_is_the_uniform_box(tmp_path):
    nx, ny, nz, L = 4, 3, 5, 2.0
    pts, f, owner, nbr, nint, starts, counts = _bench_box_layout(nx, ny, nz, L, L, L)
    d = str(tmp_path / "pm")
    _write_bench_layout(d, pts, f, owner, nbr, starts, counts)
    m = M.read_foam(d, {n: "periodic" for n in ("x0", "xL", "y0", "yL", "z0", "zL")}, warn=None)
    ref = M.uniform_box(nx, ny, nz, L, L, L)
    assert (m.n_cells, m.n_faces, m.n_int_faces) == (ref.n_cells, ref.n_faces, ref.n_int_faces)
    # same cell numbering; dyadic-free spacing still gives every cell the same volume to rounding
    np.testing.assert_allclose(m.cell_volume, ref.cell_volume, rtol=1e-14)
    np.testing.assert_allclose(m.cell_centroid, ref.cell_centroid, atol=1e-14)
    assert (m.periodic_pair >= 0).all()
    # periodic partners are translations by one box length
    b = np.arange(m.n_bnd_faces)
    d_c = m.face_centre[m.n_int_faces + m.periodic_pair[b]] - m.face_centre[m.n_int_faces + b]
    assert np.allclose(np.sort(np.abs(d_c), axis=1)[:, :2], 0.0, atol=1e-14)
    assert np.allclose(np.abs(d_c).max(axis=1), L)
    # the LS stencil reaches the same neighbour cells in both numberings
    for c in range(m.n_cells):
        s1 = slice(m.ls_offset[c], m.ls_offset[c + 1])
        s2 = slice(ref.ls_offset[c], ref.ls_offset[c + 1])
        assert sorted(m.ls_cell[s1]) == sorted(ref.ls_cell[s2])


def _corrupt(tmp_path, name, old, new):
    d = _box_files(str(tmp_path / "pm"))
    p = os.path.join(d, name)
    text = open(p).read()
    assert old in text
    with open(p, "w") as f:
        f.write(text.replace(old, new, 1))
    return d


@pytest.mark.parametrize("name,old,new,msg", [
    ("owner", "\n0\n0\n0\n0\n0\n0\n)", "\n0\n-1\n0\n0\n0\n0\n)", "negative cell index"),
    ("boundary", "startFace       5;", "startFace       6;", "outside the boundary faces"),
    ("boundary", "nFaces          1;\n        startFace       5;", "nFaces          0;\n        startFace       5;",
     "not mapped to a patch"),
    ("boundary", "type            symmetryPlane;", "type            cyclic;", "odd number of periodic"),
    ("faces", "4(4 5 6 7)", "4(4 5 6 9)", "point label out of range"),
    ("points", "FoamFile", "FoamFiel", "no FoamFile header"),
    ("faces", "format      ascii;", "format      hex;", "unknown format"),
])
def test_corrupt_files_are_rejected(tmp_path, name, old, new, msg):
    with pytest.raises(ValueError, match=msg):
        M.read_foam_input(_corrupt(tmp_path, name, old, new), warn=None)


def test_owner_not_below_neighbour_is_rejected(tmp_path):
    src = M.uniform_box(2, 1, 1, 1, 1, 1, group_types=("noSlip",) * 6)
    d = str(tmp_path / "pm")
    M.write_foam(src, d)
    p = os.path.join(d, "neighbour")
    text = open(p).read()
    with open(p, "w") as f:
        f.write(text.replace("(\n1\n)", "(\n0\n)"))
    with pytest.raises(ValueError, match="owner < neighbour"):
        M.read_foam_input(d, warn=None)


def test_missing_directory_is_an_error(tmp_path):
    with pytest.raises(ValueError, match="cannot open"):
        M.read_foam_input(str(tmp_path / "none"))


def _l_prism(start=0, h=0.5):
    """One polyhedral cell: the L-shaped hexagon of the reference's mesh-geometry test extruded by
    h; the bottom face's point list starts at vertex `start`."""
    base = [(0, 0), (2, 0), (2, 1), (1, 1), (1, 2), (0, 2)]
    pts = np.array([(x, y, 0.0) for x, y in base] + [(x, y, h) for x, y in base], dtype=float)
    bottom = [(start - k) % 6 for k in range(6)]          # clockwise from below: normal -z
    top = [6 + k for k in range(6)]                        # normal +z
    sides = [[k, (k + 1) % 6, 6 + (k + 1) % 6, 6 + k] for k in range(6)]
    faces = [bottom, top] + sides
    off = np.concatenate([[0], np.cumsum([len(f) for f in faces])])
    return {"points": pts, "face_offset": off, "face_nodes": np.concatenate(faces), "owner": np.zeros(len(faces), int),
            "neighbour": np.zeros(0, int), "patch_start": np.array([0]), "patch_nfaces": np.array([len(faces)]),
            "patch_bc": np.array([M.BC_WALL_NOSLIP]), "patch_names": ["walls"]}


@pytest.mark.parametrize("start", range(6))
def test_concave_face_geometry_known_answers(start):
    """The reference's mesh-geometry test (tests/unit/test_mesh_geometry.f90:22-41): the L-shaped
    face has area 3 and centroid (5/6, 5/6) whatever vertex the fan starts from; the extruded cell
    closes with volume 3 h and centroid (5/6, 5/6, h/2)."""
    h = 0.5
    m = M.Mesh(_l_prism(start, h))
    assert m.face_area[0] == pytest.approx(3.0, abs=1e-12)
    np.testing.assert_allclose(m.face_centre[0], [5 / 6, 5 / 6, 0.0], atol=1e-12)
    np.testing.assert_allclose(m.face_normal[0], [0.0, 0.0, -1.0], atol=1e-12)
    assert m.cell_volume[0] == pytest.approx(3.0 * h, rel=1e-12)
    np.testing.assert_allclose(m.cell_centroid[0], [5 / 6, 5 / 6, h / 2], atol=1e-12)
