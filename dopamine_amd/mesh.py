"""Mesh setup front end: element-list generators and the native mesh builder.

The reference reads Gmsh v4 / OpenFOAM meshes and builds its face list, geometry, periodic pairs
and least-squares stencil once at start-up (src/mesh/mod_gmsh_reader.f90:45-116,
src/mesh/mod_mesh.f90:215-758, src/bc/mod_bc_setup.f90:134-330). Here the same element lists come
from a Gmsh v4 ASCII file (read_gmsh: parsed by the native reader) or from generators (structured
hex boxes, with the node/element/boundary ordering that tests/regression/mk_box_mesh.py writes),
and are handed to the native builder in libdopamine_hip.so.
"""
import ctypes
import math

import numpy as np

from ._lib import lib

# Boundary-condition ids (src/mod_parameters.f90:43-54).
BC_INLET, BC_OUTLET, BC_WALL_NOSLIP, BC_WALL_MODEL, BC_SLIP, BC_PERIODIC, BC_SYMMETRY, \
    BC_FIXEDTEMP, BC_ZEROGRADIENT, BC_WALL_VELOCITY, BC_ABL_INLET = range(1, 12)

# patch_type names accepted by the reference input parser (src/io/mod_input_parser.f90:1203-1233).
BC_NAMES = {
    "velocity": BC_INLET, "inlet": BC_INLET, "fixedvalue": BC_INLET,
    "outlet": BC_OUTLET, "fixedpressure": BC_OUTLET, "outflow": BC_OUTLET,
    "zerogradient": BC_ZEROGRADIENT, "noslip": BC_WALL_NOSLIP,
    "movingwall": BC_WALL_VELOCITY, "wallvelocity": BC_WALL_VELOCITY,
    "wallmodel": BC_WALL_MODEL, "wall": BC_WALL_MODEL, "slip": BC_SLIP, "symmetry": BC_SYMMETRY,
    "periodic": BC_PERIODIC, "cyclic": BC_PERIODIC, "fixedtemperature": BC_FIXEDTEMP,
    "ablinlet": BC_ABL_INLET,
}

GMSH_QUAD4, GMSH_HEX8 = 3, 5

_P = ctypes.POINTER
_d, _i32, _i64 = ctypes.c_double, ctypes.c_int32, ctypes.c_int64


class _Input(ctypes.Structure):
    _fields_ = [("n_nodes", _i64), ("coords", _P(_d)),
                ("n_vol", _i64), ("vol_type", _P(_i32)), ("vol_offset", _P(_i64)), ("vol_nodes", _P(_i32)),
                ("n_bnd", _i64), ("bnd_type", _P(_i32)), ("bnd_offset", _P(_i64)), ("bnd_nodes", _P(_i32)),
                ("bnd_group", _P(_i32)), ("n_groups", _i32), ("group_bc", _P(_i32))]


class _View(ctypes.Structure):
    _fields_ = [("n_cells", _i32), ("n_faces", _i32), ("n_int_faces", _i32), ("n_bnd_faces", _i32),
                ("n_patches", _i32), ("n_vertices", _i32), ("n_ls", _i64), ("max_cell_faces", _i32),
                ("ls_degenerate", _i32),
                ("cell_centroid", _P(_d)), ("cell_volume", _P(_d)), ("cell_face_offset", _P(_i32)),
                ("cell_faces", _P(_i32)),
                ("face_owner", _P(_i32)), ("face_neighbour", _P(_i32)), ("face_patch", _P(_i32)),
                ("face_area", _P(_d)), ("face_centre", _P(_d)), ("face_normal", _P(_d)), ("face_sf", _P(_d)),
                ("face_dcf", _P(_d)), ("face_gc", _P(_d)), ("face_skew", _P(_d)), ("face_eps", _P(_d)),
                ("patch_bc", _P(_i32)), ("patch_start", _P(_i32)), ("patch_nfaces", _P(_i32)),
                ("periodic_pair", _P(_i32)), ("periodic_pair_dcf", _P(_d)), ("periodic_gc", _P(_d)),
                ("ls_offset", _P(_i32)), ("ls_cell", _P(_i32)), ("ls_face", _P(_i32)), ("ls_weight", _P(_d)),
                ("vertex_coords", _P(_d)), ("cell_vertex_offset", _P(_i64)), ("cell_vertices", _P(_i32)),
                ("face_node_offset", _P(_i64)), ("face_nodes", _P(_i32))]


class _PolyInput(ctypes.Structure):
    _fields_ = [("n_points", _i64), ("points", _P(_d)), ("n_faces", _i64), ("face_offset", _P(_i64)),
                ("face_nodes", _P(_i32)), ("n_int_faces", _i64), ("owner", _P(_i32)), ("neighbour", _P(_i32)),
                ("n_patches", _i32), ("patch_start", _P(_i64)), ("patch_nfaces", _P(_i64)),
                ("patch_bc", _P(_i32))]


def _ptr(a, ct):
    return a.ctypes.data_as(_P(ct))


def _copy(p, n, shape=None):
    if n == 0:
        a = np.zeros(0, dtype=np.dtype(p._type_))
    else:
        a = np.ctypeslib.as_array(p, shape=(n,)).copy()
    return a.reshape(shape) if shape else a


class Mesh:
    """A built mesh: plain numpy arrays (0-based), the data every hot-path kernel reads.

    Attribute names follow the reference's Mesh SoA mirror (src/mesh/mod_mesh.f90:164-186).
    """

    ARRAYS = ("cell_centroid", "cell_volume", "cell_face_offset", "cell_faces", "face_owner",
              "face_neighbour", "face_patch", "face_area", "face_centre", "face_normal", "face_sf",
              "face_dcf", "face_gc", "face_skew", "face_eps", "patch_bc", "patch_start", "patch_nfaces",
              "periodic_pair", "periodic_pair_dcf", "periodic_gc", "ls_offset", "ls_cell", "ls_face",
              "ls_weight")

    def __init__(self, elements):
        """``elements``: Gmsh-style element lists (box_elements, read_gmsh_elements) or face-based
        input (read_foam_input: keys points, face_offset, face_nodes, owner, neighbour,
        patch_start, patch_nfaces, patch_bc, patch_names)."""
        L = lib()
        L.dm_mesh_build.restype = ctypes.c_void_p
        L.dm_mesh_build.argtypes = [ctypes.POINTER(_Input)]
        L.dm_mesh_build_poly.restype = ctypes.c_void_p
        L.dm_mesh_build_poly.argtypes = [ctypes.POINTER(_PolyInput)]
        L.dm_mesh_get_view.argtypes = [ctypes.c_void_p, ctypes.POINTER(_View)]
        L.dm_mesh_free.argtypes = [ctypes.c_void_p]
        L.dm_last_error.restype = ctypes.c_char_p
        e = elements
        poly = "face_offset" in e
        if poly:
            keep = {
                "points": np.ascontiguousarray(e["points"], dtype=np.float64).reshape(-1),
                "face_offset": np.ascontiguousarray(e["face_offset"], dtype=np.int64),
                "face_nodes": np.ascontiguousarray(e["face_nodes"], dtype=np.int32),
                "owner": np.ascontiguousarray(e["owner"], dtype=np.int32),
                "neighbour": np.ascontiguousarray(e["neighbour"], dtype=np.int32),
                "patch_start": np.ascontiguousarray(e["patch_start"], dtype=np.int64),
                "patch_nfaces": np.ascontiguousarray(e["patch_nfaces"], dtype=np.int64),
                "patch_bc": np.ascontiguousarray(e["patch_bc"], dtype=np.int32),
            }
            inp = _PolyInput(
                n_points=len(keep["points"]) // 3, points=_ptr(keep["points"], _d),
                n_faces=len(keep["face_offset"]) - 1, face_offset=_ptr(keep["face_offset"], _i64),
                face_nodes=_ptr(keep["face_nodes"], _i32), n_int_faces=len(keep["neighbour"]),
                owner=_ptr(keep["owner"], _i32), neighbour=_ptr(keep["neighbour"], _i32),
                n_patches=len(keep["patch_bc"]), patch_start=_ptr(keep["patch_start"], _i64),
                patch_nfaces=_ptr(keep["patch_nfaces"], _i64), patch_bc=_ptr(keep["patch_bc"], _i32))
            h = L.dm_mesh_build_poly(ctypes.byref(inp))
        else:
            keep = {k: np.ascontiguousarray(v) for k, v in e.items() if isinstance(v, np.ndarray)}
            inp = _Input(
                n_nodes=len(keep["coords"]) // 3, coords=_ptr(keep["coords"], _d),
                n_vol=len(keep["vol_type"]), vol_type=_ptr(keep["vol_type"], _i32),
                vol_offset=_ptr(keep["vol_offset"], _i64), vol_nodes=_ptr(keep["vol_nodes"], _i32),
                n_bnd=len(keep["bnd_type"]), bnd_type=_ptr(keep["bnd_type"], _i32),
                bnd_offset=_ptr(keep["bnd_offset"], _i64), bnd_nodes=_ptr(keep["bnd_nodes"], _i32),
                bnd_group=_ptr(keep["bnd_group"], _i32), n_groups=len(keep["group_bc"]),
                group_bc=_ptr(keep["group_bc"], _i32))
            h = L.dm_mesh_build(ctypes.byref(inp))
        if not h:
            raise RuntimeError("mesh build failed: " + L.dm_last_error().decode())
        try:
            v = _View()
            L.dm_mesh_get_view(h, ctypes.byref(v))
            nc, nf, nb, npch = v.n_cells, v.n_faces, v.n_bnd_faces, v.n_patches
            self.n_cells, self.n_faces, self.n_int_faces, self.n_bnd_faces = nc, nf, v.n_int_faces, nb
            self.n_patches, self.n_ls, self.max_cell_faces = npch, v.n_ls, v.max_cell_faces
            self.ls_degenerate = v.ls_degenerate
            self.cell_centroid = _copy(v.cell_centroid, 3 * nc, (nc, 3))
            self.cell_volume = _copy(v.cell_volume, nc)
            self.cell_face_offset = _copy(v.cell_face_offset, nc + 1)
            self.cell_faces = _copy(v.cell_faces, int(self.cell_face_offset[-1]))
            for name in ("face_owner", "face_neighbour", "face_patch", "face_area", "face_dcf", "face_gc",
                         "face_skew"):
                setattr(self, name, _copy(getattr(v, name), nf))
            for name in ("face_centre", "face_normal", "face_sf", "face_eps"):
                setattr(self, name, _copy(getattr(v, name), 3 * nf, (nf, 3)))
            for name in ("patch_bc", "patch_start", "patch_nfaces"):
                setattr(self, name, _copy(getattr(v, name), npch))
            for name in ("periodic_pair", "periodic_pair_dcf", "periodic_gc"):
                setattr(self, name, _copy(getattr(v, name), nb))
            self.ls_offset = _copy(v.ls_offset, nc + 1)
            self.ls_cell = _copy(v.ls_cell, v.n_ls)
            self.ls_face = _copy(v.ls_face, v.n_ls)
            self.ls_weight = _copy(v.ls_weight, 3 * v.n_ls, (v.n_ls, 3))
            self.vertex_coords = _copy(v.vertex_coords, 3 * v.n_vertices, (v.n_vertices, 3))
            self.cell_vertex_offset = _copy(v.cell_vertex_offset, nc + 1)
            self.cell_vertices = _copy(v.cell_vertices, int(self.cell_vertex_offset[-1]))
            self.face_node_offset = _copy(v.face_node_offset, nf + 1)
            self.face_nodes = _copy(v.face_nodes, int(self.face_node_offset[-1]))
        finally:
            L.dm_mesh_free(h)
        if poly:
            # every patch of the input, empty ones included
            self.patch_names = list(elements.get("patch_names", []))
        else:
            # patches are the groups that hold boundary faces, in group order (an empty group gets none)
            names = list(elements.get("group_names", []))
            used = np.bincount(keep["bnd_group"], minlength=len(keep["group_bc"])) > 0
            self.patch_names = [n for n, u in zip(names, used) if u] if names else []
        self.bounds_min = self.vertex_coords.min(axis=0)
        self.bounds_max = self.vertex_coords.max(axis=0)

    def patch_of_name(self, name):
        return self.patch_names.index(name)


def box_elements(nx, ny, nz, node_xyz, group_names=("x0", "xL", "y0", "yL", "z0", "zL"),
                 group_types=("periodic",) * 6):
    """Element lists of a structured nx*ny*nz hex box.

    Node, hex and boundary-quad order follow tests/regression/mk_box_mesh.py (node id
    i + (nx+1)*(j + (ny+1)*k), hexes k-j-i nested, quads per surface x0,xL,y0,yL,z0,zL with the
    first in-surface index outermost). ``node_xyz(I, J, K)`` maps integer node indices to
    coordinates (arrays).
    """
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    # node id order: i fastest, then j, then k
    I, J, K = (a.transpose(2, 1, 0).reshape(-1) for a in (I, J, K))
    X, Y, Z = node_xyz(I, J, K)
    coords = np.stack([X, Y, Z], axis=1).reshape(-1).astype(np.float64)

    def nid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.reshape(-1), j.reshape(-1), k.reshape(-1)
    hexes = np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                      nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)],
                     axis=1).astype(np.int32)

    def surf(n1, n2, fn):
        a, b = np.meshgrid(np.arange(n1), np.arange(n2), indexing="ij")
        a, b = a.reshape(-1), b.reshape(-1)
        return np.stack(fn(a, b), axis=1)

    quads = [
        surf(ny, nz, lambda a, b: [nid(0, a, b), nid(0, a + 1, b), nid(0, a + 1, b + 1), nid(0, a, b + 1)]),
        surf(ny, nz, lambda a, b: [nid(nx, a, b), nid(nx, a + 1, b), nid(nx, a + 1, b + 1), nid(nx, a, b + 1)]),
        surf(nx, nz, lambda a, b: [nid(a, 0, b), nid(a + 1, 0, b), nid(a + 1, 0, b + 1), nid(a, 0, b + 1)]),
        surf(nx, nz, lambda a, b: [nid(a, ny, b), nid(a + 1, ny, b), nid(a + 1, ny, b + 1), nid(a, ny, b + 1)]),
        surf(nx, ny, lambda a, b: [nid(a, b, 0), nid(a + 1, b, 0), nid(a + 1, b + 1, 0), nid(a, b + 1, 0)]),
        surf(nx, ny, lambda a, b: [nid(a, b, nz), nid(a + 1, b, nz), nid(a + 1, b + 1, nz), nid(a, b + 1, nz)]),
    ]
    bnd = np.concatenate(quads).astype(np.int32)
    bnd_group = np.concatenate([np.full(len(q), g, dtype=np.int32) for g, q in enumerate(quads)])
    nv, nq = len(hexes), len(bnd)
    return {
        "coords": coords,
        "vol_type": np.full(nv, GMSH_HEX8, dtype=np.int32),
        "vol_offset": (np.arange(nv + 1, dtype=np.int64) * 8),
        "vol_nodes": hexes.reshape(-1),
        "bnd_type": np.full(nq, GMSH_QUAD4, dtype=np.int32),
        "bnd_offset": (np.arange(nq + 1, dtype=np.int64) * 4),
        "bnd_nodes": bnd.reshape(-1),
        "bnd_group": bnd_group,
        "group_bc": np.array([BC_NAMES[t.lower()] for t in group_types], dtype=np.int32),
        "group_names": list(group_names),
    }


def _round16g(a):
    """Round-trip through '%.16g', as a Gmsh ASCII file written with that format stores it."""
    return np.array([float(f"{x:.16g}") for x in np.asarray(a, dtype=np.float64).reshape(-1)]).reshape(np.shape(a))


def skewed_box(n, amp, group_types=("periodic",) * 6, group_names=("x0", "xL", "y0", "yL", "z0", "zL")):
    """Curvilinear unit box of tests/regression/mk_box_mesh.py (N^3 hexes, node displacement amp)."""
    return Mesh(skewed_box_elements(n, amp, group_types, group_names))


def skewed_box_elements(n, amp, group_types=("periodic",) * 6, group_names=("x0", "xL", "y0", "yL", "z0", "zL")):
    """Element lists of skewed_box."""
    def xyz(I, J, K):
        out = np.empty((3, len(I)))
        for q, (i, j, k) in enumerate(zip(I.tolist(), J.tolist(), K.tolist())):
            X, Y, Z = i / n, j / n, k / n
            out[0, q] = X + amp * math.sin(math.pi * X) * math.sin(2 * math.pi * Y) * math.sin(math.pi * Z)
            out[1, q] = Y + amp * math.sin(math.pi * Y) * math.sin(2 * math.pi * Z) * math.sin(math.pi * X)
            out[2, q] = Z + amp * math.sin(math.pi * Z) * math.sin(2 * math.pi * X) * math.sin(math.pi * Y)
        return _round16g(out[0]), _round16g(out[1]), _round16g(out[2])
    return box_elements(n, n, n, xyz, group_names, group_types)


def uniform_box(nx, ny, nz, lx, ly, lz, group_types=("periodic",) * 6,
                group_names=("x0", "xL", "y0", "yL", "z0", "zL")):
    """Axis-aligned box [0,lx]x[0,ly]x[0,lz] with uniform spacing."""
    def xyz(I, J, K):
        return I * (lx / nx), J * (ly / ny), K * (lz / nz)
    return Mesh(box_elements(nx, ny, nz, xyz, group_names, group_types))


def channel_box(nx, ny, nz, lx=2 * math.pi, ly=2.0, lz=math.pi, stretch=2.0):
    """Channel box with tanh wall-normal (y) stretching: y_j = ly/2 (1 + tanh(s(2j/ny-1))/tanh(s)).

    Periodic in x and z, log-law wall-model walls at y = 0 and y = ly.
    """
    yn = 0.5 * ly * (1.0 + np.tanh(stretch * (2.0 * np.arange(ny + 1) / ny - 1.0)) / math.tanh(stretch))
    yn[0], yn[-1] = 0.0, ly

    def xyz(I, J, K):
        return I * (lx / nx), yn[J], K * (lz / nz)
    return Mesh(box_elements(nx, ny, nz, xyz, ("x0", "xL", "bottom_wall", "top_wall", "z0", "zL"),
                             ("periodic", "periodic", "wallModel", "wallModel", "periodic", "periodic")))


def rb_box(nx, ny, nz, lx=4.0, ly=4.0, lz=1.0, stretch=1.5):
    """Rayleigh-Benard cell: periodic in x and y, wall-clustered tanh spacing in z
    (z_k = lz/2 (1 + tanh(s(2k/nz-1))/tanh(s))), fixed-temperature walls "bottom_wall" (z = 0) and
    "top_wall" (z = lz)."""
    zn = 0.5 * lz * (1.0 + np.tanh(stretch * (2.0 * np.arange(nz + 1) / nz - 1.0)) / math.tanh(stretch))
    zn[0], zn[-1] = 0.0, lz

    def xyz(I, J, K):
        return I * (lx / nx), J * (ly / ny), zn[K]
    return Mesh(box_elements(nx, ny, nz, xyz, ("x0", "xL", "y0", "yL", "bottom_wall", "top_wall"),
                             ("periodic", "periodic", "periodic", "periodic", "fixedTemperature",
                              "fixedTemperature")))


GMSH_TRI3, GMSH_TET4 = 2, 4

# Kuhn (Freudenthal) split of a hex into 6 tets sharing the diagonal from corner (0,0,0) to
# (1,1,1): one tet per axis order. Corner offsets (di, dj, dk) along each path.
_KUHN_PATHS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def tet_box(n, lx=1.0, ly=1.0, lz=1.0, jitter=0.15, seed=3200, shuffle_seed=None,
            group_types=("noSlip",) * 6, group_names=("x0", "xL", "y0", "yL", "z0", "zL")):
    """Mesh of tet_box_elements."""
    return Mesh(tet_box_elements(n, lx, ly, lz, jitter, seed, shuffle_seed, group_types, group_names))


def tet_box_elements(n, lx=1.0, ly=1.0, lz=1.0, jitter=0.15, seed=3200, shuffle_seed=None,
                     group_types=("noSlip",) * 6, group_names=("x0", "xL", "y0", "yL", "z0", "zL")):
    """n^3 hexes split into 6 tets each (Kuhn split, conforming across hexes), interior nodes
    jittered by a seeded uniform +-jitter of the local spacing per coordinate, tets optionally put
    in a seeded random order (cell ids shuffled). Boundary quads split along the same diagonals."""
    nx = ny = nz = n
    hx, hy, hz = lx / nx, ly / ny, lz / nz

    def nid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    I, J, K = (a.transpose(2, 1, 0).reshape(-1) for a in (I, J, K))
    X, Y, Z = I * hx, J * hy, K * hz
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        r = rng.uniform(-jitter, jitter, size=(len(X), 3))
        inner = (I > 0) & (I < nx) & (J > 0) & (J < ny) & (K > 0) & (K < nz)
        X = np.where(inner, X + r[:, 0] * hx, X)
        Y = np.where(inner, Y + r[:, 1] * hy, Y)
        Z = np.where(inner, Z + r[:, 2] * hz, Z)
    coords = np.stack([X, Y, Z], axis=1).reshape(-1).astype(np.float64)

    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.reshape(-1), j.reshape(-1), k.reshape(-1)
    tets = []
    for path in _KUHN_PATHS:
        d = np.zeros((3, len(i)), dtype=np.int64)
        verts = [nid(i, j, k)]
        for ax in path:
            d[ax] += 1
            verts.append(nid(i + d[0], j + d[1], k + d[2]))
        tets.append(np.stack(verts, axis=1))
    tets = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)  # 6 consecutive tets per hex
    if shuffle_seed is not None:
        tets = tets[np.random.default_rng(shuffle_seed).permutation(len(tets))]

    def surf(n1, n2, fn):
        a, b = np.meshgrid(np.arange(n1), np.arange(n2), indexing="ij")
        a, b = a.reshape(-1), b.reshape(-1)
        c00, c10, c11, c01 = fn(a, b)
        # split along the (0,0)-(1,1) diagonal, as the Kuhn tets do on every face
        return np.concatenate([np.stack([c00, c10, c11], axis=1), np.stack([c00, c01, c11], axis=1)])

    tris = [
        surf(ny, nz, lambda a, b: (nid(0, a, b), nid(0, a + 1, b), nid(0, a + 1, b + 1), nid(0, a, b + 1))),
        surf(ny, nz, lambda a, b: (nid(nx, a, b), nid(nx, a + 1, b), nid(nx, a + 1, b + 1), nid(nx, a, b + 1))),
        surf(nx, nz, lambda a, b: (nid(a, 0, b), nid(a + 1, 0, b), nid(a + 1, 0, b + 1), nid(a, 0, b + 1))),
        surf(nx, nz, lambda a, b: (nid(a, ny, b), nid(a + 1, ny, b), nid(a + 1, ny, b + 1), nid(a, ny, b + 1))),
        surf(nx, ny, lambda a, b: (nid(a, b, 0), nid(a + 1, b, 0), nid(a + 1, b + 1, 0), nid(a, b + 1, 0))),
        surf(nx, ny, lambda a, b: (nid(a, b, nz), nid(a + 1, b, nz), nid(a + 1, b + 1, nz), nid(a, b + 1, nz))),
    ]
    bnd = np.concatenate(tris).astype(np.int32)
    bnd_group = np.concatenate([np.full(len(t), g, dtype=np.int32) for g, t in enumerate(tris)])
    nv, nt = len(tets), len(bnd)
    return {
        "coords": coords,
        "vol_type": np.full(nv, GMSH_TET4, dtype=np.int32),
        "vol_offset": np.arange(nv + 1, dtype=np.int64) * 4,
        "vol_nodes": tets.reshape(-1),
        "bnd_type": np.full(nt, GMSH_TRI3, dtype=np.int32),
        "bnd_offset": np.arange(nt + 1, dtype=np.int64) * 3,
        "bnd_nodes": bnd.reshape(-1),
        "bnd_group": bnd_group,
        "group_bc": np.array([BC_NAMES[t.lower()] for t in group_types], dtype=np.int32),
        "group_names": list(group_names),
    }


class _GmshView(ctypes.Structure):
    _fields_ = [("n_nodes", _i64), ("coords", _P(_d)),
                ("n_vol", _i64), ("vol_type", _P(_i32)), ("vol_offset", _P(_i64)), ("vol_nodes", _P(_i32)),
                ("n_bnd", _i64), ("bnd_type", _P(_i32)), ("bnd_offset", _P(_i64)), ("bnd_nodes", _P(_i32)),
                ("bnd_group", _P(_i32)), ("n_groups", _i32)]


def read_gmsh_elements(path, patch_types=None, warn=print):
    """Element lists of a Gmsh v4 ASCII file (src/mesh/mod_gmsh_reader.f90:45-236), parsed natively.

    ``patch_types`` maps a 2-D physical group name to a patch_type name (BC_NAMES), as the
    reference's &dopamine_bc entries do. A group with no entry becomes zeroGradient with a warning;
    an entry naming no group of the file is an error (src/bc/mod_bc_setup.f90:42-106).
    """
    L = lib()
    L.dm_gmsh_read.restype = ctypes.c_void_p
    L.dm_gmsh_read.argtypes = [ctypes.c_char_p]
    L.dm_gmsh_get_view.argtypes = [ctypes.c_void_p, ctypes.POINTER(_GmshView)]
    L.dm_gmsh_group_name.restype = ctypes.c_char_p
    L.dm_gmsh_group_name.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.dm_gmsh_free.argtypes = [ctypes.c_void_p]
    L.dm_gmsh_last_error.restype = ctypes.c_char_p
    h = L.dm_gmsh_read(str(path).encode())
    if not h:
        raise ValueError(f"{path}: " + L.dm_gmsh_last_error().decode())
    try:
        v = _GmshView()
        L.dm_gmsh_get_view(h, ctypes.byref(v))
        names = [L.dm_gmsh_group_name(h, g).decode() for g in range(v.n_groups)]
        e = {
            "coords": _copy(v.coords, 3 * v.n_nodes),
            "vol_type": _copy(v.vol_type, v.n_vol),
            "vol_offset": _copy(v.vol_offset, v.n_vol + 1),
            "bnd_type": _copy(v.bnd_type, v.n_bnd),
            "bnd_offset": _copy(v.bnd_offset, v.n_bnd + 1),
            "bnd_group": _copy(v.bnd_group, v.n_bnd),
        }
        e["vol_nodes"] = _copy(v.vol_nodes, int(e["vol_offset"][-1]))
        e["bnd_nodes"] = _copy(v.bnd_nodes, int(e["bnd_offset"][-1]))
    finally:
        L.dm_gmsh_free(h)
    types = {k: v for k, v in (patch_types or {}).items()}
    missing = [k for k in types if k not in names]
    if missing:
        raise ValueError("input BC patch names not in the mesh: " + ", ".join(missing))
    bc = []
    for n in names:
        if n in types:
            bc.append(BC_NAMES[types[n].lower()])
        else:
            if warn:
                warn(f'mesh patch "{n}" has no patch type; using zeroGradient')
            bc.append(BC_ZEROGRADIENT)
    e["group_bc"] = np.array(bc, dtype=np.int32)
    e["group_names"] = names
    return e


def read_gmsh(path, patch_types=None, warn=print):
    """Mesh of a Gmsh v4 ASCII file; see read_gmsh_elements."""
    return Mesh(read_gmsh_elements(path, patch_types, warn))


def write_gmsh(elements, path):
    """Write element lists as a Gmsh 4.1 ASCII file that read_gmsh_elements reads back unchanged.

    One entity per dimension: volume elements on volume 1 (physical "fluid"), boundary group g on
    surface g+1 with physical tag g+1. Node tags are 1-based positions; coordinates are written
    with '%.17g' so they read back bit for bit.
    """
    e = elements
    names = list(e["group_names"])
    ng = len(names)
    coords = np.asarray(e["coords"], dtype=np.float64).reshape(-1, 3)
    nn = len(coords)
    lines = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$PhysicalNames", str(ng + 1)]
    lines += [f'2 {g + 1} "{n}"' for g, n in enumerate(names)]
    lines += [f'3 {ng + 1} "fluid"', "$EndPhysicalNames", "$Entities", f"0 0 {ng} 1"]
    lo, hi = coords.min(axis=0), coords.max(axis=0)
    box = " ".join(f"{x:.17g}" for x in (*lo, *hi))
    lines += [f"{g + 1} {box} 1 {g + 1} 0" for g in range(ng)]
    lines += [f"1 {box} 1 {ng + 1} 0", "$EndEntities", "$Nodes", f"1 {nn} 1 {nn}", f"3 1 0 {nn}"]
    lines += [str(i + 1) for i in range(nn)]
    lines += [f"{x:.17g} {y:.17g} {z:.17g}" for x, y, z in coords]
    lines += ["$EndNodes"]
    vt, vo, vn = e["vol_type"], e["vol_offset"], e["vol_nodes"]
    bt, bo, bn, bg = e["bnd_type"], e["bnd_offset"], e["bnd_nodes"], e["bnd_group"]
    # blocks: one per (entity, element type) run, in element order
    blocks = []

    def runs(types, key):
        start = 0
        for i in range(1, len(types) + 1):
            if i == len(types) or types[i] != types[start] or key(i) != key(start):
                blocks.append((start, i, types, key(start)))
                start = i
    runs(bt, lambda i: int(bg[i]) if i < len(bg) else -1)
    nb_blocks = len(blocks)
    runs(vt, lambda i: -1)
    tag = 1
    body = []
    for b, (s0, s1, types, grp) in enumerate(blocks):
        is_bnd = b < nb_blocks
        off, nod = (bo, bn) if is_bnd else (vo, vn)
        body.append(f"{2 if is_bnd else 3} {grp + 1 if is_bnd else 1} {int(types[s0])} {s1 - s0}")
        for i in range(s0, s1):
            body.append(f"{tag} " + " ".join(str(int(x) + 1) for x in nod[off[i]:off[i + 1]]))
            tag += 1
    lines += ["$Elements", f"{len(blocks)} {tag - 1} 1 {tag - 1}"] + body + ["$EndElements"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _resolve_patch_types(names, own_bc, patch_types, warn):
    """bc id per patch: the case configuration's patch type by name, else the mesh's own
    (src/bc/mod_bc_setup.f90:42-106); an input name missing from the mesh is an error."""
    types = dict(patch_types or {})
    missing = [k for k in types if k not in names]
    if missing:
        raise ValueError("input BC patch names not in the mesh: " + ", ".join(missing))
    out = []
    for n, b in zip(names, own_bc):
        if n in types:
            out.append(BC_NAMES[types[n].lower()])
        else:
            if warn:
                warn(f'mesh patch "{n}" is not listed in the patch types; using its mesh type (bc id {b})')
            out.append(b)
    return np.array(out, dtype=np.int32)


def read_foam_input(polymesh_dir, patch_types=None, warn=print):
    """Face-based input of an OpenFOAM constant/polyMesh directory, parsed natively
    (src/mesh/mod_foam_reader.f90:16-97). A patch's bc id comes from ``patch_types`` by name when
    listed, else from its OpenFOAM ``type`` (bc_type_to_id, :1144-1167)."""
    L = lib()
    L.dm_foam_read.restype = ctypes.c_void_p
    L.dm_foam_read.argtypes = [ctypes.c_char_p]
    L.dm_foam_get_view.argtypes = [ctypes.c_void_p, ctypes.POINTER(_PolyInput)]
    L.dm_foam_patch_name.restype = ctypes.c_char_p
    L.dm_foam_patch_name.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.dm_foam_patch_type.restype = ctypes.c_char_p
    L.dm_foam_patch_type.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.dm_foam_free.argtypes = [ctypes.c_void_p]
    L.dm_foam_last_error.restype = ctypes.c_char_p
    h = L.dm_foam_read(str(polymesh_dir).encode())
    if not h:
        raise ValueError(f"{polymesh_dir}: " + L.dm_foam_last_error().decode())
    try:
        v = _PolyInput()
        L.dm_foam_get_view(h, ctypes.byref(v))
        nf, npch = v.n_faces, v.n_patches
        e = {
            "points": _copy(v.points, 3 * v.n_points, (v.n_points, 3)),
            "face_offset": _copy(v.face_offset, nf + 1),
            "owner": _copy(v.owner, nf),
            "neighbour": _copy(v.neighbour, v.n_int_faces),
            "patch_start": _copy(v.patch_start, npch),
            "patch_nfaces": _copy(v.patch_nfaces, npch),
            "patch_names": [L.dm_foam_patch_name(h, p).decode() for p in range(npch)],
            "patch_types": [L.dm_foam_patch_type(h, p).decode() for p in range(npch)],
        }
        e["face_nodes"] = _copy(v.face_nodes, int(e["face_offset"][-1]))
        own_bc = _copy(v.patch_bc, npch)
    finally:
        L.dm_foam_free(h)
    e["patch_bc"] = _resolve_patch_types(e["patch_names"], own_bc, patch_types, warn)
    return e


def read_foam(polymesh_dir, patch_types=None, warn=print):
    """Mesh of an OpenFOAM polyMesh directory; see read_foam_input."""
    return Mesh(read_foam_input(polymesh_dir, patch_types, warn))


# OpenFOAM patch type written for each bc id: the inverse of bc_type_to_id where it has one
# (src/mesh/mod_foam_reader.f90:1144-1167); other ids are written as "patch", whose bc id then has
# to come from the case's patch types.
_FOAM_TYPE = {BC_INLET: "inlet", BC_OUTLET: "outlet", BC_ZEROGRADIENT: "zeroGradient", BC_WALL_MODEL: "wall",
              BC_FIXEDTEMP: "fixedTemperature", BC_WALL_VELOCITY: "movingWall", BC_SYMMETRY: "symmetryPlane",
              BC_SLIP: "empty", BC_PERIODIC: "cyclic"}


def write_foam(mesh, polymesh_dir, binary=False, label_bytes=4, scalar_bytes=8, msb=False, compact=True):
    """Write a built Mesh as an OpenFOAM polyMesh directory (points, faces, owner, neighbour,
    boundary) that read_foam reads back with the same faces, cells and patches.

    Faces keep the mesh's order (internal faces first, each with owner < neighbour; boundary faces
    by patch) and their points run so that the right-hand normal points from owner to neighbour
    or out of the domain. ASCII writes coordinates with '%.17g'; binary writes raw labels and
    scalars in the requested widths and byte order, faces as a compact offset/label pair
    (``compact``) or as (n, labels) records.
    """
    import os
    os.makedirs(polymesh_dir, exist_ok=True)
    order = ">" if msb else "<"
    lt, st = np.dtype(f"{order}i{label_bytes}"), np.dtype(f"{order}f{scalar_bytes}")
    arch = f'"{"MSB" if msb else "LSB"};label={8 * label_bytes};scalar={8 * scalar_bytes}"'

    def header(cls, obj, fmt):
        return ("FoamFile\n{\n    version     2.0;\n    format      " + fmt + ";\n    arch        " + arch +
                ";\n    class       " + cls + ";\n    location    \"constant/polyMesh\";\n    object      " +
                obj + ";\n}\n\n").encode()

    def write(name, cls, lists, ascii_body=None):
        with open(os.path.join(polymesh_dir, name), "wb") as f:
            f.write(header(cls, name, "binary" if binary else "ascii"))
            if binary:
                for a in lists:  # a list of vectors is counted in vectors
                    f.write(f"{len(a)}\n(".encode())
                    f.write(np.ascontiguousarray(a).tobytes())
                    f.write(b")\n")
            else:
                f.write(ascii_body.encode())

    nf, nint = mesh.n_faces, mesh.n_int_faces
    off, nodes = mesh.face_node_offset, mesh.face_nodes
    pts = mesh.vertex_coords
    write("points", "vectorField", [pts.astype(st)],
          f"{len(pts)}\n(\n" + "".join(f"({x:.17g} {y:.17g} {z:.17g})\n" for x, y, z in pts) + ")\n")
    if binary and compact:
        write("faces", "faceCompactList", [off.astype(lt), nodes.astype(lt)])
    elif binary:
        rec = np.concatenate([np.concatenate([[off[f + 1] - off[f]], nodes[off[f]:off[f + 1]]]) for f in range(nf)])
        with open(os.path.join(polymesh_dir, "faces"), "wb") as fh:
            fh.write(header("faceList", "faces", "binary"))
            fh.write(f"{nf}\n(".encode())
            fh.write(rec.astype(lt).tobytes())
            fh.write(b")\n")
    else:
        write("faces", "faceList", None, f"{nf}\n(\n" + "".join(
            f"{off[f + 1] - off[f]}(" + " ".join(str(int(x)) for x in nodes[off[f]:off[f + 1]]) + ")\n"
            for f in range(nf)) + ")\n")
    write("owner", "labelList", [mesh.face_owner.astype(lt)],
          f"{nf}\n(\n" + "".join(f"{int(x)}\n" for x in mesh.face_owner) + ")\n")
    nb = mesh.face_neighbour[:nint]
    write("neighbour", "labelList", [nb.astype(lt)], f"{nint}\n(\n" + "".join(f"{int(x)}\n" for x in nb) + ")\n")
    names = mesh.patch_names or [f"patch{p}" for p in range(mesh.n_patches)]
    body = f"{mesh.n_patches}\n(\n"
    for p in range(mesh.n_patches):
        body += (f"    {names[p]}\n    {{\n        type            {_FOAM_TYPE.get(int(mesh.patch_bc[p]), 'patch')};\n"
                 f"        nFaces          {int(mesh.patch_nfaces[p])};\n"
                 f"        startFace       {nint + int(mesh.patch_start[p])};\n    }}\n")
    body += ")\n"
    with open(os.path.join(polymesh_dir, "boundary"), "wb") as f:
        f.write(header("polyBoundaryMesh", "boundary", "ascii"))
        f.write(body.encode())


def obstacle_box_elements(nx, ny, nz, lx, ly, lz, obstacles, group_names=("x0", "xL", "y0", "yL", "z0", "zL"),
                          group_types=("periodic", "periodic", "periodic", "periodic", "wallModel", "slip"),
                          obstacle_name="obstacles", obstacle_type="wallModel"):
    """Uniform nx*ny*nz hex box with cuboid obstacles: every cell whose centre lies inside an
    obstacle ((x0, x1, y0, y1, z0, z1) in coordinates) is removed, and the faces between kept and
    removed cells form one more patch, `obstacle_name`. Node numbering and the order of the kept
    hexes and boundary quads follow box_elements; boundary quads of removed cells are dropped."""
    def xyz(I, J, K):
        return I * (lx / nx), J * (ly / ny), K * (lz / nz)
    e = box_elements(nx, ny, nz, xyz, group_names, group_types)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.reshape(-1), j.reshape(-1), k.reshape(-1)
    cx, cy, cz = (i + 0.5) * lx / nx, (j + 0.5) * ly / ny, (k + 0.5) * lz / nz
    removed = np.zeros(nx * ny * nz, dtype=bool)
    for (x0, x1, y0, y1, z0, z1) in obstacles:
        removed |= (cx > x0) & (cx < x1) & (cy > y0) & (cy < y1) & (cz > z0) & (cz < z1)
    if removed.all():
        raise ValueError("the obstacles remove every cell")
    keep = ~removed
    hexes = e["vol_nodes"].reshape(-1, 8)

    def cid(ii, jj, kk):
        return ii + nx * (jj + ny * kk)

    # outer quads in box_elements' order, each tagged with its cell
    cells = [
        (np.zeros(ny * nz, int), *[a.reshape(-1) for a in np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")]),
        (np.full(ny * nz, nx - 1), *[a.reshape(-1) for a in np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij")]),
    ]
    a, b = [q.reshape(-1) for q in np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")]
    cells += [(a, np.zeros_like(a), b), (a, np.full_like(a, ny - 1), b)]
    a, b = [q.reshape(-1) for q in np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")]
    cells += [(a, b, np.zeros_like(a)), (a, b, np.full_like(a, nz - 1))]
    owner = np.concatenate([cid(ci, cj, ck) for ci, cj, ck in cells])
    quads = e["bnd_nodes"].reshape(-1, 4)
    ok = keep[owner]
    bnd = [quads[ok]]
    grp = [e["bnd_group"][ok]]
    # faces between a kept cell and a removed neighbour: direction by direction (-x +x -y +y -z
    # +z), kept cells ascending within each
    corner = {(-1, 0, 0): (0, 3, 7, 4), (1, 0, 0): (1, 2, 6, 5), (0, -1, 0): (0, 1, 5, 4), (0, 1, 0): (3, 2, 6, 7),
              (0, 0, -1): (0, 1, 2, 3), (0, 0, 1): (4, 5, 6, 7)}
    g_obst = len(group_names)
    for (di, dj, dk), loc in corner.items():
        ni, nj, nk = i + di, j + dj, k + dk
        inb = (ni >= 0) & (ni < nx) & (nj >= 0) & (nj < ny) & (nk >= 0) & (nk < nz)
        sel = np.zeros_like(keep)
        sel[inb] = keep[inb] & removed[cid(ni[inb], nj[inb], nk[inb])]
        if sel.any():
            bnd.append(hexes[sel][:, list(loc)].astype(np.int32))
            grp.append(np.full(int(sel.sum()), g_obst, dtype=np.int32))
    bnd = np.concatenate(bnd).astype(np.int32)
    grp = np.concatenate(grp).astype(np.int32)
    kept = hexes[keep]
    nv, nq = len(kept), len(bnd)
    return {
        "coords": e["coords"],
        "vol_type": np.full(nv, GMSH_HEX8, dtype=np.int32),
        "vol_offset": np.arange(nv + 1, dtype=np.int64) * 8,
        "vol_nodes": kept.reshape(-1),
        "bnd_type": np.full(nq, GMSH_QUAD4, dtype=np.int32),
        "bnd_offset": np.arange(nq + 1, dtype=np.int64) * 4,
        "bnd_nodes": bnd.reshape(-1),
        "bnd_group": grp,
        "group_bc": np.array([BC_NAMES[t.lower()] for t in (*group_types, obstacle_type)], dtype=np.int32),
        "group_names": [*group_names, obstacle_name],
    }


def seeded_obstacles(n, lx, ly, hmin, hmax, seed, wmin=0.04, wmax=0.1, margin=0.05):
    """n cuboids on the ground: footprint sides uniform in [wmin, wmax] of the box's horizontal
    size, heights uniform in [hmin, hmax], positions uniform inside the box at least `margin` of
    its size from the sides, so that periodic side patches keep matching faces (seeded)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        wx, wy = rng.uniform(wmin, wmax) * lx, rng.uniform(wmin, wmax) * ly
        x0 = rng.uniform(margin * lx, (1.0 - margin) * lx - wx)
        y0 = rng.uniform(margin * ly, (1.0 - margin) * ly - wy)
        out.append((x0, x0 + wx, y0, y0 + wy, -1.0, rng.uniform(hmin, hmax)))
    return out


def rcm_elements(elements):
    """The same element lists with the volume elements in reverse Cuthill-McKee order of the
    cells' face graph (dm_mesh_rcm_order), for meshes whose cell numbering scatters neighbours
    (shuffled tets): the rebuilt mesh keeps its geometry and faces, with neighbours close in
    memory. Boundary elements are unchanged."""
    L = lib()
    L.dm_mesh_rcm_order.argtypes = [ctypes.c_void_p, _P(_i32)]
    L.dm_mesh_build.restype = ctypes.c_void_p
    L.dm_mesh_build.argtypes = [ctypes.POINTER(_Input)]
    L.dm_mesh_free.argtypes = [ctypes.c_void_p]
    L.dm_last_error.restype = ctypes.c_char_p
    e = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in elements.items()}
    inp = _Input(
        n_nodes=len(e["coords"]) // 3, coords=_ptr(e["coords"], _d),
        n_vol=len(e["vol_type"]), vol_type=_ptr(e["vol_type"], _i32),
        vol_offset=_ptr(e["vol_offset"], _i64), vol_nodes=_ptr(e["vol_nodes"], _i32),
        n_bnd=len(e["bnd_type"]), bnd_type=_ptr(e["bnd_type"], _i32),
        bnd_offset=_ptr(e["bnd_offset"], _i64), bnd_nodes=_ptr(e["bnd_nodes"], _i32),
        bnd_group=_ptr(e["bnd_group"], _i32), n_groups=len(e["group_bc"]), group_bc=_ptr(e["group_bc"], _i32))
    h = L.dm_mesh_build(ctypes.byref(inp))
    if not h:
        raise RuntimeError("mesh build failed: " + L.dm_last_error().decode())
    try:
        order = np.zeros(len(e["vol_type"]), dtype=np.int32)
        if L.dm_mesh_rcm_order(h, _ptr(order, _i32)) != 0:
            raise RuntimeError("RCM ordering failed: " + L.dm_last_error().decode())
    finally:
        L.dm_mesh_free(h)
    off = e["vol_offset"]
    cnt = np.diff(off)[order]
    new_off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    idx = np.concatenate([np.arange(off[c], off[c + 1]) for c in order]) if len(order) else np.zeros(0, np.int64)
    out = dict(elements)
    out["vol_type"] = e["vol_type"][order]
    out["vol_offset"] = new_off
    out["vol_nodes"] = e["vol_nodes"][idx]
    return out
