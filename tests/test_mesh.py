"""Mesh builder (libdopamine_hip.so, dm_mesh_build): geometry and connectivity checked against
closed forms, plus the least-squares stencil's exactness for linear fields. Reference:
src/mesh/mod_gmsh_reader.f90 (face numbering, geometry), src/mesh/mod_mesh.f90 (LS stencil)."""
import math

import numpy as np
import pytest

from dopamine_amd import mesh as M

WALLS = ("noSlip",) * 6


def _cell_sums(m):
    """sum_f sigma_f S_f per cell (zero for a closed cell)."""
    s = np.zeros((m.n_cells, 3))
    sf = m.face_sf.reshape(-1, 3)
    np.add.at(s, m.face_owner, sf)
    i = m.face_neighbour >= 0
    np.add.at(s, m.face_neighbour[i], -sf[i])
    return s


@pytest.mark.parametrize("n", [1, 3, 5])
def test_uniform_box_counts_and_geometry(n):
    m = M.uniform_box(n, n + 1, n + 2, 1.0, 2.0, 3.0, group_types=WALLS)
    nx, ny, nz = n, n + 1, n + 2
    assert m.n_cells == nx * ny * nz
    assert m.n_int_faces == (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    assert m.n_bnd_faces == 2 * (nx * ny + ny * nz + nx * nz)
    h = np.array([1.0 / nx, 2.0 / ny, 3.0 / nz])
    assert np.allclose(m.cell_volume, h.prod(), rtol=1e-14)
    assert math.isclose(m.cell_volume.sum(), 6.0, rel_tol=1e-13)
    # internal faces: owner has the lower index; unit normal along the centroid offset; gc = 1/2
    o, nb = m.face_owner[: m.n_int_faces], m.face_neighbour[: m.n_int_faces]
    assert np.all(o < nb) and np.all(m.face_neighbour[m.n_int_faces:] < 0)
    d = m.cell_centroid.reshape(-1, 3)[nb] - m.cell_centroid.reshape(-1, 3)[o]
    nrm = m.face_normal.reshape(-1, 3)[: m.n_int_faces]
    assert np.allclose(np.einsum("ij,ij->i", d, nrm), np.linalg.norm(d, axis=1), rtol=1e-13)
    assert np.allclose(m.face_gc[: m.n_int_faces], 0.5, atol=1e-14)
    assert np.allclose(np.linalg.norm(m.face_normal.reshape(-1, 3), axis=1), 1.0, atol=1e-14)
    assert np.allclose(_cell_sums(m), 0.0, atol=1e-14)


def test_boundary_patches_are_contiguous_and_outward():
    m = M.uniform_box(3, 4, 5, 1.0, 1.0, 1.0, group_types=WALLS)
    total = 0
    for p in range(m.n_patches):
        start, cnt = m.patch_start[p], m.patch_nfaces[p]
        assert start == total  # patch_start counts boundary faces
        total += cnt
        f = m.n_int_faces + np.arange(start, start + cnt)
        assert np.all(m.face_patch[f] == p)
        out = m.face_centre.reshape(-1, 3)[f] - m.cell_centroid.reshape(-1, 3)[m.face_owner[f]]
        assert np.all(np.einsum("ij,ij->i", out, m.face_normal.reshape(-1, 3)[f]) > 0)
    assert total == m.n_bnd_faces
    assert [m.patch_nfaces[p] for p in range(6)] == [20, 20, 15, 15, 12, 12]


@pytest.mark.parametrize("n,amp", [(6, 0.08), (8, 0.04)])
def test_skewed_box_closed_cells_and_volume(n, amp):
    m = M.skewed_box(n, amp, group_types=WALLS)
    assert np.allclose(_cell_sums(m), 0.0, atol=1e-14)
    # the boundary is the undisplaced unit box's (plane faces), so the cells tile volume 1
    assert math.isclose(m.cell_volume.sum(), 1.0, rel_tol=1e-12)
    assert np.all(m.cell_volume > 0)
    assert np.all(m.face_skew >= 0) and m.face_skew[: m.n_int_faces].max() > 0


def test_periodic_pairs_are_translations():
    m = M.skewed_box(6, 0.08)
    nb = m.n_bnd_faces
    pair = m.periodic_pair
    assert np.all(pair >= 0)
    assert np.array_equal(pair[pair], np.arange(nb))
    fc = m.face_centre.reshape(-1, 3)[m.n_int_faces:]
    shift = fc[pair] - fc
    # each partner sits one box length away along exactly one axis
    assert np.allclose(np.sort(np.abs(shift), axis=1), [[0, 0, 1]] * nb, atol=1e-12)
    sf = m.face_sf.reshape(-1, 3)[m.n_int_faces:]
    assert np.allclose(sf[pair], -sf, atol=1e-14)
    # the pair's distance is the partner's owner-to-face distance
    assert np.allclose(m.periodic_pair_dcf, m.face_dcf[m.n_int_faces:][pair], rtol=1e-14)


def _ls_moment(m):
    """sum_k w_k (x_k - x_c)^T per cell: the identity when the stencil reproduces linear fields."""
    ctr = m.cell_centroid.reshape(-1, 3)
    fc = m.face_centre.reshape(-1, 3)
    w = m.ls_weight.reshape(-1, 3)
    out = np.zeros((m.n_cells, 3, 3))
    for c in range(m.n_cells):
        for k in range(m.ls_offset[c], m.ls_offset[c + 1]):
            xk = ctr[m.ls_cell[k]] if m.ls_cell[k] >= 0 else fc[m.ls_face[k]]
            out[c] += np.outer(w[k], xk - ctr[c])
    return out


@pytest.mark.parametrize("maker", [lambda: M.uniform_box(4, 5, 6, 1.0, 2.0, 0.5, group_types=WALLS),
                                   lambda: M.skewed_box(6, 0.08, group_types=WALLS)])
def test_least_squares_stencil_reproduces_linear_fields(maker):
    m = maker()
    mom = _ls_moment(m)
    assert np.abs(mom - np.eye(3)).max() < 1e-11
    assert np.all(np.diff(m.ls_offset) >= 6)


def test_channel_box_stretching_and_wall_patches():
    m = M.channel_box(8, 12, 6)
    y = np.unique(np.round(m.vertex_coords.reshape(-1, 3)[:, 1], 12))
    assert len(y) == 13 and y[0] == 0.0 and math.isclose(y[-1], 2.0, rel_tol=1e-15)
    j = np.arange(13)
    assert np.allclose(y, 1.0 + np.tanh(2.0 * (2 * j / 12 - 1)) / math.tanh(2.0), atol=1e-12)
    names = m.patch_names
    assert names[2] == "bottom_wall" and names[3] == "top_wall"
    assert math.isclose(m.cell_volume.sum(), 2 * math.pi * 2.0 * math.pi, rel_tol=1e-12)


def test_obstacle_box_removes_cells_and_walls_them():
    """Cells whose centres lie inside the cuboids are removed; the faces between kept and removed
    cells form the obstacle patch; every cell stays closed and the volume is the kept cells'."""
    nx, ny, nz, L, H = 16, 16, 8, 1000.0, 200.0
    obs = M.seeded_obstacles(3, L, L, 20.0, 120.0, seed=5, wmin=0.1, wmax=0.2)
    m = M.Mesh(M.obstacle_box_elements(nx, ny, nz, L, L, H, obs))
    h = np.array([L / nx, L / ny, H / nz])
    ctr = (np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3) + 0.5) * h
    inside = np.zeros(len(ctr), bool)
    for (x0, x1, y0, y1, z0, z1) in obs:
        inside |= ((ctr[:, 0] > x0) & (ctr[:, 0] < x1) & (ctr[:, 1] > y0) & (ctr[:, 1] < y1) & (ctr[:, 2] > z0)
                   & (ctr[:, 2] < z1))
    assert m.n_cells == nx * ny * nz - inside.sum() and inside.sum() > 0
    np.testing.assert_allclose(m.cell_volume, np.prod(h), rtol=1e-13)
    assert np.abs(_cell_sums(m)).max() < 1e-9
    assert m.patch_names[-1] == "obstacles" and m.patch_bc[-1] == M.BC_WALL_MODEL
    # ground faces lost under the obstacles' footprints
    p = m.patch_names.index("z0")
    footprint = inside.reshape(nx, ny, nz)[:, :, 0].sum()
    assert m.patch_nfaces[p] == nx * ny - footprint
    # obstacle face area equals the exposed surface of the removed block set
    rem = inside.reshape(nx, ny, nz)
    exposed = 0.0
    for ax, area in ((0, h[1] * h[2]), (1, h[0] * h[2]), (2, h[0] * h[1])):
        d = np.diff(rem.astype(int), axis=ax)
        exposed += np.abs(d).sum() * area
    # faces on the domain boundary of removed cells are not obstacle faces (top of the domain
    # never reached; the ground under an obstacle is removed with it)
    q = m.patch_names.index("obstacles")
    s = slice(m.n_int_faces + m.patch_start[q], m.n_int_faces + m.patch_start[q] + m.patch_nfaces[q])
    assert m.face_area[s].sum() == pytest.approx(exposed, rel=1e-12)


def test_rcm_renumbering_is_a_permutation_with_a_narrow_band():
    e = M.tet_box_elements(6, jitter=0.15, shuffle_seed=1)
    a, b = M.Mesh(e), M.Mesh(M.rcm_elements(e))
    assert a.n_cells == b.n_cells and a.n_faces == b.n_faces
    # same cells: match by centroid
    ka = np.lexsort(a.cell_centroid.T)
    kb = np.lexsort(b.cell_centroid.T)
    np.testing.assert_array_equal(a.cell_centroid[ka], b.cell_centroid[kb])
    np.testing.assert_array_equal(a.cell_volume[ka], b.cell_volume[kb])

    def band(m):
        return np.abs(m.face_neighbour[:m.n_int_faces].astype(np.int64) - m.face_owner[:m.n_int_faces])
    assert band(b).max() * 10 < band(a).max()
    assert b.patch_names == a.patch_names and (b.patch_nfaces == a.patch_nfaces).all()
