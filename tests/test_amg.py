"""The aggregation-multigrid preconditioner on the CPU: the NumPy statement (oracle/amg.py) pinned
by what the method fixes, and the product's implementation (host-check build of amg_impl.h, the
same code the GPU runs) against it."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle.amg as om
from dopamine_amd import cases
from hostcheck import AMG_DEFAULTS, amg_solve, amg_vcycle, ell_off16


def chain(n, c=1.0, dirichlet=0.0):
    """1-D chain Laplacian as ELL: slot 0 the left neighbour, slot 1 the right one."""
    nbr = -np.ones((2, n), dtype=np.int64)
    coef = np.zeros((2, n))
    nbr[0, 1:] = np.arange(n - 1)
    nbr[1, :-1] = np.arange(1, n)
    coef[nbr >= 0] = c
    diag = (nbr >= 0).sum(axis=0) * c
    diag[0] += dirichlet
    return nbr, coef, diag.astype(float)


def operator(case):
    """The pressure operator of a case as ELL arrays (as the solver assembles it)."""
    import oracle
    mesh = case.make_mesh()
    o = oracle.Solver(mesh, case.params)
    own, nb, skip = o.int_array("ps_owner").copy(), o.int_array("ps_nb").copy(), o.int_array("ps_skip").copy()
    coeff, diag = o.array("ps_coeff").copy(), o.array("ps_diag").copy()
    n = mesh.n_cells
    keep = (skip == 0) & (nb >= 0)
    own, nb, coeff = own[keep], nb[keep], coeff[keep]
    rows, cols, vals = np.concatenate([own, nb]), np.concatenate([nb, own]), np.concatenate([coeff, coeff])
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    cnt = np.bincount(rows, minlength=n)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    slot = np.arange(rows.size) - start[rows]
    nbr = -np.ones((cnt.max(), n), dtype=np.int64)
    cf = np.zeros((cnt.max(), n))
    nbr[slot, rows] = cols
    cf[slot, rows] = vals
    ns = bool(np.all(np.abs(diag - np.bincount(rows, vals, n)) <= 1e-10 * diag))
    return nbr, cf, diag, ns


OPERATORS = {
    "tgv_skew": lambda: cases.regression_cases()["tgv_skew"],   # periodic skewed hexes, singular
    "hydro": lambda: cases.regression_cases()["hydro"],         # outlet rows: non-singular
    "channel": lambda: cases.small_channel(16, 12, 8),          # stretched walls, singular
    "tets": lambda: cases.cavity_tets(6, rcm=True),             # jittered tets, singular
    "full_stack": lambda: cases.full_stack(16, 16, 8, n_obstacles=3, seed=5),  # flat cells, obstacles
}
# the test configuration with the product's aggressive first coarsening on
AMG_PRODUCT = AMG_DEFAULTS[:4] + (0.02,)


# ---- the statement pinned by what the method fixes

def test_aggregation_of_a_chain_known_answer():
    """Equal couplings, cells visited in order. Six cells: cell 0 roots {0, 1}; cell 2 is no root
    (its neighbour 1 is taken); cell 3 roots {2, 3, 4}; cell 5 (neighbour 4 taken) joins
    aggregate 1 in phase 2. Seven cells: the same, then cell 6 roots {5, 6}."""
    nbr, coef, diag = chain(6)
    agg, na = om.aggregate(nbr, coef, diag, 0.08)
    assert na == 2 and list(agg) == [0, 0, 1, 1, 1, 1]
    nbr2, coef2, diag2 = chain(7)
    agg2, na2 = om.aggregate(nbr2, coef2, diag2, 0.08)
    assert na2 == 3 and list(agg2) == [0, 0, 1, 1, 1, 2, 2]


def test_weak_couplings_are_not_aggregated_together():
    """A chain whose middle coupling is 1e-3 of the others: with theta = 0.08 no aggregate spans it."""
    nbr, coef, diag = chain(8)
    coef[1, 3] = coef[0, 4] = 1e-3
    diag = np.zeros(8)
    for k in range(2):
        diag += np.where(nbr[k] >= 0, coef[k], 0.0)
    agg, _ = om.aggregate(nbr, coef, diag, 0.08)
    assert set(agg[:4]).isdisjoint(set(agg[4:]))


@pytest.mark.parametrize("name", ["tgv_skew", "hydro"])
def test_statement_invariants(name):
    nbr, coef, diag, ns = operator(OPERATORS[name]())
    H = om.Hierarchy(nbr, coef, diag)
    A = H.A[0]
    n = A.shape[0]
    # every cell in one aggregate, every aggregate connected
    agg = H.aggs[0]
    assert agg.min() == 0 and agg.max() == H.sizes[1] - 1
    Afine = A.tolil()
    for a in range(min(H.sizes[1], 200)):
        cells = np.flatnonzero(agg == a)
        sub = A[cells][:, cells]
        ncomp, _ = __import__("scipy.sparse.csgraph", fromlist=["x"]).connected_components(sub != 0, directed=False)
        assert ncomp == 1, a
    # Galerkin: the coarse operator is P^T A P and keeps the row-sum (null space) structure
    P = H.P[0]
    Ac = (P.T @ A @ P).toarray()
    assert np.allclose(H.A[1].toarray(), Ac, rtol=0, atol=1e-12 * np.abs(Ac).max())
    assert np.allclose(Ac, Ac.T)
    if ns:
        assert np.abs(Ac.sum(axis=1)).max() <= 1e-10 * np.abs(np.diag(Ac)).max()
    # the cycle is a symmetric positive (semi-)definite linear map
    rng = np.random.default_rng(1)
    u, v = rng.standard_normal(n), rng.standard_normal(n)
    if ns:
        u -= u.mean()
        v -= v.mean()
    Mu, Mv = H.vcycle(u), H.vcycle(v)
    assert abs(Mu @ v - u @ Mv) <= 1e-10 * abs(Mu @ v)
    assert u @ Mu > 0 and v @ Mv > 0


def test_two_level_without_smoothing_is_the_galerkin_projection():
    """omega = 0 leaves M = alpha P A_c^-1 P^T: M A is alpha times an A-orthogonal projection,
    (M A)^2 = alpha M A."""
    nbr, coef, diag = chain(40, dirichlet=1.0)
    H = om.Hierarchy(nbr, coef, diag, omega=0.0, alpha=1.0, coarse_max=1000, theta0=-1.0)
    assert len(H.A) == 2
    A = H.A[0].toarray()
    M = np.column_stack([H.vcycle(e) for e in np.eye(40)])
    MA = M @ A
    assert np.allclose(MA @ MA, MA, atol=1e-10)
    P = H.P[0].toarray()
    assert np.allclose(M, P @ np.linalg.inv(P.T @ A @ P) @ P.T, atol=1e-12)


@pytest.mark.parametrize("name", ["tgv_skew", "hydro"])
def test_statement_fcg_solves_the_system(name):
    nbr, coef, diag, ns = operator(OPERATORS[name]())
    H = om.Hierarchy(nbr, coef, diag)
    A = H.A[0]
    b = np.random.default_rng(2).standard_normal(diag.size)
    if ns:
        b -= b.mean()
    x, its = om.fcg(A, b, H.vcycle, rtol=1e-10, nullspace=ns)
    assert its < 40
    if ns:
        import scipy.sparse as sp
        xd = spla.spsolve((A + sp.coo_matrix(([1.0], ([0], [0])), shape=A.shape)).tocsc(), b)
        xd -= xd.mean()
        x -= x.mean()
    else:
        xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)


# ---- the product (host-check build of the same code the GPU runs) against the statement

@pytest.mark.parametrize("name", list(OPERATORS))
def test_product_hierarchy_and_cycle_match_the_statement(name):
    nbr, coef, diag, ns = operator(OPERATORS[name]())
    H = om.Hierarchy(nbr, coef, diag, *AMG_DEFAULTS)
    r = np.random.default_rng(3).standard_normal(diag.size)
    z, agg, sizes = amg_vcycle(nbr, coef, diag, r)
    assert sizes == H.sizes
    assert np.array_equal(agg, H.aggs[0])
    zo = H.vcycle(r)
    err = np.linalg.norm(z - zo) / np.linalg.norm(zo)
    # the product's cycle runs in FP32 with its couplings stored as binary16 (relative rounding
    # 2^-12 on average, 2^-11 at most) against the float64 statement
    assert err < 5e-4, err


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("name", list(OPERATORS))
def test_product_multigrid_cg_reaches_the_fp64_tolerance(name, mixed):
    nbr, coef, diag, ns = operator(OPERATORS[name]())
    A = om.ell_to_csr(nbr, coef, diag)
    b = np.random.default_rng(4).standard_normal(diag.size)
    if ns:
        b -= b.mean()
    rc, x, its, res, sizes, sweeps = amg_solve(nbr, coef, diag, b, 1e-8, nullspace=ns, mixed=mixed)
    H = om.Hierarchy(nbr, coef, diag, *AMG_DEFAULTS)
    _, its_o = om.fcg(A, b, H.vcycle, rtol=1e-8, nullspace=ns)
    assert rc == 0
    assert res == pytest.approx(np.linalg.norm(b - A @ x), rel=1e-6)
    assert res <= 1e-8 * np.linalg.norm(b)
    assert its <= its_o + 3 + (4 if mixed else 0)  # FP32 sweeps restart the Krylov space once
    assert (sweeps > 0) == mixed


def test_aggressive_first_coarsening_on_flat_cells():
    """The full-stack cells are 2.5 times wider than high: theta = 0.08 keeps only the vertical
    couplings strong and the first aggregates are short vertical lines (under 3.5 cells each), so
    the first level is aggregated again with theta0 = 0.02 (compact aggregates) and the cycle
    takes alpha_aggr = 2.4. Product and statement agree on that branch, and the flexible CG
    converges."""
    nbr, coef, diag, ns = operator(OPERATORS["full_stack"]())
    line, na_line = om.aggregate(nbr, coef, diag, 0.08)
    assert diag.size < 3.5 * na_line
    H = om.Hierarchy(nbr, coef, diag, *AMG_PRODUCT)
    assert H.aggressive and H.alpha == 2.4 and H.sizes[1] < diag.size / 3.5
    r = np.random.default_rng(5).standard_normal(diag.size)
    z, agg, sizes = amg_vcycle(nbr, coef, diag, r, cfg=AMG_PRODUCT)
    assert sizes == H.sizes and np.array_equal(agg, line)  # agg_out is the theta = 0.08 aggregation
    zo = H.vcycle(r)
    assert np.linalg.norm(z - zo) <= 5e-4 * np.linalg.norm(zo)  # FP32 cycle, binary16 couplings
    b = r - r.mean()
    rc, x, its, res, sizes2, _ = amg_solve(nbr, coef, diag, b, 1e-8, nullspace=ns, mixed=True, cfg=AMG_PRODUCT)
    assert rc == 0 and res <= 1e-8 * np.linalg.norm(b) and sizes2 == H.sizes


def test_channel_keeps_the_strength_threshold():
    """The stretched channel's first aggregation already coarsens by more than 3.5: no second pass."""
    nbr, coef, diag, ns = operator(cases.small_channel(32, 24, 16))
    H = om.Hierarchy(nbr, coef, diag, *AMG_PRODUCT)
    assert not H.aggressive and H.alpha == 1.8


def test_binary16_coupling_storage_matches_ieee_half():
    """The couplings' binary16 storage rounds to nearest even as IEEE 754 half precision does
    (numpy float16): normals, subnormals, ties, overflow to infinity."""
    from hostcheck import coef_roundtrip
    rng = np.random.default_rng(16)
    v = np.concatenate([rng.random(20000), rng.random(2000) * 1e-5, rng.random(2000) * 1e-7,
                        -rng.random(1000), np.array([0.0, -0.0, 1.0, 65504.0, 65519.0, 65520.0, 1e6, 2.0 ** -24,
                                                     2.0 ** -25, 3 * 2.0 ** -26, 1.0 + 2.0 ** -11,
                                                     1.0 + 3 * 2.0 ** -11])]).astype(np.float32)
    h, back = coef_roundtrip(v)
    with np.errstate(over="ignore"):  # 65520 and 1e6 overflow to infinity on purpose
        want = v.astype(np.float16)
    assert np.array_equal(h, want.view(np.uint16))
    assert np.array_equal(back, want.astype(np.float32))


# ---- the 16-bit neighbour offsets of the level-0 passes (dm::ell_off16)

def box_ell(nx, ny, nz, seed=0):
    """7-point box Laplacian neighbour lists in natural order, each cell's slots shuffled."""
    n = nx * ny * nz
    c = np.arange(n)
    i, j, k = c % nx, (c // nx) % ny, c // (nx * ny)
    rows = []
    for ok, off in ((i > 0, -1), (i < nx - 1, 1), (j > 0, -nx), (j < ny - 1, nx), (k > 0, -nx * ny),
                    (k < nz - 1, nx * ny)):
        rows.append(np.where(ok, c + off, -1))
    nbr = np.array(rows)
    rng = np.random.default_rng(seed)
    for cc in range(n):
        nbr[:, cc] = nbr[rng.permutation(6), cc]
    return nbr


def decode(nbr, W, base, d, src):
    n = nbr.shape[1]
    j = np.arange(n)[None, :] + base[:, None] + d.astype(np.int64)
    live = src >= 0
    assert np.array_equal(j[~live], np.broadcast_to(np.arange(n), (W, n))[~live])  # empty slot: the cell
    return j, live


def test_off16_box_slots_are_seated_by_direction():
    nx, ny, nz = 9, 7, 5
    nbr = box_ell(nx, ny, nz)
    W, base, d, src = ell_off16(nbr)
    assert W == 6
    j, live = decode(nbr, W, base, d, src)
    c = 4 + nx * (3 + ny * 2)  # an interior cell
    assert list(j[:, c] - c) == [-nx * ny, -nx, -1, 1, nx, nx * ny]
    c0 = 0  # the corner: its three (positive) neighbours in the upper slots, the lower ones empty
    assert list(live[:, c0]) == [False] * 3 + [True] * 3
    assert list(j[3:, c0]) == [1, nx, nx * ny]


@pytest.mark.parametrize("name", list(OPERATORS))
def test_off16_rows_hold_the_same_neighbours(name):
    nbr = operator(OPERATORS[name]())[0]
    W, base, d, src = ell_off16(nbr)
    if W == 0:
        pytest.skip("offsets do not fit 16 bits on this operator (the int32 lists are used)")
    j, live = decode(nbr, W, base, d, src)
    n = nbr.shape[1]
    for c in range(n):
        s = src[live[:, c], c]
        assert sorted(s) == [k for k in range(nbr.shape[0]) if nbr[k, c] >= 0]
        assert np.array_equal(j[live[:, c], c], nbr[s, c])
        off = (j[:, c] - c)[live[:, c]]
        assert np.all(np.diff(off) > 0)  # the slots keep the offsets in ascending order


def test_off16_falls_back_when_offsets_do_not_fit():
    def periodic_chain(n):
        nbr, _, _ = chain(n)
        nbr[0, 0], nbr[1, n - 1] = n - 1, 0
        return nbr
    assert ell_off16(periodic_chain(60_000))[0] == 4
    assert ell_off16(periodic_chain(70_000))[0] == 0  # the wrap spans more than 65,536 offsets
    nbr = -np.ones((9, 10), dtype=np.int64)  # nine slots: wider than the compact form's eight
    nbr[:, 9] = np.arange(9)
    nbr[0, :9] = 9
    assert ell_off16(nbr)[0] == 0


@pytest.mark.parametrize("name", ["full_stack", "channel"])
def test_off16_solve_matches_the_int32_lists(name, monkeypatch):
    """The compact and the int32 forms differ only in the order of each row's terms."""
    nbr, coef, diag, ns = operator(OPERATORS[name]())
    assert ell_off16(nbr)[0] > 0
    b = np.random.default_rng(4).standard_normal(diag.size)
    if ns:
        b -= b.mean()
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("DOPAMINE_ELL_OFF16", flag)
        out[flag] = amg_solve(nbr, coef, diag, b, 1e-8, nullspace=ns, mixed=True)
    (rc1, x1, it1, r1, *_), (rc0, x0, it0, r0, *_) = out["1"], out["0"]
    assert rc1 == rc0 == 0 and abs(it1 - it0) <= 1
    assert np.linalg.norm(x1 - x0) <= 1e-6 * np.linalg.norm(x0)
