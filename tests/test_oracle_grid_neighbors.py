"""Oracle pins for the grid (GridHash) and the neighbor search (NeighborQuery).

Checked against closed forms on hand-picked points and an O(N^2) brute force,
not against a retyped copy of the oracle's grid logic."""
import numpy as np
import pytest


def brute_sets(x, h):
    """All pairs with sqrtf(((dx*dx + dy*dy) + dz*dz)) < h in float32 (the
    reference's norm, glm func_geometric.inl:64-71,95-100)."""
    x = np.asarray(x, np.float32)
    d = x[:, None, :] - x[None, :, :]
    s = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    r = np.sqrt(s)
    return [np.flatnonzero(r[i] < np.float32(h)) for i in range(len(x))]


def csr_rows(off, el):
    return [np.sort(el[off[i]:off[i + 1]]) for i in range(len(off) - 1)]


def test_default_grid(oracle):
    p = oracle.make_params()
    ds, lo, dims = oracle.grid(p)
    assert dims == (85, 85, 85)  # ceil(1/0.0125) + 1 + 2*2 (GridHash.cu:57-59)
    assert np.float32(ds) == np.float32(0.0125)
    np.testing.assert_allclose(lo, [-0.025] * 3, rtol=0, atol=1e-9)


def test_cell_centres_map_to_their_cell(oracle):
    p = oracle.make_params()
    ds, lo, dims = oracle.grid(p)
    k = np.arange(dims[0])
    centres = (np.float64(lo[0]) + (k + 0.5) * 0.0125).astype(np.float32)
    x = np.stack([centres, centres[::-1], np.full_like(centres, centres[7])], 1)
    c3, cid = oracle.cell3(p, x)
    assert (c3[:, 0] == k).all() and (c3[:, 1] == k[::-1]).all() and (c3[:, 2] == 7).all()
    np.testing.assert_array_equal(cid, c3[:, 0] + 85 * (c3[:, 1] + 85 * c3[:, 2]))


def test_out_of_grid_cells(oracle):
    p = oracle.make_params()
    # grid spans [-0.025, -0.025 + 85*0.0125 = 1.0375)
    x = np.array([[-0.026, 0.5, 0.5], [1.04, 0.5, 0.5], [0.5, 1e6, 0.5], [0.5, 0.5, -1e6],
                  [np.nan, 0.5, 0.5], [-0.024, 1.037, 0.5]], np.float32)
    c3, cid = oracle.cell3(p, x)
    assert c3[0, 0] == -1 and cid[0] == -1
    assert c3[1, 0] == 85 and cid[1] == -1
    assert c3[2, 1] == 86 and cid[2] == -1  # clamped to n+1: no valid stencil cell
    assert c3[3, 2] == -2 and cid[3] == -1
    assert c3[4, 0] == -2 and cid[4] == -1
    assert c3[5, 0] == 0 and c3[5, 1] == 84 and cid[5] >= 0


@pytest.mark.parametrize("seed,n,width", [(1, 300, 0.05), (2, 500, 0.03), (3, 200, 0.1), (4, 400, 0.02)])
def test_neighbor_sets_equal_brute_force(oracle, scene, seed, n, width):
    p = oracle.make_params()
    u = scene.uniform01(seed, 3 * n).reshape(n, 3)
    x = (0.4 + width * u).astype(np.float32)
    off, el = oracle.neighbor_query(p, x)
    got = csr_rows(off, el)
    want = brute_sets(x, 0.0125)
    for i in range(n):
        np.testing.assert_array_equal(got[i], want[i])


def test_neighbor_sets_near_grid_faces(oracle, scene):
    """Particles within one cell of the padded grid's faces still see every
    in-grid neighbor."""
    p = oracle.make_params()
    n = 400
    u = scene.uniform01(9, 3 * n).reshape(n, 3)
    x = (-0.024 + 0.04 * u).astype(np.float32)  # corner region of the grid
    x[n // 2:] = (1.0375 - 0.04 + 0.0399 * u[n // 2:]).astype(np.float32)
    off, el = oracle.neighbor_query(p, x)
    got = csr_rows(off, el)
    want = brute_sets(x, 0.0125)
    for i in range(n):
        np.testing.assert_array_equal(got[i], want[i])


def test_pair_three_half_cells_apart(oracle):
    """x_i = 0.106249996 and x_j = 0.11874999 are closer than h (the strict test
    passes in float32) and lie in adjacent reference cells 10 and 11, although
    2q puts them 3 half cells apart (20.999996 and 23.0): the reference lists
    the pair, as the brute force does."""
    p = oracle.make_params()
    x = np.array([[0.106249996, 0.5, 0.5], [0.11874999, 0.5, 0.5]], dtype=np.float32)
    lo, ds = np.float32(-0.025), np.float32(0.0125)
    q = ((x[:, 0] - lo).astype(np.float32) / ds).astype(np.float32)
    assert np.floor(2 * q).tolist() == [20.0, 23.0] and np.floor(q).tolist() == [10.0, 11.0]
    off, el = oracle.neighbor_query(p, x)
    got, want = csr_rows(off, el), brute_sets(x, 0.0125)
    for g, w in zip(got, want):
        assert g.tolist() == w.tolist() == [0, 1]


def test_symmetry_and_self(oracle, scene):
    p = oracle.make_params()
    x = scene.jittered_lattice((10, 9, 8), (0.3, 0.3, 0.3), 0.005, 0.1, seed=5)
    off, el = oracle.neighbor_query(p, x)
    rows = csr_rows(off, el)
    for i, r in enumerate(rows):
        assert i in r
        for j in r:
            assert i in rows[j]


def test_sets_independent_of_grid_placement(oracle, scene):
    """Shifting the scene bounds moves every cell boundary; the sets (a
    property of distances only) must not change."""
    x = scene.jittered_lattice((9, 9, 9), (0.4, 0.4, 0.4), 0.005, 0.1, seed=6)
    a = csr_rows(*oracle.neighbor_query(oracle.make_params(), x))
    lo = (-0.0031, 0.0017, -0.0043)
    hi = tuple(v + 1.0 for v in lo)
    b = csr_rows(*oracle.neighbor_query(oracle.make_params(lo=lo, hi=hi), x))
    for r1, r2 in zip(a, b):
        np.testing.assert_array_equal(r1, r2)


def test_strict_cutoff_one_ulp(oracle):
    """|x_i - x_j| == h is not a neighbor (strict <, NeighborQuery.cu:267);
    one ulp further in h it is."""
    x = np.array([[0.5, 0.5, 0.5], [0.5, 0.5, 0.5]], np.float32)
    x[1, 0] = np.float32(0.5) + np.float32(0.0123)
    d = np.float32(np.sqrt(np.float32((x[0, 0] - x[1, 0]) ** 2)))
    h_eq = float(d)
    h_up = float(np.nextafter(d, np.float32(1)))
    off, el = oracle.neighbor_query(oracle.make_params(h=h_eq), x)
    assert list(np.diff(off)) == [1, 1]
    off, el = oracle.neighbor_query(oracle.make_params(h=h_up), x)
    assert list(np.diff(off)) == [2, 2]


def test_out_of_grid_particle_queries_but_is_not_listed(oracle):
    """GridHash::construct drops particles outside the grid (GridHash.cu:85,194)
    but they still query: the lists become one-sided."""
    p = oracle.make_params()
    x = np.array([[-0.0245, 0.5, 0.5], [-0.0255, 0.5, 0.5]], np.float32)  # in grid / just outside
    off, el = oracle.neighbor_query(p, x)
    rows = csr_rows(off, el)
    assert list(rows[0]) == [0]
    assert list(rows[1]) == [0]
