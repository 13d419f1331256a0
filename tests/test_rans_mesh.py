"""C-grid meshes for the RANS solver (aero2d_amd.rans_mesh: cgrid,
refine_cwall, cgrid_mesh), CPU only: closed forms of the stretching, the
wall and wake spacings the generator promises, and the invariants of the
mesh the residual reads (areas add up to the domain, every boundary face
labelled, the wake cut joined by interior faces)."""
import numpy as np
import pytest

from aero2d_amd import naca4
from aero2d_amd.rans_mesh import _two_sided, cgrid, cgrid_mesh, refine_cwall


def _shoelace(p):
    x, y = p[:, 0], p[:, 1]
    return 0.5 * abs(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y))


@pytest.fixture(scope="module")
def grid():
    return cgrid(naca4("0012", 101), n_side=17, nj=11, ds_le=1e-2, ds_te=1e-2, ds0=1e-2, radius=6.0,
                 wake_ratio=1.3)


def test_two_sided_stretching_end_steps():
    # the end steps are d0 and d1 (Vinokur targets corrected by fixed point);
    # symmetric ends give a symmetric distribution; steps grow monotonically
    # from the smaller end step towards the middle
    for n, d0, d1 in ((97, 2e-3, 5e-3), (200, 1e-3, 1e-3), (60, 4e-3, 1e-3)):
        s = _two_sided(n, d0, d1)
        h = np.diff(s)
        assert s[0] == 0.0 and abs(s[-1] - 1.0) < 1e-14 and np.all(h > 0)
        assert abs(h[0] / d0 - 1) < 1e-10 and abs(h[-1] / d1 - 1) < 1e-10
        k = int(np.argmax(h))
        assert np.all(np.diff(h[:k + 1]) > -1e-15) and np.all(np.diff(h[k:]) < 1e-15)
    s = _two_sided(100, 2e-3, 2e-3)
    assert np.allclose(s + s[::-1], 1.0, atol=1e-13)
    with pytest.raises(ValueError):
        _two_sided(10, 0.2, 0.2)


def test_cgrid_wall_wake_and_ends(grid):
    x, y, K = grid.x, grid.y, grid.n_wake
    ni, nj = x.shape
    assert ni == 2 * K + 2 * 17 + 1
    # the wall row is the contour, its ends the trailing edge; the cut rows coincide
    assert np.allclose([x[K, 0], y[K, 0], x[ni - 1 - K, 0], y[ni - 1 - K, 0]], [1.0, 0.0, 1.0, 0.0], atol=1e-12)
    assert np.allclose(x[:K + 1, 0], x[ni - 1 - np.arange(K + 1), 0]) and np.all(y[:K, 0] == 0.0)
    assert abs(x[0, 0] - 7.0) < 1e-9                  # the cut ends at x_te + radius
    # the two cut ends march straight up and down
    assert np.allclose(x[0], x[0, 0]) and np.allclose(x[-1], x[-1, 0])
    assert np.all(np.diff(y[0]) > 0) and np.all(np.diff(y[-1]) < 0)
    assert abs(y[0, -1] - 6.0) < 1e-9 and abs(y[-1, -1] + 6.0) < 1e-9
    assert grid.cell_areas().min() > 0
    # leading- and trailing-edge steps of the re-panelled contour: arc-length
    # steps of 1e-2, whose chords are shorter by s^2/(24 R^2) (R the radius of
    # curvature: ~0.016 at the leading edge of NACA 0012, large at the trailing edge)
    w = np.hypot(np.diff(x[K:ni - K, 0]), np.diff(y[K:ni - K, 0]))
    assert abs(w[0] / 1e-2 - 1) < 1e-5 and abs(w[-1] / 1e-2 - 1) < 1e-5
    assert 0.97 < w[16] / 1e-2 < 1.0 and 0.97 < w[17] / 1e-2 < 1.0
    # the cut's first step is the trailing-edge step, then grows geometrically
    c = np.diff(x[ni - 1 - K:, 0])
    q = c[1:] / c[:-1]
    assert abs(c[0] / 1e-2 - 1) < 1e-12 and np.allclose(q, q[0]) and 1.2 < q[0] < 1.4
    with pytest.raises(ValueError):
        cgrid(naca4("0012", 33, closed_te=False))


def test_refine_cwall_first_heights(grid):
    ds, g = 1e-5, 1.15
    r = refine_cwall(grid, ds, layers=1, ratio=1.2, wake_growth=g)
    K, ni = grid.n_wake, grid.x.shape[0]
    h = np.hypot(r.x[:, 1] - r.x[:, 0], r.y[:, 1] - r.y[:, 0])
    H = np.hypot(grid.x[:, 1] - grid.x[:, 0], grid.y[:, 1] - grid.y[:, 0])
    n = r.x.shape[1] - grid.x.shape[1] + 1
    col = np.zeros(ni)
    col[:K] = np.arange(K, 0, -1)
    col[ni - K:] = np.arange(1, K + 1)
    want = np.minimum(ds * g ** col, H / n)
    assert np.allclose(h[K:ni - K], ds, rtol=1e-9)          # the wall: ds everywhere
    assert np.allclose(h, want, rtol=1e-6)                  # the cut: grows by g per column, capped
    # the layers beyond the re-sampled one are the marched ones
    assert np.array_equal(r.x[:, n + 1:], grid.x[:, 2:]) and np.allclose(r.x[:, n], grid.x[:, 1], atol=1e-14)
    assert r.cell_areas().min() > 0
    with pytest.raises(ValueError):
        refine_cwall(grid, 1.0)


@pytest.mark.parametrize("quad_layers", [None, 4])
def test_cgrid_mesh_invariants(grid, quad_layers):
    m = cgrid_mesh(grid, quad_layers=quad_layers)
    x, y, K = grid.x, grid.y, grid.n_wake
    ni, nj = x.shape
    nq = nj - 1 if quad_layers is None else quad_layers
    assert m.n_cells == (ni - 1) * nq + 2 * (ni - 1) * (nj - 1 - nq)
    # coincident cut nodes are one node
    assert len(m.nodes) == ni * nj - (K + 1)
    tag = np.bincount(m.face_tag, minlength=3)
    assert tag[1] == 2 * 17 and tag[2] == (ni - 1) + 2 * (nj - 1)
    # areas: the region inside the outer loop (the upper end line up, the outer
    # row, the lower end line down; the cut is interior) less the airfoil
    loop = np.concatenate([np.column_stack([x[0], y[0]]), np.column_stack([x[1:, -1], y[1:, -1]]),
                           np.column_stack([x[-1, -2:0:-1], y[-1, -2:0:-1]])])
    body = np.column_stack([x[K:ni - K, 0], y[K:ni - K, 0]])
    assert abs(m.volume.sum() - (_shoelace(loop) - _shoelace(body))) < 1e-9 * _shoelace(loop)
    # the wake cut: the first cells above and below each cut column share an interior face
    fc = {tuple(sorted(p)) for p in m.face_cells[m.face_tag == 0].tolist()}
    for i in range(K):
        if quad_layers is None:
            a, b = i, ni - 2 - i
            assert (min(a, b), max(a, b)) in fc
    # faces' normals point from the left cell to the right one
    it = m.face_tag == 0
    d = m.centroid[m.face_cells[it, 1]] - m.centroid[m.face_cells[it, 0]]
    assert np.all(np.sum(d * m.face_normal[it], 1) > 0)
    # closed cells: the outward face normals of every cell sum to zero
    acc = np.zeros((m.n_cells, 2))
    nA = m.face_normal * m.face_area[:, None]
    np.add.at(acc, m.face_cells[:, 0], nA)
    right = m.face_cells[:, 1] >= 0
    np.add.at(acc, m.face_cells[right, 1], -nA[right])
    assert np.abs(acc).max() < 1e-12
    assert m.wall_dist.min() > 0 and np.all(m.wall_dist <= 7.5)


def test_transfer_state_between_meshes(grid):
    """transfer_state: the identity on the same mesh; onto a mesh with a finer
    wall layer the nearest-cell values keep density and pressure, the
    velocity and rho chi shrink by the wall-distance ratio, and a uniform
    state away from walls carries over exactly (the reference's test pattern,
    tests/test_rans.py:128-137: finite, positive density, the new mesh's size)."""
    from aero2d_amd.rans import transfer_state
    GAM = 1.4
    coarse = cgrid_mesh(grid)
    fine = cgrid_mesh(refine_cwall(grid, 1e-4))
    rng = np.random.default_rng(3)
    w = np.column_stack([1 + 0.1 * rng.random(coarse.n_cells), 0.5 + 0.1 * rng.random(coarse.n_cells),
                         0.1 * rng.random(coarse.n_cells), 1 / GAM + 0.05 * rng.random(coarse.n_cells),
                         3 * rng.random(coarse.n_cells)])
    Q = np.column_stack([w[:, 0], w[:, 0] * w[:, 1], w[:, 0] * w[:, 2],
                         w[:, 3] / (GAM - 1) + 0.5 * w[:, 0] * (w[:, 1] ** 2 + w[:, 2] ** 2), w[:, 0] * w[:, 4]])
    assert np.array_equal(transfer_state(coarse, Q, coarse), Q)
    Qf = transfer_state(coarse, Q, fine)
    assert Qf.shape == (fine.n_cells, 5) and np.all(np.isfinite(Qf)) and np.all(Qf[:, 0] > 0)
    p = lambda q: (GAM - 1) * (q[:, 3] - 0.5 * (q[:, 1] ** 2 + q[:, 2] ** 2) / q[:, 0])   # noqa: E731
    from scipy.spatial import cKDTree
    _, near = cKDTree(coarse.centroid).query(fine.centroid)
    assert np.allclose(p(Qf), p(Q)[near], rtol=1e-12, atol=1e-14)
    assert np.array_equal(Qf[:, 0], Q[near, 0])
    s = np.clip(fine.wall_dist / coarse.wall_dist[near], 0, 1)
    assert np.allclose(Qf[:, 1], Q[near, 1] * s) and np.allclose(Qf[:, 4], Q[near, 4] * s)
    assert (s < 1).any()
