"""Unstructured hybrid mesh for the RANS solver (host side, NumPy FP64).

`Mesh` carries the arrays the RANS residual reads: nodes, cells (counter-
clockwise vertex lists, triangles and quadrilaterals), cell centroids and
areas, faces with their left/right cells, unit normals (left -> right), lengths
and midpoints, boundary tags (0 interior, 1 wall, 2 far field), the body of a
wall face and each cell's distance to the nearest wall (aero2d/rans_mesh.py:19-36).

`connectivity` builds a Mesh from nodes, cells and labelled boundary edges
(aero2d/rans_mesh.py:283-337). The reference meshes with gmsh, which this
image does not have. `cgrid` marches a C-grid with a wake cut (the topology of
the reference's wall block, aero2d/cblock.py) and `cgrid_mesh` turns it into a
Mesh, the cut's coincident nodes merged; `ogrid_mesh` does the same for a
marched O-grid (aero2d_amd.grid). Both can keep quadrilaterals in the wall
layers and split each outer quadrilateral into two triangles. On O-grids the
cells round a sharp trailing edge form a fan of skewed cells and the SA solves
do not converge; on C-grids they do (NACA 0012, M 0.15, Re 6e6).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["Mesh", "connectivity", "ogrid_mesh", "refine_wall", "CGrid", "cgrid", "refine_cwall", "cgrid_mesh"]


@dataclass
class Mesh:
    nodes: np.ndarray          # (Nn, 2)
    cells: list                # vertex lists, counter-clockwise
    centroid: np.ndarray       # (Nc, 2)
    volume: np.ndarray         # (Nc,)
    face_cells: np.ndarray     # (Nf, 2) left, right cell (-1 on a boundary)
    face_nodes: np.ndarray     # (Nf, 2)
    face_normal: np.ndarray    # (Nf, 2) unit normal from left to right
    face_area: np.ndarray      # (Nf,)
    face_center: np.ndarray    # (Nf, 2)
    face_tag: np.ndarray       # (Nf,) 0 interior, 1 wall, 2 far field
    face_body: np.ndarray      # (Nf,) body of a wall face, -1 otherwise
    wall_dist: np.ndarray      # (Nc,) distance to the nearest wall

    @property
    def n_cells(self) -> int:
        return len(self.volume)


def connectivity(nodes, cells, wall_edges: dict, far_edges: set) -> Mesh:
    """Mesh arrays from nodes, cells and boundary edges (aero2d/rans_mesh.py:283-337).

    Each cell is turned counter-clockwise; its centroid is the polygon's centre
    of mass. Faces are numbered in the order their edges are first met walking
    the cells; the first cell met is the face's left cell and the face's nodes
    follow that cell's counter-clockwise walk, so the normal (t_y, -t_x)/|t|
    points out of it. A boundary edge must be in `wall_edges` (key: sorted
    node pair, value: body) or `far_edges`, else RuntimeError. The wall
    distance is the distance from each centroid to the nearest of nine points
    spaced evenly along every wall face."""
    from scipy.spatial import cKDTree

    nodes = np.asarray(nodes, dtype=np.float64)
    ccw, cen, vol = [], np.zeros((len(cells), 2)), np.zeros(len(cells))
    for k, c in enumerate(cells):
        c = list(c)
        p = nodes[c]
        px, py = p[:, 0], p[:, 1]
        qx, qy = np.roll(px, -1), np.roll(py, -1)
        cr = px * qy - qx * py
        a = 0.5 * cr.sum()
        if a < 0:
            c, p = c[::-1], p[::-1]
            px, py = p[:, 0], p[:, 1]
            qx, qy = np.roll(px, -1), np.roll(py, -1)
            cr = px * qy - qx * py
            a = -a
        ccw.append(c)
        cen[k] = [np.sum((px + qx) * cr) / (6 * a), np.sum((py + qy) * cr) / (6 * a)]
        vol[k] = a
    # every (cell, edge) pair, in walking order
    owner, ea, eb = [], [], []
    for k, c in enumerate(ccw):
        n = len(c)
        for i in range(n):
            owner.append(k)
            ea.append(c[i])
            eb.append(c[(i + 1) % n])
    owner, ea, eb = np.asarray(owner), np.asarray(ea), np.asarray(eb)
    nn = len(nodes)
    key = np.minimum(ea, eb).astype(np.int64) * nn + np.maximum(ea, eb)
    _, first, inv, cnt = np.unique(key, return_index=True, return_inverse=True, return_counts=True)
    if cnt.max() > 2:
        raise RuntimeError("an edge is shared by more than two cells")
    order = np.argsort(first)                     # faces in first-met order
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    nf = len(order)
    fc = -np.ones((nf, 2), dtype=np.int64)
    fn = np.zeros((nf, 2), dtype=np.int64)
    fidx = rank[inv]                              # face of every (cell, edge) pair
    firstpair = first[order]
    fc[:, 0] = owner[firstpair]
    fn[:, 0], fn[:, 1] = ea[firstpair], eb[firstpair]
    second = np.ones(len(owner), bool)
    second[firstpair] = False
    fc[fidx[second], 1] = owner[second]
    tag = np.zeros(nf, dtype=np.int64)
    body = -np.ones(nf, dtype=np.int64)
    for f in np.where(fc[:, 1] < 0)[0]:
        kk = (int(min(fn[f])), int(max(fn[f])))
        if kk in wall_edges:
            tag[f], body[f] = 1, wall_edges[kk]
        elif kk in far_edges:
            tag[f] = 2
        else:
            raise RuntimeError("boundary edge without a label")
    pa, pb = nodes[fn[:, 0]], nodes[fn[:, 1]]
    t = pb - pa
    area = np.hypot(t[:, 0], t[:, 1])
    normal = np.column_stack([t[:, 1], -t[:, 0]]) / area[:, None]
    fcen = 0.5 * (pa + pb)
    wf = np.where(tag == 1)[0]
    sub = np.linspace(0.0, 1.0, 9)
    wp = (pa[wf, None, :] + sub[None, :, None] * t[wf, None, :]).reshape(-1, 2)
    d, _ = cKDTree(wp).query(cen)
    return Mesh(nodes, ccw, cen, vol, fc, fn, normal, area, fcen, tag, body, d)


def refine_wall(grid, ds_wall: float, layers: int = 1, ratio: float = 1.2):
    """O-grid whose first `layers` wall layers are re-sampled so the first cell
    is ds_wall high (y+ ~ 1 meshes, which the marching generator cannot start
    with directly). Each wall-normal node line is cut at the same fractions of
    its arc length up to node `layers`: a geometric sequence from ds_wall / H
    (H the mean arc length to that node) whose count gives a ratio close to
    `ratio`. Returns an OGrid with the added node layers."""
    from .grid import OGrid, growth_ratio
    x, y = np.asarray(grid.x, float), np.asarray(grid.y, float)
    seg = np.hypot(np.diff(x, axis=1), np.diff(y, axis=1))          # (ni, nj-1)
    s = np.concatenate([np.zeros((x.shape[0], 1)), np.cumsum(seg, 1)], 1)
    H = float(s[:, layers].mean())
    if not 0 < ds_wall < H:
        raise ValueError("ds_wall must be positive and below the height of the re-sampled layers")
    n = max(layers, int(np.ceil(np.log(1 + H * (ratio - 1) / ds_wall) / np.log(ratio))))
    q = growth_ratio(ds_wall / H, 1.0, n)
    t = np.concatenate([[0.0], np.cumsum((ds_wall / H) * q ** np.arange(n))])
    t /= t[-1]
    nx = np.empty((x.shape[0], n + 1))
    ny = np.empty_like(nx)
    for i in range(x.shape[0]):
        st = t * s[i, layers]
        nx[i] = np.interp(st, s[i, :layers + 1], x[i, :layers + 1])
        ny[i] = np.interp(st, s[i, :layers + 1], y[i, :layers + 1])
    X = np.concatenate([nx, x[:, layers + 1:]], 1)
    Y = np.concatenate([ny, y[:, layers + 1:]], 1)
    g = OGrid(X, Y)
    if g.cell_areas().min() <= 0:
        raise RuntimeError("refined O-grid has inverted cells")
    return g


def ogrid_mesh(grid, quad_layers: int | None = None) -> Mesh:
    """Hybrid mesh from an O-grid's nodes (grid.x, grid.y of shape (ni, nj)).

    Cells of wall layers j < quad_layers stay quadrilaterals; every cell beyond
    is cut into two triangles along alternating diagonals. Node (i, j) is number
    j*ni + i; cells are numbered layer by layer from the wall, i fastest, so
    cells close in memory are close on the mesh. The wall is body 0, the
    outermost node ring the far field. quad_layers=None keeps every cell a
    quadrilateral."""
    x, y = np.asarray(grid.x, np.float64), np.asarray(grid.y, np.float64)
    ni, nj = x.shape
    nodes = np.column_stack([x.T.ravel(), y.T.ravel()])
    nq = nj - 1 if quad_layers is None else int(quad_layers)
    nid = lambda i, j: j * ni + (i % ni)          # noqa: E731
    cells = []
    for j in range(nj - 1):
        for i in range(ni):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            if j < nq:
                cells.append([a, b, c, d])
            elif (i + j) % 2 == 0:
                cells += [[a, b, c], [a, c, d]]
            else:
                cells += [[a, b, d], [b, c, d]]
    wall = {(min(nid(i, 0), nid(i + 1, 0)), max(nid(i, 0), nid(i + 1, 0))): 0 for i in range(ni)}
    far = {(min(nid(i, nj - 1), nid(i + 1, nj - 1)), max(nid(i, nj - 1), nid(i + 1, nj - 1)))
           for i in range(ni)}
    return connectivity(nodes, cells, wall, far)


@dataclass
class CGrid:
    """Structured C-grid: node arrays (ni, nj); i runs along the C from the
    downstream end of the upper wake cut, over the airfoil (trailing edge, upper
    surface, leading edge, lower surface, trailing edge) to the downstream end
    of the lower wake cut; j runs outward. Nodes (i, 0) and (ni-1-i, 0) for
    i <= n_wake coincide (the wake cut and the trailing edge)."""
    x: np.ndarray
    y: np.ndarray
    n_wake: int

    def cell_areas(self) -> np.ndarray:
        """Signed areas (ni-1, nj-1) of the cells (i..i+1, j..j+1); positive when
        the C runs clockwise round the airfoil (the contour order of naca4)."""
        x, y = self.x, self.y
        d1x, d1y = x[1:, 1:] - x[:-1, :-1], y[1:, 1:] - y[:-1, :-1]
        d2x, d2y = x[:-1, 1:] - x[1:, :-1], y[:-1, 1:] - y[1:, :-1]
        return -0.5 * (d1x * d2y - d1y * d2x)


def _vinokur(n: int, d0: float, d1: float) -> np.ndarray:
    B = 1.0 / (n * np.sqrt(d0 * d1))
    if B <= 1.0:
        raise ValueError("two-sided stretching needs steps below the even division")
    lo, hi = 1e-9, 50.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if np.sinh(mid) / mid < B else (lo, mid)
    delta = 0.5 * (lo + hi)
    xi = np.arange(n + 1) / n
    u = 0.5 * (1.0 + np.tanh(delta * (xi - 0.5)) / np.tanh(0.5 * delta))
    A = np.sqrt(d1 / d0)
    return u / (A + (1.0 - A) * u)


def _two_sided(n: int, d0: float, d1: float) -> np.ndarray:
    """n+1 points on [0, 1] with first step d0 and last step d1 (fractions of
    the length). Vinokur's two-sided tanh stretching (J. Comput. Phys. 50,
    1983): u = (1 + tanh(delta (xi - 1/2)) / tanh(delta / 2)) / 2 with
    sinh(delta)/delta = 1/(n sqrt(d0' d1')), s = u / (A + (1 - A) u) with
    A = sqrt(d1'/d0'), gives the end steps d0', d1' to first order in 1/n;
    the targets d0', d1' are corrected by fixed-point iteration until the
    actual end steps are d0 and d1. Needs n sqrt(d0 d1) < 1 (clustering)."""
    t0, t1 = d0, d1
    for _ in range(50):
        s = _vinokur(n, t0, t1)
        h0, h1 = s[1] - s[0], s[-1] - s[-2]
        if abs(h0 / d0 - 1) < 1e-12 and abs(h1 / d1 - 1) < 1e-12:
            break
        t0, t1 = t0 * d0 / h0, t1 * d1 / h1
    return s


def cgrid(af, n_side: int = 97, nj: int = 65, ds_le: float = 2e-3, ds_te: float = 2e-3, ds0: float = 1e-3,
          radius: float = 20.0, wake_ratio: float = 1.12, smooth_width: float = 0.5,
          spread: float = 1.0, spread_from: float = 0.5) -> CGrid:
    """March a C-grid from an airfoil with a closed (sharp) trailing edge.

    Each side of the contour (trailing edge to leading edge) is re-panelled to
    n_side intervals along a cubic arc-length spline with two-sided tanh
    stretching (steps ds_le at the leading edge, ds_te at the trailing edge,
    in chords). The wake cut runs along +x from the trailing edge to
    x_te + radius on each side, its first step ds_te growing geometrically
    (ratio close to wake_ratio). Each node layer is the previous one moved
    along its smoothed normal by a geometric step sequence (first step ds0,
    total `radius`); a node's normal is the Gaussian-weighted mean of the
    layer's normals within smooth_width times the distance marched (in arc
    length). Each new layer's nodes are then re-spaced along it, a fraction
    spread * (j / (nj - 1))^2 of the way to even spacing, so the lines
    crowded at the trailing edge open out towards the far field instead of
    reaching it as cells thousands of times longer than wide. The two cut
    ends march straight up and down (the downstream boundary is the line
    x = x_te + radius). This is the C topology of the reference's wall block
    (aero2d/cblock.py:314-486: layers marched along smoothed normals, the block
    continued along the wake cut), carried out to the far field; refine_cwall
    re-samples the wall layers to y+ ~ 1. Raises ValueError for an open
    trailing edge and RuntimeError when a cell inverts."""
    from .geometry import _ArcSpline
    from .grid import growth_ratio
    pts = np.asarray(af.coords(), dtype=np.float64)
    te = pts[0].copy()
    if np.hypot(*(pts[-1] - te)) > 1e-12:
        raise ValueError("cgrid needs a closed trailing edge")
    arc = np.concatenate([[0.0], np.cumsum(np.hypot(*np.diff(pts, axis=0).T))])
    spl = _ArcSpline(arc, pts[:, 0], pts[:, 1], [])
    probe = np.linspace(0.0, arc[-1], 40001)
    s_le = probe[np.argmin(spl(probe)[0])]
    up = s_le * _two_sided(n_side, ds_te / s_le, ds_le / s_le)
    lo = s_le + (arc[-1] - s_le) * _two_sided(n_side, ds_le / (arc[-1] - s_le), ds_te / (arc[-1] - s_le))
    st = np.concatenate([up, lo[1:]])
    cx, cy = spl(st)
    cx[0], cy[0], cx[-1], cy[-1] = te[0], te[1], te[0], te[1]
    cont = np.column_stack([cx, cy])
    K = int(np.ceil(np.log(1 + radius * (wake_ratio - 1) / ds_te) / np.log(wake_ratio)))
    q = growth_ratio(ds_te, radius, K)
    s = ds_te * np.cumsum(q ** np.arange(K))
    wake = te[None, :] + s[:, None] * np.array([1.0, 0.0])[None, :]
    C = np.concatenate([wake[::-1], cont, wake])
    ni = len(C)
    r = growth_ratio(ds0, radius, nj - 1)
    d = ds0 * r ** np.arange(nj - 1)
    X = np.zeros((ni, nj, 2))
    X[:, 0] = C
    dist = 0.0
    for j in range(1, nj):
        cur = X[:, j - 1]
        t = np.empty_like(cur)
        t[1:-1] = cur[2:] - cur[:-2]
        t[0], t[-1] = cur[1] - cur[0], cur[-1] - cur[-2]
        t /= np.linalg.norm(t, axis=1, keepdims=True)
        n = np.column_stack([t[:, 1], -t[:, 0]])
        # normals averaged over an arc-length window that widens with the distance
        # marched: steps near the trailing edge are far smaller than the layer
        # heights, so index-space smoothing cannot open the concave corners there
        dist += d[j - 1]
        sig = np.concatenate([[0.0], np.cumsum(np.hypot(*np.diff(cur, axis=0).T))])
        w = np.exp(-0.5 * ((sig[:, None] - sig[None, :]) / (smooth_width * dist)) ** 2)
        ns = (w * np.gradient(sig)[None, :]) @ n
        ns[0], ns[-1] = (0.0, 1.0), (0.0, -1.0)
        ns /= np.linalg.norm(ns, axis=1, keepdims=True)
        new = cur + d[j - 1] * ns
        h = np.hypot(*np.diff(new, axis=0).T)
        sg = np.concatenate([[0.0], np.cumsum(h)])
        lh = np.log(h)
        for _ in range(int(spread * max(0.0, dist - spread_from) / d[j - 1] * 4)):
            lh[1:-1] += 0.25 * (lh[:-2] + lh[2:] - 2.0 * lh[1:-1])
        tgt = np.concatenate([[0.0], np.cumsum(np.exp(lh))])
        tgt *= sg[-1] / tgt[-1]
        X[:, j, 0] = np.interp(tgt, sg, new[:, 0])
        X[:, j, 1] = np.interp(tgt, sg, new[:, 1])
    g = CGrid(X[..., 0], X[..., 1], K)
    if g.cell_areas().min() <= 0:
        raise RuntimeError(f"C-grid has inverted cells: min area {g.cell_areas().min():.3e}")
    return g


def refine_cwall(grid: CGrid, ds_wall: float, layers: int = 1, ratio: float = 1.2,
                 wake_growth: float = 1.15) -> CGrid:
    """C-grid whose first `layers` node layers are re-sampled for a first cell
    ds_wall high at the wall. Along the wake cut the first cell grows by
    wake_growth per column from the trailing edge (as the reference's
    wake_h1_growth, aero2d/rans_mesh.py:94 and aero2d/cblock.py:440), up to an
    even division of the re-sampled height: cells a few y+ high and chords long
    behind the trailing edge would hold the pseudo-time step of the whole wake
    at the acoustic limit of their height. Every line gets the same node count
    (set by the wall spacing and `ratio`); a line's own geometric ratio makes
    its first step its own target height."""
    from .grid import growth_ratio
    x, y = np.asarray(grid.x, float), np.asarray(grid.y, float)
    ni = x.shape[0]
    K = grid.n_wake
    seg = np.hypot(np.diff(x, axis=1), np.diff(y, axis=1))
    s = np.concatenate([np.zeros((ni, 1)), np.cumsum(seg, 1)], 1)
    H = s[:, layers]
    if not 0 < ds_wall < H.min():
        raise ValueError("ds_wall must be positive and below the height of the re-sampled layers")
    n = max(layers, int(np.ceil(np.log(1 + H.mean() * (ratio - 1) / ds_wall) / np.log(ratio))))
    col = np.zeros(ni)                       # wake column counted from the trailing edge
    col[:K] = np.arange(K, 0, -1)
    col[ni - K:] = np.arange(1, K + 1)
    target = np.minimum(ds_wall * wake_growth ** col, H / n)
    nx = np.empty((ni, n + 1))
    ny = np.empty_like(nx)
    for i in range(ni):
        if target[i] * n >= H[i] * (1 - 1e-12):
            t = np.linspace(0.0, 1.0, n + 1)
        else:
            q = growth_ratio(target[i] / H[i], 1.0, n)
            t = np.concatenate([[0.0], np.cumsum((target[i] / H[i]) * q ** np.arange(n))])
            t /= t[-1]
        st = t * s[i, layers]
        nx[i] = np.interp(st, s[i, :layers + 1], x[i, :layers + 1])
        ny[i] = np.interp(st, s[i, :layers + 1], y[i, :layers + 1])
    g = CGrid(np.concatenate([nx, x[:, layers + 1:]], 1), np.concatenate([ny, y[:, layers + 1:]], 1), K)
    if g.cell_areas().min() <= 0:
        raise RuntimeError("refined C-grid has inverted cells")
    return g


def cgrid_mesh(grid: CGrid, quad_layers: int | None = None) -> Mesh:
    """Hybrid mesh from a C-grid: cells of layers j < quad_layers (the wall and
    wake-cut block) stay quadrilaterals, every cell beyond is cut into two
    triangles along alternating diagonals (quad_layers=None: all quadrilaterals).
    The coincident nodes of the wake cut are merged, so the cells above and
    below the cut are neighbours across interior faces. Wall: the contour's
    edges (body 0); far field: the outer layer and the two downstream end
    lines. Cells are numbered layer by layer from the wall, i fastest."""
    x, y = np.asarray(grid.x, np.float64), np.asarray(grid.y, np.float64)
    ni, nj = x.shape
    K = grid.n_wake
    full = np.arange(ni * nj).reshape(nj, ni)     # (j, i) -> raw id
    raw = full.copy()
    raw[0, :K + 1] = full[0, ni - 1 - np.arange(K + 1)]
    used = np.unique(raw)
    remap = -np.ones(ni * nj, dtype=np.int64)
    remap[used] = np.arange(len(used))
    nid = remap[raw]                              # (j, i) -> compact node id
    xy = np.column_stack([x.T.ravel(), y.T.ravel()])
    nodes = xy[used]
    nq = nj - 1 if quad_layers is None else int(quad_layers)
    cells = []
    for j in range(nj - 1):
        for i in range(ni - 1):
            a, b, c, d = nid[j, i], nid[j, i + 1], nid[j + 1, i + 1], nid[j + 1, i]
            if j < nq:
                cells.append([a, b, c, d])
            elif (i + j) % 2 == 0:
                cells += [[a, b, c], [a, c, d]]
            else:
                cells += [[a, b, d], [b, c, d]]
    key = lambda p, q: (int(min(p, q)), int(max(p, q)))            # noqa: E731
    wall = {key(nid[0, i], nid[0, i + 1]): 0 for i in range(K, ni - 1 - K)}
    far = {key(nid[nj - 1, i], nid[nj - 1, i + 1]) for i in range(ni - 1)}
    far |= {key(nid[j, 0], nid[j + 1, 0]) for j in range(nj - 1)}
    far |= {key(nid[j, ni - 1], nid[j + 1, ni - 1]) for j in range(nj - 1)}
    return connectivity(nodes, cells, wall, far)
