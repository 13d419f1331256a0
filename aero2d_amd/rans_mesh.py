"""Unstructured hybrid mesh for the RANS solver (host side, NumPy FP64).

`Mesh` carries the arrays the RANS residual reads: nodes, cells (counter-
clockwise vertex lists, triangles and quadrilaterals), cell centroids and
areas, faces with their left/right cells, unit normals (left -> right), lengths
and midpoints, boundary tags (0 interior, 1 wall, 2 far field), the body of a
wall face and each cell's distance to the nearest wall (aero2d/rans_mesh.py:19-36).

`connectivity` builds a Mesh from nodes, cells and labelled boundary edges
(aero2d/rans_mesh.py:283-337). The reference meshes with gmsh, which this
image does not have; `ogrid_mesh` turns a marched O-grid (aero2d_amd.grid) into
the same kind of hybrid mesh: quadrilaterals in the wall layers, each outer
quadrilateral split into two triangles.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["Mesh", "connectivity", "ogrid_mesh"]


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
