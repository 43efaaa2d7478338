"""Gmsh 2.2 ASCII writer for the synthetic periodic box meshes the solver runs on.

The file format follows the reference's own meshes (meshes/QuadBox_32x32.msh,
meshes/HexBox.msh): a ``$PhysicalNames`` block with a ``PERIODIC`` boundary group and a
``FLUID`` volume group, nodes numbered 1..N in order (the reader ignores the ids,
src/geometry.cpp:436-438), linear quads (Gmsh type 3) or hexes (type 5), and one
boundary element (type 1 lines in 2D, type 3 quads in 3D) per box face.  Coordinates
are printed with ``%.17g`` so every reader recovers the same doubles.

Nodes are numbered with x fastest, then y, then z; elements likewise.  Quad nodes are
counter-clockwise; hex nodes are the bottom quad (z-) then the top quad (z+), which is
the Gmsh convention the reference's face tables (src/geometry.cpp:218-265) assume.
"""

from __future__ import annotations

import numpy as np

__all__ = ["box_nodes", "write_box", "write_quad_box", "write_hex_box"]


def box_nodes(n, lo, hi):
    """Node coordinate arrays along each axis: ``lo + (hi - lo) * i / n``."""
    return [lo[d] + (hi[d] - lo[d]) * np.arange(n[d] + 1, dtype=np.float64) / n[d]
            for d in range(len(n))]


def _fmt_rows(ids, cols):
    """Join integer ids and float columns into Gmsh node lines."""
    parts = [ids.astype(str)]
    for c in cols:
        parts.append(np.char.mod("%.17g", c))
    out = parts[0]
    for p in parts[1:]:
        out = np.char.add(np.char.add(out, " "), p)
    return "\n".join(out.tolist())


def write_quad_box(path, nx, ny, lo=(0.0, 0.0), hi=(1.0, 1.0), perturb=None):
    """Write an ``nx`` x ``ny`` periodic quad box on ``[lo, hi]``."""
    xs, ys = box_nodes((nx, ny), lo, hi)
    X, Y = np.meshgrid(xs, ys, indexing="xy")  # [ny+1][nx+1], x fastest when raveled
    X = X.ravel()
    Y = Y.ravel()
    if perturb is not None:
        X, Y = perturb(X, Y)
    nnode = X.size
    node_ids = np.arange(1, nnode + 1)
    nid = lambda i, j: j * (nx + 1) + i + 1  # noqa: E731

    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", "2",
             '1 1 "PERIODIC"', '2 2 "FLUID"', "$EndPhysicalNames", "$Nodes", str(nnode)]
    lines.append(_fmt_rows(node_ids, [X, Y, np.zeros_like(X)]))
    lines.append("$EndNodes")

    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    i = i.ravel(); j = j.ravel()
    quads = np.stack([nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)], axis=1)
    bnd = []
    for a in range(nx):
        bnd.append((nid(a, 0), nid(a + 1, 0)))
        bnd.append((nid(a + 1, ny), nid(a, ny)))
    for b in range(ny):
        bnd.append((nid(nx, b), nid(nx, b + 1)))
        bnd.append((nid(0, b + 1), nid(0, b)))
    ne = quads.shape[0]
    nb = len(bnd)
    lines.append("$Elements")
    lines.append(str(ne + nb))
    eid = np.arange(1, ne + 1).astype(str)
    body = np.char.add(eid, " 3 2 2 2")
    for k in range(4):
        body = np.char.add(np.char.add(body, " "), quads[:, k].astype(str))
    lines.append("\n".join(body.tolist()))
    for k, (a, b) in enumerate(bnd):
        lines.append(f"{ne + k + 1} 1 2 1 1 {a} {b}")
    lines.append("$EndElements")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def write_hex_box(path, nx, ny, nz, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), perturb=None):
    """Write an ``nx`` x ``ny`` x ``nz`` periodic hex box on ``[lo, hi]``."""
    xs, ys, zs = box_nodes((nx, ny, nz), lo, hi)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")  # x fastest when raveled
    X = X.ravel(); Y = Y.ravel(); Z = Z.ravel()
    if perturb is not None:
        X, Y, Z = perturb(X, Y, Z)
    nnode = X.size
    node_ids = np.arange(1, nnode + 1)

    def nid(i, j, k):
        return (k * (ny + 1) + j) * (nx + 1) + i + 1

    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", "2",
             '2 1 "PERIODIC"', '3 2 "FLUID"', "$EndPhysicalNames", "$Nodes", str(nnode)]
    lines.append(_fmt_rows(node_ids, [X, Y, Z]))
    lines.append("$EndNodes")

    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i = i.ravel(); j = j.ravel(); k = k.ravel()
    hexes = np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                      nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1),
                      nid(i, j + 1, k + 1)], axis=1)
    bq = []
    # z faces
    a, b = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    a = a.ravel(); b = b.ravel()
    bq.append(np.stack([nid(a, b, 0), nid(a, b + 1, 0), nid(a + 1, b + 1, 0), nid(a + 1, b, 0)], 1))
    bq.append(np.stack([nid(a, b, nz), nid(a + 1, b, nz), nid(a + 1, b + 1, nz), nid(a, b + 1, nz)], 1))
    # x faces
    b, c = np.meshgrid(np.arange(ny), np.arange(nz), indexing="xy")
    b = b.ravel(); c = c.ravel()
    bq.append(np.stack([nid(0, b, c), nid(0, b, c + 1), nid(0, b + 1, c + 1), nid(0, b + 1, c)], 1))
    bq.append(np.stack([nid(nx, b, c), nid(nx, b + 1, c), nid(nx, b + 1, c + 1), nid(nx, b, c + 1)], 1))
    # y faces
    a, c = np.meshgrid(np.arange(nx), np.arange(nz), indexing="xy")
    a = a.ravel(); c = c.ravel()
    bq.append(np.stack([nid(a, 0, c), nid(a + 1, 0, c), nid(a + 1, 0, c + 1), nid(a, 0, c + 1)], 1))
    bq.append(np.stack([nid(a, ny, c), nid(a, ny, c + 1), nid(a + 1, ny, c + 1), nid(a + 1, ny, c)], 1))
    bq = np.concatenate(bq, axis=0)
    ne = hexes.shape[0]
    nb = bq.shape[0]
    lines.append("$Elements")
    lines.append(str(ne + nb))
    eid = np.arange(1, ne + 1).astype(str)
    body = np.char.add(eid, " 5 2 2 2")
    for m in range(8):
        body = np.char.add(np.char.add(body, " "), hexes[:, m].astype(str))
    lines.append("\n".join(body.tolist()))
    bid = np.arange(ne + 1, ne + nb + 1).astype(str)
    bbody = np.char.add(bid, " 3 2 1 1")
    for m in range(4):
        bbody = np.char.add(np.char.add(bbody, " "), bq[:, m].astype(str))
    lines.append("\n".join(bbody.tolist()))
    lines.append("$EndElements")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def write_box(path, n, lo, hi, perturb=None):
    """Dispatch on ``len(n)``: 2 -> quads, 3 -> hexes."""
    if len(n) == 2:
        return write_quad_box(path, n[0], n[1], lo, hi, perturb)
    if len(n) == 3:
        return write_hex_box(path, n[0], n[1], n[2], lo, hi, perturb)
    raise ValueError("n must have 2 or 3 entries")
