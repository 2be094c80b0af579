"""The generated marching-cubes table (host code of libvoxhip.so, no GPU) and the CPU
mesh assembly of the oracle, pinned by what any marching-cubes mesh must satisfy:
triangles use exactly the edges whose ends straddle the threshold, face the below side,
close up into a 2-manifold without boundary when the surface stays inside the lattice
(also on random fields full of ambiguous faces), and a sphere comes out with Euler
characteristic 2, area within 1% of 4 pi r^2 and outward normals.  The classic table
has at most 5 triangles per cube and 1 for a single cut-off corner; so does this one."""
import numpy as np

from oracle import torch_ref as tr
from voxurf_amd.mesh import mc_tables_host


def _corner(c):
    return np.array([c & 1, c >> 1 & 1, c >> 2 & 1], dtype=np.float64)


def _edge_ends(e):
    ax = e >> 2
    others = [a for a in range(3) if a != ax]
    p = [0, 0, 0]
    p[others[0]] = e & 1
    p[others[1]] = e >> 1 & 1
    c0 = p[0] | p[1] << 1 | p[2] << 2
    return c0, c0 | 1 << ax


def test_table_uses_crossing_edges_and_faces_below_side():
    count, edges = mc_tables_host()
    assert count.max() == 5 and count[1] == 1 and count[0] == 0 and count[255] == 0
    for cfg in range(256):
        below = [(cfg >> c) & 1 for c in range(8)]
        crossing = {e for e in range(12) if below[_edge_ends(e)[0]] != below[_edge_ends(e)[1]]}
        used = set()
        for t in range(count[cfg]):
            tri = [int(edges[cfg][3 * t + r]) for r in range(3)]
            used |= set(tri)
            pts = [0.5 * (_corner(_edge_ends(e)[0]) + _corner(_edge_ends(e)[1])) for e in tri]
            n = np.cross(pts[1] - pts[0], pts[2] - pts[0])
            toward = 0.0
            for e in tri:
                c0, c1 = _edge_ends(e)
                b, a = (c0, c1) if below[c0] else (c1, c0)
                toward += float(n @ (_corner(b) - _corner(a)))
            assert toward > 0, (cfg, t)
        assert used == crossing, cfg
        # no triangle lies in a cube face (a neighbour could duplicate it there)
        for t in range(count[cfg]):
            tri = [int(edges[cfg][3 * t + r]) for r in range(3)]
            pts = np.asarray([0.5 * (_corner(_edge_ends(e)[0]) + _corner(_edge_ends(e)[1]))
                              for e in tri])
            assert not any(np.all(pts[:, a] == s_) for a in range(3) for s_ in (0.0, 1.0)), cfg


def _mesh(tris):
    """Vertex ids by exact position; returns (V, faces [F, 3])."""
    ids = {}
    faces = []
    for tri in tris:
        f = []
        for v in tri:
            key = tuple(v)
            if key not in ids:
                ids[key] = len(ids)
            f.append(ids[key])
        faces.append(f)
    return len(ids), np.asarray(faces, dtype=np.int64).reshape(-1, 3)


def _closed_oriented(faces):
    directed = {}
    for f in faces:
        for a, b in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            directed[(a, b)] = directed.get((a, b), 0) + 1
    ok = all(c == 1 for c in directed.values()) and all((b, a) in directed for a, b in directed)
    return ok, len(directed) // 2


def test_random_fields_close_up():
    count, edges = mc_tables_host()
    rng = np.random.default_rng(0)
    for trial in range(4):
        u = rng.uniform(-1, 1, (9, 8, 10)).astype(np.float32)
        u[0], u[-1], u[:, 0], u[:, -1], u[:, :, 0], u[:, :, -1] = 1, 1, 1, 1, 1, 1
        tris = tr.mc_assemble(u, 0.0, count, edges)
        V, faces = _mesh(tris)
        ok, E = _closed_oriented(faces)
        assert ok, trial
        # the vertices are exactly the brute-force crossing-edge points
        verts = np.asarray(sorted({tuple(v) for tri in tris for v in tri}))
        assert np.array_equal(verts, tr.mc_vertices(u, 0.0))


def test_sphere_mesh():
    count, edges = mc_tables_host()
    n, r = 24, 0.6
    x = np.linspace(-1, 1, n, dtype=np.float64)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = -(np.sqrt(X * X + Y * Y + Z * Z) - r).astype(np.float32)   # -sdf, as meshed
    tris = tr.mc_assemble(u, 0.0, count, edges)
    V, faces = _mesh(tris)
    ok, E = _closed_oriented(faces)
    assert ok and V - E + len(faces) == 2
    h = 2.0 / (n - 1)
    p = tris * h - 1.0
    n_ = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    area = 0.5 * np.linalg.norm(n_, axis=1).sum()
    assert abs(area - 4 * np.pi * r * r) < 0.01 * 4 * np.pi * r * r
    outward = (n_ * p.mean(1)).sum(1) > 0
    assert outward.all()
