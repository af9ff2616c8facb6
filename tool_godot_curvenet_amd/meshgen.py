"""Seeded synthetic meshes and cut curves for the benchmark workloads.

These are input generators, not part of the solver path.  Every mesh is a
consistently oriented triangle list (int32, CCW seen from outside) with FP64
positions.  Cut curves are closed edge loops of the mesh, returned as an
(m, 2) int32 array of undirected edges (the input of ``cn_cut_seams``).

Workloads
---------
* ``open_cylinder``   64 x 64 quad grid cylinder, one seam ring (4,096 verts)
* ``geodesic_sphere`` frequency-nu geodesic sphere, V = 10 nu^2 + 2
  (nu = 2^L gives the level-L icosphere vertex count)
* ``tilted_loops``    latitude loops with random tilt snapped to edge paths
* ``capsule_union``   union of capsules with ring cuts (~100k verts)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

BASE_SEED = 20261203


@dataclass
class Mesh:
    pos: np.ndarray            # (nv, 3) float64
    tris: np.ndarray           # (nf, 3) int32
    cut_edges: np.ndarray      # (m, 2) int32
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def nv(self) -> int:
        return int(self.pos.shape[0])

    @property
    def nf(self) -> int:
        return int(self.tris.shape[0])


def mean_edge_length(pos: np.ndarray, tris: np.ndarray) -> float:
    e = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]])
    d = pos[e[:, 0]] - pos[e[:, 1]]
    return float(np.sqrt((d * d).sum(1)).mean())


def jitter(pos: np.ndarray, amount: float, h: float, rng: np.random.Generator) -> np.ndarray:
    """Uniform +-amount*h per component."""
    return pos + rng.uniform(-amount * h, amount * h, size=pos.shape)


# ---------------------------------------------------------------- cylinder --
def open_cylinder(n_around: int = 64, n_rings: int = 64, radius: float = 1.0,
                  height: float = 1.0, jitter_frac: float = 0.1,
                  seam_ring: int | None = None, seed: int = BASE_SEED) -> Mesh:
    """Quad-grid open cylinder, each quad split along one diagonal.

    Vertex (ring j, index i) has id j*n_around + i.  The seam is the closed
    ring ``seam_ring`` (default n_rings // 2)."""
    rng = np.random.default_rng(seed)
    j, i = np.meshgrid(np.arange(n_rings), np.arange(n_around), indexing="ij")
    th = 2 * np.pi * i / n_around
    z = height * j / (n_rings - 1)
    pos = np.stack([radius * np.cos(th), radius * np.sin(th), z], -1).reshape(-1, 3)
    tris = []
    for r in range(n_rings - 1):
        for a in range(n_around):
            b = (a + 1) % n_around
            v00, v01 = r * n_around + a, r * n_around + b
            v10, v11 = (r + 1) * n_around + a, (r + 1) * n_around + b
            tris.append((v00, v01, v11))
            tris.append((v00, v11, v10))
    tris = np.asarray(tris, dtype=np.int32)
    h = mean_edge_length(pos, tris)
    pos = jitter(pos, jitter_frac, h, rng)
    sr = n_rings // 2 if seam_ring is None else seam_ring
    cut = np.array([(sr * n_around + a, sr * n_around + (a + 1) % n_around)
                    for a in range(n_around)], dtype=np.int32)
    return Mesh(pos, tris, cut, "open_cylinder", {"n_around": n_around, "n_rings": n_rings,
                                                   "seam_ring": sr, "h": h})


# ---------------------------------------------------------- geodesic sphere --
def _icosahedron():
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def geodesic_sphere(nu: int, radius: float = 1.0, jitter_frac: float = 0.0,
                    seed: int = BASE_SEED) -> Mesh:
    """Class-I geodesic sphere of frequency nu: every icosahedron face is cut
    into nu^2 triangles on a barycentric grid, points pushed to the sphere.
    V = 10 nu^2 + 2, F = 20 nu^2."""
    rng = np.random.default_rng(seed)
    cv, cf = _icosahedron()
    ii, jj = np.meshgrid(np.arange(nu + 1), np.arange(nu + 1), indexing="ij")
    msk = ii + jj <= nu
    ii, jj = ii[msk], jj[msk]               # local grid points
    npt = ii.size
    lid = -np.ones((nu + 1, nu + 1), dtype=np.int64)
    lid[ii, jj] = np.arange(npt)
    # local triangles on the grid (CCW like the face)
    lt = []
    for i in range(nu):
        for j in range(nu - i):
            lt.append((lid[i, j], lid[i + 1, j], lid[i, j + 1]))
            if i + j + 1 < nu:
                lt.append((lid[i + 1, j], lid[i + 1, j + 1], lid[i, j + 1]))
    lt = np.asarray(lt, dtype=np.int64)
    # global keys: weights on the three corners, canonicalised by corner id
    wa = nu - ii - jj
    keys = np.zeros((20 * npt, 6), dtype=np.int64)
    for f in range(20):
        a, b, c = cf[f]
        cid = np.stack([np.full(npt, a), np.full(npt, b), np.full(npt, c)], 1)
        w = np.stack([wa, ii, jj], 1)
        w = np.where(w == 0, 0, w)
        cid = np.where(w == 0, 99, cid)      # drop zero weights from the key
        order = np.argsort(cid, axis=1, kind="stable")
        cid_s = np.take_along_axis(cid, order, 1)
        w_s = np.take_along_axis(w, order, 1)
        keys[f * npt:(f + 1) * npt, :3] = cid_s
        keys[f * npt:(f + 1) * npt, 3:] = w_s
    uk, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    # order vertices by first appearance (deterministic, face-major)
    ordr = np.argsort(first, kind="stable")
    rank = np.empty_like(ordr)
    rank[ordr] = np.arange(ordr.size)
    gid = rank[inv]
    nvert = uk.shape[0]
    pos = np.zeros((nvert, 3))
    allp = np.zeros((20 * npt, 3))
    for f in range(20):
        a, b, c = cf[f]
        p = (cv[a][None] * wa[:, None] + cv[b][None] * ii[:, None] + cv[c][None] * jj[:, None]) / nu
        allp[f * npt:(f + 1) * npt] = p
    pos[gid] = allp
    pos = radius * pos / np.linalg.norm(pos, axis=1, keepdims=True)
    tris = np.concatenate([gid[f * npt + lt] for f in range(20)]).astype(np.int32)
    h = mean_edge_length(pos, tris)
    if jitter_frac > 0:
        pos = jitter(pos, jitter_frac, h, rng)
    return Mesh(pos, tris, np.zeros((0, 2), np.int32), f"geodesic_nu{nu}", {"nu": nu, "h": h})


# ------------------------------------------------------------- edge paths --
def _edge_graph(pos, tris):
    import scipy.sparse as sp
    e = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]]).astype(np.int64)
    e = np.sort(e, axis=1)
    e = np.unique(e, axis=0)
    w = np.linalg.norm(pos[e[:, 0]] - pos[e[:, 1]], axis=1)
    n = pos.shape[0]
    g = sp.coo_matrix((np.concatenate([w, w]), (np.concatenate([e[:, 0], e[:, 1]]),
                                                np.concatenate([e[:, 1], e[:, 0]]))),
                      shape=(n, n)).tocsr()
    return g


def _simplify_loop(verts: list[int]) -> list[int]:
    """Remove sub-loops so every vertex appears once (keeps the outer loop)."""
    out: list[int] = []
    where: dict[int, int] = {}
    for v in verts:
        if v in where:
            k = where[v]
            for u in out[k + 1:]:
                where.pop(u, None)
            out = out[:k + 1]
        else:
            where[v] = len(out)
            out.append(v)
    return out


def loop_edges(loop: list[int]) -> np.ndarray:
    m = len(loop)
    return np.array([(loop[a], loop[(a + 1) % m]) for a in range(m)], dtype=np.int32)


def _nearest_vertex(tree, pos, p) -> int:
    """Vertex nearest to p by squared distance ((pos - p)**2).sum(1), the
    lowest index among ties (SurfaceProjection's strict < scan): the tree
    proposes every vertex within a hair of its nearest distance, and the
    exact rule picks among those."""
    d0, _ = tree.query(p)
    cand = np.array(sorted(tree.query_ball_point(p, d0 * (1 + 1e-9) + 1e-15)), dtype=np.int64)
    d = ((pos[cand] - p[None]) ** 2).sum(1)
    return int(cand[int(np.argmin(d))])


def loop_knots(mesh: Mesh, n_loops: int = 8, max_tilt_deg: float = 30.0,
               n_way: int = 32, seed: int = BASE_SEED + 1) -> list:
    """The curve knots of `tilted_loops`: for each of n_loops seeded latitude
    circles of random tilt, n_way points on the circle (n_way x 3)."""
    rng = np.random.default_rng(seed)
    pos = mesh.pos
    rad = float(np.linalg.norm(pos, axis=1).mean())
    lats = np.sort(rng.uniform(-0.8, 0.8, size=n_loops))
    out = []
    for L in range(n_loops):
        tilt = math.radians(rng.uniform(0, max_tilt_deg))
        az = rng.uniform(0, 2 * math.pi)
        nrm = np.array([math.sin(tilt) * math.cos(az), math.sin(tilt) * math.sin(az), math.cos(tilt)])
        u = np.cross(nrm, [1.0, 0.0, 0.0])
        if np.linalg.norm(u) < 1e-6:
            u = np.cross(nrm, [0.0, 1.0, 0.0])
        u /= np.linalg.norm(u)
        v = np.cross(nrm, u)
        h = lats[L] * rad
        rr = math.sqrt(max(rad * rad - h * h, 1e-12))
        ang = np.linspace(0, 2 * math.pi, n_way, endpoint=False)
        out.append(h * nrm[None] + rr * (np.cos(ang)[:, None] * u[None]
                                         + np.sin(ang)[:, None] * v[None]))
    return out


def tilted_loops(mesh: Mesh, n_loops: int = 8, max_tilt_deg: float = 30.0,
                 n_way: int = 32, seed: int = BASE_SEED + 1) -> Mesh:
    """Latitude loops of random tilt, snapped to the mesh: waypoints on the
    tilted circle go to their nearest vertex (strict <, lowest index), and
    consecutive waypoints are joined by shortest edge paths."""
    from scipy.sparse.csgraph import dijkstra
    from scipy.spatial import cKDTree
    pos = mesh.pos
    g = _edge_graph(pos, mesh.tris)
    tree = cKDTree(pos)
    edges = []
    loops = []
    for pts in loop_knots(mesh, n_loops, max_tilt_deg, n_way, seed):
        way = [_nearest_vertex(tree, pos, p) for p in pts]
        way = [w for k, w in enumerate(way) if k == 0 or w != way[k - 1]]
        if way[0] == way[-1]:
            way = way[:-1]
        path: list[int] = []
        for k in range(len(way)):
            a, b = way[k], way[(k + 1) % len(way)]
            # Dijkstra settles vertices in order of distance, so a search cut
            # off past a few chord lengths (a mesh path between neighbouring
            # waypoints is barely longer than the chord) gives the same
            # predecessors along the path to b; the full search is the
            # fallback when b was not reached
            # graph distance is never shorter than Euclidean distance, so the
            # cut-off search stays inside the ball of radius lim around a:
            # it runs on that ball's subgraph (vertices kept in ascending
            # order).  An edge path between neighbouring waypoints is ~1.1x
            # the chord on these meshes; a longer one takes the fallback.
            lim = 1.5 * float(np.linalg.norm(pos[a] - pos[b])) + 1e-12
            ball = np.array(sorted(tree.query_ball_point(pos[a], lim * (1 + 1e-9))), dtype=np.int64)
            sub = g[ball][:, ball]
            la, lb = int(np.searchsorted(ball, a)), int(np.searchsorted(ball, b))
            seg = [b]
            if lb < len(ball) and ball[lb] == b:
                _, pred = dijkstra(sub, indices=la, return_predecessors=True, limit=lim)
                if pred[lb] >= 0 or a == b:
                    while seg[-1] != a:
                        seg.append(int(ball[pred[int(np.searchsorted(ball, seg[-1]))]]))
            if seg[-1] != a:
                _, pred = dijkstra(g, indices=a, return_predecessors=True)
                seg = [b]
                while seg[-1] != a:
                    seg.append(int(pred[seg[-1]]))
            seg.reverse()
            path.extend(seg[:-1])
        loop = _simplify_loop(path)
        if len(loop) < 3:
            continue
        loops.append(loop)
        edges.append(loop_edges(loop))
    cut = np.unique(np.sort(np.concatenate(edges), axis=1), axis=0).astype(np.int32)
    meta = dict(mesh.meta)
    meta["loops"] = [len(x) for x in loops]
    meta["loop_verts"] = [list(map(int, x)) for x in loops]
    return Mesh(mesh.pos, mesh.tris, cut, mesh.name + "_loops", meta)


# ---------------------------------------------------------------- capsules --
def _zip_rings(a_ids, a_ang, b_ids, b_ang, out):
    """Triangulate between closed rings a (lower) and b (upper)."""
    na, nb = len(a_ids), len(b_ids)
    i = j = 0
    aa = np.concatenate([a_ang, a_ang[:1] + 2 * np.pi])
    bb = np.concatenate([b_ang, b_ang[:1] + 2 * np.pi])
    # align b's start to the point nearest a[0]
    while i < na or j < nb:
        if j >= nb or (i < na and aa[i + 1] <= bb[j + 1]):
            out.append((a_ids[i % na], a_ids[(i + 1) % na], b_ids[j % nb]))
            i += 1
        else:
            out.append((a_ids[i % na], b_ids[(j + 1) % nb], b_ids[j % nb]))
            j += 1


def capsule(radius: float, length: float, h: float, jitter_frac: float,
            rng: np.random.Generator, start: int = 0):
    """Closed capsule along +z from z=0 to z=length, target edge length h.
    Returns (pos, tris, rings) where rings[k] lists the vertex ids of the
    cylinder rings (for cut curves)."""
    m = max(6, int(round(2 * np.pi * radius / h)))
    n_cyl = max(2, int(round(length / h)) + 1)
    n_hemi = max(1, int(round(0.5 * np.pi * radius / h)))
    ring_z, ring_r = [], []
    for k in range(1, n_hemi):                       # bottom cap, pole excluded
        th = 0.5 * np.pi * k / n_hemi                # from the pole
        ring_z.append(-radius * np.cos(th))
        ring_r.append(radius * np.sin(th))
    cyl_first = len(ring_z)
    for k in range(n_cyl):
        ring_z.append(length * k / (n_cyl - 1))
        ring_r.append(radius)
    cyl_last = len(ring_z) - 1
    for k in range(n_hemi - 1, 0, -1):
        th = 0.5 * np.pi * k / n_hemi
        ring_z.append(length + radius * np.cos(th))
        ring_r.append(radius * np.sin(th))
    pos = [np.array([0.0, 0.0, -radius])]
    ids = 1
    rings, angs = [], []
    for z, r in zip(ring_z, ring_r):
        cnt = max(6, int(round(m * r / radius)))
        phase = rng.uniform(0, 2 * np.pi / cnt) if cnt != m else 0.0
        a = phase + 2 * np.pi * np.arange(cnt) / cnt
        pos.extend(np.stack([r * np.cos(a), r * np.sin(a), np.full(cnt, z)], 1))
        rings.append(np.arange(ids, ids + cnt))
        angs.append(a - a[0])
        ids += cnt
    top = ids
    pos.append(np.array([0.0, 0.0, length + radius]))
    pos = np.asarray(pos)
    tris: list = []
    r0 = rings[0]
    for i in range(len(r0)):
        tris.append((0, r0[(i + 1) % len(r0)], r0[i]))
    for k in range(len(rings) - 1):
        a_ids, b_ids = rings[k], rings[k + 1]
        # angles relative to a common origin
        a_ang = np.mod(np.arctan2(pos[a_ids, 1], pos[a_ids, 0]), 2 * np.pi)
        b_ang = np.mod(np.arctan2(pos[b_ids, 1], pos[b_ids, 0]), 2 * np.pi)
        oa, ob = np.argsort(a_ang), np.argsort(b_ang)
        a_ids, a_ang = a_ids[oa], a_ang[oa]
        b_ids, b_ang = b_ids[ob], b_ang[ob]
        # rotate b so that its first point is the first at/after a[0]
        sh = int(np.searchsorted(b_ang, a_ang[0]))
        b_ids = np.roll(b_ids, -sh)
        b_ang = np.roll(b_ang, -sh)
        b_ang = np.where(b_ang < a_ang[0], b_ang + 2 * np.pi, b_ang)
        _zip_rings(list(a_ids), a_ang, list(b_ids), b_ang, tris)
    rl = rings[-1]
    for i in range(len(rl)):
        tris.append((rl[i], rl[(i + 1) % len(rl)], top))
    tris = np.asarray(tris, dtype=np.int64)
    if jitter_frac > 0:
        pos = jitter(pos, jitter_frac, h, rng)
    cyl_rings = [rings[k] + start for k in range(cyl_first, cyl_last + 1)]
    return pos, tris + start, cyl_rings


def _rotation_to(d):
    d = d / np.linalg.norm(d)
    z = np.array([0.0, 0.0, 1.0])
    v = np.cross(z, d)
    s, c = np.linalg.norm(v), float(z @ d)
    if s < 1e-12:
        return np.eye(3) if c > 0 else np.diag([1.0, -1.0, -1.0])
    vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
    return np.eye(3) + vx + vx @ vx * ((1 - c) / (s * s))


def capsule_union(n_caps: int = 12, n_cuts: int = 16, target_verts: int = 100_000,
                  jitter_frac: float = 0.05, seed: int = BASE_SEED + 3) -> Mesh:
    """Capsules (radius U(0.2,1), length U(1,4)) on a random skeleton; each
    capsule's start sits on an earlier capsule's end sphere.  Components stay
    topologically separate.  Cut curves are cylinder rings: one per capsule,
    the remaining n_cuts - n_caps on randomly chosen capsules."""
    rng0 = np.random.default_rng(seed)
    radii = rng0.uniform(0.2, 1.0, n_caps)
    lens = rng0.uniform(1.0, 4.0, n_caps)
    area = float(np.sum(2 * np.pi * radii * lens + 4 * np.pi * radii ** 2))
    h = math.sqrt(area / (target_verts * math.sqrt(3) / 2))
    # one correction pass: vertex count scales like h^-2
    cnt = sum(capsule(radii[c], lens[c], h, 0.0, np.random.default_rng(0))[0].shape[0]
              for c in range(n_caps))
    h *= math.sqrt(cnt / target_verts)
    rng = np.random.default_rng(seed + 1000)
    joints = [np.zeros(3)]
    frames = []
    for c in range(n_caps):
        base = joints[int(rng.integers(0, len(joints)))]
        d = rng.normal(size=3)
        frames.append((base, _rotation_to(d)))
        joints.append(base + _rotation_to(d) @ np.array([0.0, 0.0, lens[c]]))
    extra = list(rng.choice(n_caps, size=max(0, n_cuts - n_caps), replace=True))
    ncut_per = [1 + extra.count(c) for c in range(n_caps)]
    pos_all, tri_all, cut_all = [], [], []
    start = 0
    for c in range(n_caps):
        p, t, rings = capsule(radii[c], lens[c], h, jitter_frac, rng, start)
        base, R = frames[c]
        p = p @ R.T + base[None]
        pos_all.append(p)
        tri_all.append(t)
        sel = rng.choice(np.arange(1, len(rings) - 1), size=min(ncut_per[c], len(rings) - 2),
                         replace=False)
        for s in sorted(sel):
            rg = rings[int(s)]
            # ring ids are in angular order already
            cut_all.append(loop_edges(list(rg)))
        start += p.shape[0]
    pos = np.concatenate(pos_all)
    tris = np.concatenate(tri_all).astype(np.int32)
    cut = np.concatenate(cut_all).astype(np.int32)
    return Mesh(pos, tris, cut, "capsule_union",
                {"n_caps": n_caps, "n_cuts": int(sum(min(ncut_per[c], 99) for c in range(n_caps))),
                 "h": h})


# ------------------------------------------------------------- workloads --
def config_mesh(name: str) -> Mesh:
    """The seeded meshes of the benchmark configurations."""
    if name == "cylinder64":
        return open_cylinder(64, 64, seed=BASE_SEED)
    if name == "icosphere6":
        m = geodesic_sphere(64, jitter_frac=0.05, seed=BASE_SEED + 2)
        return tilted_loops(m, 8, 30.0, seed=BASE_SEED + 1)
    if name == "capsules100k":
        return capsule_union(12, 16, 100_000, seed=BASE_SEED + 3)
    if name in ("sphere45", "sphere224"):
        # geodesic sphere (V = 10 nu^2 + 2) with 8 tilted latitude loops, the
        # curves of the B-spline gradient workload
        nu = 45 if name == "sphere45" else 224
        m = geodesic_sphere(nu, jitter_frac=0.05, seed=BASE_SEED + 4)
        return tilted_loops(m, 8, 30.0, seed=BASE_SEED + 5)
    if name == "icosphere9":
        # level-9 icosphere (nu = 512, V = 2,621,442) with 64 tilted loops
        m = geodesic_sphere(512, jitter_frac=0.05, seed=BASE_SEED + 8)
        return tilted_loops(m, 64, 30.0, n_way=128, seed=BASE_SEED + 9)
    raise KeyError(name)


def meshes64(n: int = 64, lo: int = 2_000, hi: int = 200_000,
             seed: int = BASE_SEED + 6) -> list:
    """Many uneven meshes (the C3b workload): n capsule unions whose vertex
    counts are exp(U(ln lo, ln hi)); a union of t vertices has
    clip(round(t / 9000), 1, 12) capsules and a third more ring cuts."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(lo), math.log(hi), size=n))
    out = []
    for i, t in enumerate(sizes):
        caps = int(np.clip(round(t / 9000), 1, 12))
        out.append(capsule_union(caps, caps + 1 + caps // 3, int(t), seed=seed + 1 + i))
    return out


def seeded_boundary(n_con: int, k: int = 3, amp: float = 0.2, seed: int = BASE_SEED + 10) -> np.ndarray:
    """Seeded boundary displacements U(-amp, amp), n_con x k."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-amp, amp, size=(n_con, k))
