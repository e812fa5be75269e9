"""Procedural test meshes: pulsating-sphere icospheres and closed horn bodies.

The reference ships no mesh files and builds its fixtures procedurally; these
generators produce the workloads of the benchmark configurations:

* ``icosphere(radius, subdivisions)`` -- closed sphere, outward winding
  (subdivision 3: 642 vertices, 1280 triangles).
* ``horn_body(...)`` -- closed body of revolution: a flat throat disc (tag 2)
  feeding a flared horn wall, a cylindrical outer shell back to a flat rear
  disc ``rear_depth`` behind the throat (tag 1).  Flares: exponential,
  conical, tractrix-like.  ``quarter=True`` keeps the x >= 0, y >= 0 quadrant
  for a ``yz+xz`` mirror-symmetric solve.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from numpy.typing import NDArray

from .mesh import LoadedMesh, loaded_mesh_from_arrays


def icosphere(radius: float = 0.1, subdivisions: int = 3) -> tuple[NDArray[np.float64], NDArray[np.int32]]:
    """Icosahedron refined ``subdivisions`` times by edge midpoints and
    projected onto the sphere; triangles wound outward."""
    t = (1.0 + 5.0**0.5) / 2.0
    verts = [
        (-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
        (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
        (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1),
    ]
    faces = [
        (0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
        (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
        (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
        (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1),
    ]
    v = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in verts]
    for _ in range(subdivisions):
        cache: dict[tuple[int, int], int] = {}

        def mid(a: int, b: int) -> int:
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[key] = len(v) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(v) * radius
    F = np.array(faces, dtype=np.int32)
    # outward: normal . centroid > 0
    c = V[F].mean(axis=1)
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = np.einsum("ij,ij->i", n, c) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F


def sphere_mesh(radius: float = 0.1, subdivisions: int = 3, tag: int = 2) -> LoadedMesh:
    V, F = icosphere(radius, subdivisions)
    return loaded_mesh_from_arrays(V, F, np.full(F.shape[0], tag, np.int32))


# ---------------------------------------------------------------------------
# Horn bodies
# ---------------------------------------------------------------------------


@dataclass(frozen=True)
class HornSpec:
    throat_radius: float = 0.0127
    mouth_radius: float = 0.150
    length: float = 0.150
    rear_depth: float = 0.020
    flare: str = "exponential"  # "exponential" | "conical" | "tractrix"


def _flare_radius(spec: HornSpec, z: NDArray[np.float64]) -> NDArray[np.float64]:
    r0, R, L = spec.throat_radius, spec.mouth_radius, spec.length
    u = np.clip(z / L, 0.0, 1.0)
    if spec.flare == "exponential":
        return r0 * np.exp(math.log(R / r0) * u)
    if spec.flare == "conical":
        return r0 + (R - r0) * u
    if spec.flare == "tractrix":
        # tractrix-like: fast opening near the mouth, r = r0 + (R - r0) * (1 - sqrt(1 - u^2))^0.7
        return r0 + (R - r0) * np.power(1.0 - np.sqrt(np.clip(1.0 - u * u, 0.0, 1.0)), 0.7)
    raise ValueError(f"unknown flare {spec.flare!r}")


def throat_edge(spec: HornSpec, h: float) -> float:
    """Edge length on the throat disc: at least three segments across it."""
    return min(h, spec.throat_radius / 3.0)


def _profile(spec: HornSpec, h: float) -> tuple[list[NDArray[np.float64]], list[str]]:
    """Profile pieces in the (r, z) half-plane, each discretised into nodes of
    spacing ~h.  Traversal: axis at the throat -> throat rim -> flare -> mouth
    rim -> down the outer shell -> rear disc -> axis."""
    r0, R, L, d = spec.throat_radius, spec.mouth_radius, spec.length, spec.rear_depth

    def sample(fn, t_dense):
        pts = fn(t_dense)
        seg = np.linalg.norm(np.diff(pts, axis=0), axis=1)
        s = np.concatenate([[0.0], np.cumsum(seg)])
        n = max(1, int(round(s[-1] / h)))
        ts = np.interp(np.linspace(0.0, s[-1], n + 1), s, t_dense)
        return fn(ts)

    dense = np.linspace(0.0, 1.0, 4001)
    h_outer = h
    h = min(h_outer, throat_edge(spec, h_outer))
    throat = sample(lambda t: np.stack([r0 * t, np.zeros_like(t)], axis=1), dense)
    h = h_outer
    flare = sample(lambda t: np.stack([_flare_radius(spec, t * L), t * L], axis=1), dense)
    shell = sample(lambda t: np.stack([np.full_like(t, R), L - t * (L + d)], axis=1), dense)
    rear = sample(lambda t: np.stack([R * (1.0 - t), np.full_like(t, -d)], axis=1), dense)
    return [throat, flare, shell, rear], ["throat", "flare", "shell", "rear"]


def _ring(r: float, h: float, quarter: bool, min_seg: int) -> NDArray[np.float64]:
    """Angles of one ring (quarter: both plane angles included)."""
    if r <= 1e-12:
        return np.array([0.0])
    span = 0.5 * math.pi if quarter else 2.0 * math.pi
    n = max(min_seg, int(round(span * r / h)))
    if quarter:
        return np.linspace(0.0, span, n + 1)
    return np.linspace(0.0, span, n, endpoint=False)


def horn_body(
    spec: HornSpec = HornSpec(),
    *,
    edge_length: float = 0.008,
    quarter: bool = True,
    jitter: float = 0.05,
    seed: int = 0,
) -> tuple[NDArray[np.float64], NDArray[np.int32], NDArray[np.int32]]:
    """Closed horn body; returns (vertices (N,3), triangles (M,3), tags (M,))."""
    h = float(edge_length)
    rng = np.random.default_rng(seed)
    pieces, names = _profile(spec, h)
    min_seg = 1 if quarter else 3
    verts: list[NDArray[np.float64]] = []
    tris: list[tuple[int, int, int]] = []
    tags: list[int] = []
    # profile nodes with piece bookkeeping; corners are shared between pieces
    nodes: list[tuple[float, float]] = []
    node_piece: list[int] = []
    for pi, pts in enumerate(pieces):
        for j, (r, z) in enumerate(pts):
            if pi > 0 and j == 0:
                continue  # corner already emitted by the previous piece
            nodes.append((float(r), float(z)))
            node_piece.append(pi)
    corner = set()
    acc = 0
    for pi, pts in enumerate(pieces):
        corner.add(acc)
        acc += len(pts) - 1
    corner.add(len(nodes) - 1)
    rings: list[tuple[NDArray[np.int64], NDArray[np.float64]]] = []
    h_throat = throat_edge(spec, h)
    for k, (r, z) in enumerate(nodes):
        ang = _ring(r, h_throat if node_piece[k] == 0 else h, quarter, min_seg)
        ids = []
        for a in ang:
            on_plane = quarter and (abs(a) < 1e-12 or abs(a - 0.5 * math.pi) < 1e-12)
            aa = a
            rr, zz = r, z
            if jitter > 0 and r > 1e-12:
                if not on_plane:
                    aa = a + rng.uniform(-jitter, jitter) * h / r
                if k not in corner:
                    # move along the local profile tangent
                    pr, pz = nodes[k - 1]
                    nr, nz = nodes[k + 1]
                    tr, tz = nr - pr, nz - pz
                    tn = math.hypot(tr, tz)
                    du = rng.uniform(-jitter, jitter) * h / tn
                    rr, zz = r + du * tr, z + du * tz
            ids.append(len(verts))
            verts.append(np.array([rr * math.cos(aa), rr * math.sin(aa), zz]))
        rings.append((np.array(ids, dtype=np.int64), ang))

    def tag_of(k: int) -> int:
        return 2 if node_piece[k + 1] == 0 else 1

    V = np.array(verts)
    for k in range(len(nodes) - 1):
        ia, aa = rings[k]
        ib, ab = rings[k + 1]
        tg = tag_of(k)
        local: list[tuple[int, int, int]] = []
        if ia.size == 1 or ib.size == 1:
            c, ring, ang = (ia[0], ib, ab) if ia.size == 1 else (ib[0], ia, aa)
            m = ring.size
            for j in range(m - 1 if quarter else m):
                local.append((c, ring[j], ring[(j + 1) % m]))
        else:
            # merge the two angle sequences into a triangle strip
            A = list(zip(ia, aa))
            Bq = list(zip(ib, ab))
            if not quarter:
                A.append((ia[0], aa[0] + 2 * math.pi))
                Bq.append((ib[0], ab[0] + 2 * math.pi))
            i = j = 0
            while i < len(A) - 1 or j < len(Bq) - 1:
                adv_a = j >= len(Bq) - 1 or (i < len(A) - 1 and A[i + 1][1] <= Bq[j + 1][1])
                if adv_a:
                    local.append((A[i][0], Bq[j][0], A[i + 1][0]))
                    i += 1
                else:
                    local.append((A[i][0], Bq[j][0], Bq[j + 1][0]))
                    j += 1
        # orient against the analytic outward normal of this profile segment
        (r1, z1), (r2, z2) = nodes[k], nodes[k + 1]
        nr_, nz_ = -(z2 - z1), (r2 - r1)
        for t in local:
            p = V[list(t)]
            gn = np.cross(p[1] - p[0], p[2] - p[0])
            c = p.mean(axis=0)
            phi = math.atan2(c[1], c[0])
            out = np.array([nr_ * math.cos(phi), nr_ * math.sin(phi), nz_])
            if float(gn @ out) < 0:
                t = (t[0], t[2], t[1])
            tris.append(t)
            tags.append(tg)
    F = np.array(tris, dtype=np.int32)
    # drop unreferenced vertices (none expected) and snap plane coordinates
    if quarter:
        V[np.abs(V[:, 0]) < 1e-12, 0] = 0.0
        V[np.abs(V[:, 1]) < 1e-12, 1] = 0.0
    return V, F, np.array(tags, dtype=np.int32)


def horn_mesh(
    spec: HornSpec = HornSpec(),
    *,
    edge_length: float = 0.008,
    quarter: bool = True,
    jitter: float = 0.05,
    seed: int = 0,
) -> LoadedMesh:
    V, F, T = horn_body(spec, edge_length=edge_length, quarter=quarter, jitter=jitter, seed=seed)
    return loaded_mesh_from_arrays(
        V, F, T, physical_groups={1: "wall", 2: "throat"},
        native_symmetry_plane="yz+xz" if quarter else None,
    )


def edge_length_for_dofs(
    spec: HornSpec, target_dofs: int, *, quarter: bool = True, seed: int = 0, rtol: float = 0.002,
    bracket: tuple[float, float] = (1e-4, 0.2),
) -> float:
    """Edge length whose mesh has about ``target_dofs`` P1 dofs (bisection in
    log h over ``bracket`` down to a relative width of ``rtol``;
    deterministic)."""
    lo, hi = bracket
    for _ in range(40):
        mid = math.sqrt(lo * hi)
        V, _, _ = horn_body(spec, edge_length=mid, quarter=quarter, jitter=0.0, seed=seed)
        if V.shape[0] > target_dofs:
            lo = mid
        else:
            hi = mid
        if hi / lo < 1.0 + rtol:
            break
    return math.sqrt(lo * hi)


def mirror_quarter_to_full(
    vertices: NDArray[np.float64], triangles: NDArray[np.int32], tags: NDArray[np.int32],
    tol: float = 1e-9,
) -> tuple[NDArray[np.float64], NDArray[np.int32], NDArray[np.int32]]:
    """Full body from an x >= 0, y >= 0 quarter by the three mirror images;
    winding reversed on single reflections so normals stay outward."""
    key_map: dict[tuple[int, int, int], int] = {}
    out_v: list[NDArray[np.float64]] = []
    out_t, out_g = [], []
    for sx, sy in ((1, 1), (-1, 1), (1, -1), (-1, -1)):
        remap = np.empty(vertices.shape[0], dtype=np.int64)
        for i, p in enumerate(vertices):
            q = np.array([sx * p[0], sy * p[1], p[2]])
            key = tuple(np.round(q / tol).astype(np.int64).tolist())
            if key not in key_map:
                key_map[key] = len(out_v)
                out_v.append(q)
            remap[i] = key_map[key]
        t = remap[triangles]
        if sx * sy < 0:
            t = t[:, [0, 2, 1]]
        out_t.append(t)
        out_g.append(tags)
    return np.array(out_v), np.vstack(out_t).astype(np.int32), np.concatenate(out_g).astype(np.int32)
