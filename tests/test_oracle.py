"""Pins of the CPU oracle (oracle/hbem_oracle.c) against known answers of the
reference's tests, closed forms and mathematical identities.  CPU only."""
from __future__ import annotations

import math
from math import factorial

import numpy as np
import pytest

import oracle as O
from meshes import (
    horn_buffers,
    horn_quarter_and_full,
    mesh_buffers,
    rel_l2,
    sphere_buffers,
    tiny_two_triangles,
    tiny_yz_xz_full,
    tiny_yz_xz_quarter,
)


def _mono(a, b):
    """Integral of x^a y^b over the reference triangle."""
    return factorial(a) * factorial(b) / factorial(a + b + 2)


def test_rule6_integrates_degree_four_exactly():
    x, y, w = O.rule6()
    assert abs(w.sum() - 0.5) < 1e-14
    for a in range(5):
        for b in range(5 - a):
            assert abs(np.sum(w * x**a * y**b) - _mono(a, b)) < 1e-14


@pytest.mark.parametrize("kind,count", [(1, 1536), (2, 1280), (3, 512)])
def test_duffy_rule_counts_weights_and_exactness(kind, count):
    tx, ty, sx, sy, w = O.duffy_rule(kind)
    assert w.size == count
    assert abs(w.sum() - 0.25) < 1e-14
    for px, py in ((tx, ty), (sx, sy)):
        assert np.all((px >= 0) & (py >= 0) & (px + py <= 1 + 1e-14))
    # tensor products of polynomials on T x T up to total degree 4
    for a in range(5):
        for b in range(5 - a):
            for c in range(5 - a - b):
                for d in range(5 - a - b - c):
                    got = np.sum(w * tx**a * ty**b * sx**c * sy**d)
                    assert abs(got - _mono(a, b) * _mono(c, d)) < 1e-15


def test_kernel_matches_closed_form_and_normal_derivative():
    rng = np.random.default_rng(1)
    for _ in range(20):
        d = rng.standard_normal(3) * 0.3
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        k, kim = rng.uniform(0.5, 300.0), rng.uniform(0.0, 2.0)
        g, dl = O.kernel(d, n, k, kim)
        r = np.linalg.norm(d)
        kc = k + 1j * kim
        G = lambda dd: np.exp(1j * kc * np.linalg.norm(dd)) / (4 * np.pi * np.linalg.norm(dd))
        assert abs(g - G(d)) <= 1e-12 * abs(G(d))
        # dG/dn_y with r = y - x: derivative along n at y, central difference
        h = 1e-6 * r
        fd = (G(d + h * n) - G(d - h * n)) / (2 * h)
        assert abs(dl - fd) <= 1e-6 * abs(fd) + 1e-9 * abs(G(d)) / r
    assert O.kernel([0.0, 0.0, 0.0], [0, 0, 1.0], 3.0) == (0j, 0j)


def test_tiny_fixture_duffy_plan_matches_reference_known_answer():
    """test_metal_native.py:1544-1550 and :1809-1816."""
    g = O.Geometry.from_buffers(tiny_two_triangles())
    pairs, counts = O.duffy_pairs(g)
    assert counts == {"coincident": 2, "edge": 2, "vertex": 0}
    _, _, st = O.assemble(g, 1.8318326, np.array([1.0, 0.5j]))
    assert st["unique_triplets"] == 16
    assert 9 * len(pairs) == 36


def test_quarter_single_triangle_has_fifteen_image_pairs():
    """test_metal_native.py:1712-1731: image_adjacent_pairs == 15."""
    g = O.Geometry.from_buffers(tiny_yz_xz_quarter(), "yz+xz")
    pairs, _ = O.duffy_pairs(g)
    assert int(np.sum((pairs[:, 2] != 0) | (pairs[:, 3] != 0))) == 15


def test_tiny_quarter_matches_full_orbit_reduction():
    """Reference quarter-vs-full parity (test_metal_native.py:1700-1710),
    held here at FP64 level."""
    k = 1.8318326
    gq = O.Geometry.from_buffers(tiny_yz_xz_quarter(), "yz+xz")
    gf = O.Geometry.from_buffers(tiny_yz_xz_full())
    Aq, rq, _ = O.assemble(gq, k, np.array([1.0 + 0j]))
    Af, rf, _ = O.assemble(gf, k, np.ones(4, complex))
    orbits = [[0], [1, 3], [2, 4]]
    Ar = np.array([[Af[np.ix_(r, c)].sum() for c in orbits] for r in orbits])
    rr = np.array([rf[r].sum() for r in orbits])
    assert rel_l2(Aq, Ar) < 1e-12
    assert rel_l2(rq, rr) < 1e-12


def test_quarter_horn_matches_full_horn_orbit_reduction():
    bq, bf, (Vq, Vf) = horn_quarter_and_full(edge_length=0.035)
    k = float(np.float32(2 * math.pi * 3000.0 / 343.0))
    qq = np.where(bq.physical_tags_i32 == 2, 1.0 + 0j, 0.0)
    qf = np.where(bf.physical_tags_i32 == 2, 1.0 + 0j, 0.0)
    Aq, rq, _ = O.assemble(O.Geometry.from_buffers(bq, "yz+xz"), k, qq)
    Af, rf, _ = O.assemble(O.Geometry.from_buffers(bf), k, qf)
    key = lambda p: (round(abs(p[0]) * 1e7), round(abs(p[1]) * 1e7), round(p[2] * 1e7))
    orbit: dict = {}
    for j, p in enumerate(Vf):
        orbit.setdefault(key(p), []).append(j)
    rows = [orbit[key(p)] for p in Vq]
    Ar = np.array([[Af[np.ix_(r, c)].sum() for c in rows] for r in rows])
    rr = np.array([rf[r].sum() for r in rows])
    assert rel_l2(Aq, Ar) < 1e-12
    assert rel_l2(rq, rr) < 1e-12


def test_icosphere_plan_counts():
    g = O.Geometry.from_buffers(sphere_buffers(0.1, 3))
    _, counts = O.duffy_pairs(g)
    assert counts == {"coincident": 1280, "edge": 3840, "vertex": 11460}


def _mass_row_sums(g):
    m = np.zeros(g.n_dofs)
    np.add.at(m, g.l2g.reshape(-1), np.repeat(g.areas.astype(np.float64) / 3.0, 3))
    return m


def test_closed_surface_double_layer_identity():
    """Gauss: for a closed polyhedron and x on a face, int dG0/dn_y = -1/2,
    so at k = 0 every row of A = K - M/2 sums to -(row sum of M).  Pins the
    kernel sign, normals and the singular quadrature of the double layer."""
    g = O.Geometry.from_buffers(sphere_buffers(0.1, 2))
    A, _, _ = O.assemble(g, 0.0, None)
    assert rel_l2(A.sum(axis=1), -_mass_row_sums(g)) < 2e-5


def test_closed_horn_identity_improves_with_finer_quadrature():
    """The horn's zero-thickness mouth lip makes touching pairs fold sharply
    and puts non-touching triangles very close together; 4-point
    Sauter-Schwab and the 6x6 rule are then the accuracy limit.  Raising the
    singular order and adding the subdivided near-field pass
    (main.swift:3072-3175) together cut the identity error by > 10x."""
    g = O.Geometry.from_buffers(horn_buffers(0.03, quarter=False))
    m = _mass_row_sums(g)
    A0, _, _ = O.assemble(g, 0.0, None)
    A1, _, st = O.assemble(g, 0.0, None, duffy_order=8, near_level=2, near_threshold=1.5)
    e0 = rel_l2(A0.sum(axis=1), -m)
    e1 = rel_l2(A1.sum(axis=1), -m)
    assert st["near_pairs"] > 0
    assert e0 < 1e-3
    assert e1 < 0.1 * e0


def _inplane_potential(x, tri):
    """int_T 1/|x - y| dA_y for x in the plane of T (closed form per edge)."""
    v0, v1, v2 = tri
    n = np.cross(v1 - v0, v2 - v0)
    n /= np.linalg.norm(n)
    total = 0.0
    for a, b in ((v0, v1), (v1, v2), (v2, v0)):
        t = (b - a) / np.linalg.norm(b - a)
        m = np.cross(t, n)  # outward edge normal
        d = float(np.dot(a - x, m))
        if abs(d) < 1e-15:
            continue
        sa, sb = float(np.dot(a - x, t)), float(np.dot(b - x, t))
        total += d * (np.arcsinh(sb / abs(d)) - np.arcsinh(sa / abs(d)))
    return total


def _subdivided_points(tri, level):
    x, y, w = O.rule6()
    subs = [np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])]
    for _ in range(level):
        nxt = []
        for a, b, c in subs:
            ab, bc, ca = (a + b) / 2, (b + c) / 2, (c + a) / 2
            nxt += [np.array([a, ab, ca]), np.array([ab, b, bc]), np.array([ca, bc, c]), np.array([ab, bc, ca])]
        subs = nxt
    pts, wts = [], []
    area2 = np.linalg.norm(np.cross(tri[1] - tri[0], tri[2] - tri[0]))
    for a, b, c in subs:
        det = abs((b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]))
        for i in range(6):
            r = a + x[i] * (b - a) + y[i] * (c - a)
            pts.append(tri[0] + r[0] * (tri[1] - tri[0]) + r[1] * (tri[2] - tri[0]))
            wts.append(w[i] * det * area2)
    return np.array(pts), np.array(wts)


@pytest.mark.parametrize("config", ["coincident", "edge", "vertex"])
def test_singular_single_layer_matches_semi_analytic(config):
    """Coplanar touching pairs at k = 0: int int 1/(4 pi r) from the
    Sauter-Schwab blocks vs the closed-form inner integral integrated by a
    fine subdivided rule."""
    v = {
        "coincident": ([[0, 0, 0], [1, 0, 0], [0.3, 0.8, 0]], [[0, 1, 2], [0, 1, 2]]),
        "edge": ([[0, 0, 0], [1, 0, 0], [0.3, 0.8, 0], [0.6, -0.7, 0]], [[0, 1, 2], [1, 0, 3]]),
        "vertex": ([[0, 0, 0], [1, 0, 0], [0.3, 0.8, 0], [-0.9, -0.2, 0], [-0.4, -0.9, 0]], [[0, 1, 2], [0, 3, 4]]),
    }[config]
    V = np.array(v[0], float) * 0.05
    F = np.array(v[1] if config != "coincident" else [v[1][0]], np.int32)
    b = mesh_buffers(V, F, np.full(F.shape[0], 2, np.int32))
    g = O.Geometry.from_buffers(b)
    trial = 0 if config == "coincident" else 1
    slp, _ = O.pair_block(g, 0, trial, mode=1, k=0.0)
    duffy = slp.sum().real
    T = [np.array(b.vertices_3xn_f32, float).T[b.triangles_3xm_i32[:, t]] for t in (0, trial)]
    pts, wts = _subdivided_points(T[0], 5)
    ref = sum(w_ * _inplane_potential(p, T[1]) for p, w_ in zip(pts, wts)) / (4 * math.pi)
    assert abs(duffy - ref) / abs(ref) < 2e-4, (duffy, ref)


def test_regular_block_converges_to_subdivided_for_separated_pairs():
    b = horn_buffers(0.03, quarter=False)
    g = O.Geometry.from_buffers(b)
    cent = np.array(b.vertices_3xn_f32, float).T[b.triangles_3xm_i32.T].mean(axis=1)
    d = np.linalg.norm(cent - cent[0], axis=1)
    far = int(np.argmax(d))
    s0, d0 = O.pair_block(g, 0, far, mode=0, k=40.0)
    s3, d3 = O.pair_block(g, 0, far, mode=2, order=3, k=40.0)
    assert rel_l2(s0, s3) < 3e-5
    assert rel_l2(d0, d3) < 3e-5


def test_singular_rule_converges_with_gauss_order():
    """Touching, non-coplanar pairs of a closed horn: 4-point Sauter-Schwab
    blocks against the same construction with 12 Gauss points."""
    b = horn_buffers(0.03, quarter=False)
    g = O.Geometry.from_buffers(b)
    pairs, _ = O.duffy_pairs(g)
    rng = np.random.default_rng(0)
    for pi in rng.choice(len(pairs), 12, replace=False):
        t, s = int(pairs[pi, 0]), int(pairs[pi, 1])
        s4, d4 = O.pair_block(g, t, s, mode=1, order=4, k=60.0)
        s8, d8 = O.pair_block(g, t, s, mode=1, order=8, k=60.0)
        s16, d16 = O.pair_block(g, t, s, mode=1, order=16, k=60.0)
        # a wrong region map or remap gives O(1) errors and no convergence;
        # the 4-point rule's own single-layer error here is ~1e-3
        assert rel_l2(s4, s16) < 5e-3
        assert rel_l2(s8, s16) < 0.5 * rel_l2(s4, s16) + 1e-12
        dscale = np.linalg.norm(d16) + np.linalg.norm(s16) / 0.03
        assert np.linalg.norm(d8 - d16) < 0.5 * np.linalg.norm(d4 - d16) + 1e-9 * dscale


def test_pulsating_sphere_matches_analytic_monopole():
    """q = i rho omega U, p(r) = rho c U (a/r) ika/(ika-1) e^{ik(r-a)}."""
    a, rho, c, U = 0.1, 1.21, 343.0, 1.0
    ang = np.deg2rad(np.linspace(0, 180, 7))
    obs = np.stack([2 * np.sin(ang), 0 * ang, 2 * np.cos(ang)]).astype(np.float32)
    errs = []
    for sub in (2, 3):
        b = sphere_buffers(a, sub)
        g = O.Geometry.from_buffers(b)
        f = 600.0
        k = 2 * np.pi * f / c
        q = np.full(b.n_triangles, 1j * rho * 2 * np.pi * f * U)
        A, rhs, _ = O.assemble(g, k, q)
        p = O.solve(A, rhs)
        fo = O.field(g, k, p, q, obs)
        ka = k * a
        exact = rho * c * U * (1j * ka) / (1j * ka - 1) * (a / 2.0) * np.exp(1j * k * (2.0 - a))
        errs.append(np.max(np.abs(fo - exact)) / abs(exact))
        surf = rho * c * U * (1j * ka) / (1j * ka - 1)
        avg, area = O.surface_average(g, p, 2)
        assert abs(avg - surf) / abs(surf) < 0.05
    assert errs[1] < 0.02 and errs[1] < errs[0]


def test_surface_average_known_answers():
    """Uniform pressure averages to itself; the tiny fixture's tag-2 mean is
    (p0 + p2 + p3)/3 (test_metal_native.py:3347-3356)."""
    g = O.Geometry.from_buffers(tiny_two_triangles())
    p = np.array([1 + 2j, -3 + 0.5j, 0.25 - 1j, 2 + 0j])
    avg, _ = O.surface_average(g, p, 2)
    assert abs(avg - (p[0] + p[2] + p[3]) / 3) < 1e-14
    avg1, _ = O.surface_average(g, np.full(4, 0.7 - 0.2j), 1)
    assert abs(avg1 - (0.7 - 0.2j)) < 1e-14
    assert O.surface_average(g, p, 9) == (0j, 0.0)


def test_fp32_port_tracks_fp64_oracle():
    b = sphere_buffers(0.1, 2)
    k = float(np.float32(2 * math.pi * 1500.0 / 343.0))
    q = np.full(b.n_triangles, 1j * 1000.0)
    A64, r64, _ = O.assemble(O.Geometry.from_buffers(b), k, q, bits=64)
    A32, r32, _ = O.assemble(O.Geometry.from_buffers(b), k, q, bits=32)
    assert rel_l2(A32, A64) < 1e-5
    assert rel_l2(r32, r64) < 1e-5


def test_near_field_subdivision_moves_toward_fine_reference():
    """Near (non-touching, close) pairs: level-1 subdivision is closer to a
    level-3 reference than the plain 6x6 rule (test_metal_native.py:1882-1949)."""
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 0.05], [1, 0, 0.05], [0, 1, 0.05]], float)
    F = np.array([[0, 1, 2], [3, 5, 4]], np.int32)
    b = mesh_buffers(V, F, np.array([1, 2], np.int32))
    g = O.Geometry.from_buffers(b)
    assert len(O.near_pairs(g, 1.5)) == 2
    s0, d0 = O.pair_block(g, 0, 1, mode=0, k=2.0)
    s1, d1 = O.pair_block(g, 0, 1, mode=2, order=1, k=2.0)
    s3, d3 = O.pair_block(g, 0, 1, mode=2, order=3, k=2.0)
    assert rel_l2(s1, s3) < rel_l2(s0, s3)
    assert rel_l2(d1, d3) < rel_l2(d0, d3)


def test_singular_mirror_classes_are_invariant():
    """The GPU evaluates each class of touching combinations once: members
    (tm, sm) with equal tm^sm, equal locals and, for coincident and edge
    pairs, equal mask parity give the same block (both triangles moved by one
    isometry); vertex pairs are parity-free.  Coincident and edge pairs of
    opposite parity differ at quadrature level, which shows the check can
    fail.  FP64 oracle blocks, quarter horn."""
    buf = horn_buffers(edge_length=0.03)
    g = O.Geometry.from_buffers(buf, "yz+xz")
    pairs, _ = O.duffy_pairs(g)
    k = 150.0
    same = {1: 0.0, 2: 0.0, 3: 0.0}
    cross = {1: 0.0, 2: 0.0}
    rng = np.random.default_rng(3)
    for idx in rng.choice(len(pairs), 1500, replace=False):
        t, s, tm, sm, kind = (int(v) for v in pairs[idx, :5])
        if tm == 0:
            continue
        odd = bin(tm).count("1") % 2
        # same-parity partner: tm' = tm ^ 3 keeps the parity, tm' = 0 flips odd ones
        partners = [(tm ^ 3, sm ^ 3)] if odd else [(0, tm ^ sm)]
        if not odd or kind == 3:
            partners.append((0, tm ^ sm))
        a_slp, a_dlp = O.pair_block(g, t, s, tm, sm, mode=1, k=k)
        scale = np.abs(a_slp).max() + np.abs(a_dlp).max()
        for ptm, psm in partners:
            b_slp, b_dlp = O.pair_block(g, t, s, ptm, psm, mode=1, k=k)
            e = max(np.abs(a_slp - b_slp).max(), np.abs(a_dlp - b_dlp).max()) / scale
            same[kind] = max(same[kind], e)
        if odd and kind != 3:
            b_slp, b_dlp = O.pair_block(g, t, s, 0, tm ^ sm, mode=1, k=k)
            e = max(np.abs(a_slp - b_slp).max(), np.abs(a_dlp - b_dlp).max()) / scale
            cross[kind] = max(cross[kind], e)
    assert max(same.values()) < 1e-10, same
    assert cross[1] > 1e-8 and cross[2] > 1e-8, cross


def test_robin_sphere_matches_closed_form():
    """Uniform Neumann data q with a Robin wall on the whole sphere:
    dp/dn = q + i k beta p (neumannWithRobin, main.swift:1646-1674).  The
    monopole relation p(a) = a / (i k a - 1) dp/dn gives
    p(a) = a q / (i k a - 1 - i k beta a), and p(r) = p(a) (a / r) e^{ik(r-a)}.
    The oracle's Robin fold -(i k beta / 3) V and the field's
    q + i k beta mean(p) must reproduce it (beta = 0 is the monopole test)."""
    a, rho, c, U = 0.1, 1.21, 343.0, 1.0
    beta = 0.4 + 0.3j
    ang = np.deg2rad(np.linspace(0, 180, 7))
    obs = np.stack([2 * np.sin(ang), 0 * ang, 2 * np.cos(ang)]).astype(np.float32)
    errs = []
    for sub in (2, 3):
        b = sphere_buffers(a, sub)
        g = O.Geometry.from_buffers(b)
        f = 600.0
        k = 2 * np.pi * f / c
        q = np.full(b.n_triangles, 1j * rho * 2 * np.pi * f * U)
        bt = np.full(b.n_triangles, beta)
        A, rhs, _ = O.assemble(g, k, q, beta=bt)
        p = O.solve(A, rhs)
        l2g = np.asarray(b.p1_local2global_i32).reshape(-1, 3)
        q_tot = q + 1j * k * bt * p[l2g].mean(axis=1)
        fo = O.field(g, k, p, q_tot, obs)
        pa = a * q[0] / (1j * k * a - 1 - 1j * k * beta * a)
        exact = pa * (a / 2.0) * np.exp(1j * k * (2.0 - a))
        errs.append(np.max(np.abs(fo - exact)) / abs(exact))
        avg, _ = O.surface_average(g, p, 2)
        assert abs(avg - pa) / abs(pa) < 0.05
        # the wall changes the answer: beta = 0 differs by far more than the error
        pa0 = a * q[0] / (1j * k * a - 1)
        assert abs(pa - pa0) / abs(pa0) > 0.2
    assert errs[1] < 0.02 and errs[1] < errs[0]


def test_subset_assembly_is_the_part_that_moves_with_a_vertex():
    """assemble_subset (pairs touching the selected triangles) changes under a
    vertex move exactly as the full corrected assembly does: the local
    finite differences of the shape-gradient check rest on this."""
    b = horn_buffers(edge_length=0.04, quarter=True)
    g0 = O.Geometry.from_vertices(np.asarray(b.vertices_3xn_f32, np.float64),
                                  O.Geometry.from_buffers(b, "yz+xz"))
    k = 40.0
    q = np.where(b.physical_tags_i32 == 2, 1j * 100.0, 0.0)
    A, r, _ = O.assemble(g0, k, q)
    for v in (7, 33):
        sel = np.any(g0.triangles == v, axis=0)
        V = g0.vertices.copy()
        V[:, v] += [1e-4, -2e-4, 3e-4]
        g1 = O.Geometry.from_vertices(V, g0)
        A1, r1, _ = O.assemble(g1, k, q)
        As0, rs0 = O.assemble_subset(g0, sel, k, q)
        As1, rs1 = O.assemble_subset(g1, sel, k, q)
        assert np.abs((A1 - A) - (As1 - As0)).max() <= 1e-12 * np.abs(A1 - A).max()
        assert np.abs((r1 - r) - (rs1 - rs0)).max() <= 1e-12 * np.abs(r1 - r).max()
        assert np.abs(A1 - A).max() > 0


def test_chief_rows_and_least_squares_on_the_pulsating_sphere():
    """CHIEF (assembleChiefRows MS:1930-2036, zgels MS:7220-7416): the rows
    are the field operator at interior points (C p - d = field(p, q) for any
    p); the exact monopole satisfies the interior identity C p - d = 0 up to
    discretisation error; at the first interior resonance ka = pi the
    least-squares solution is closer to the monopole than the square solve,
    and away from it CHIEF changes nothing of note."""
    a, rho, c, U = 0.1, 1.21, 343.0, 1.0
    b = sphere_buffers(a, 2)
    g = O.Geometry.from_buffers(b)
    pts = (np.random.default_rng(0).uniform(-0.5, 0.5, (3, 6)) * a).astype(np.float32)
    for ka, gain in ((2.0, None), (math.pi, 2.0)):
        k = ka / a
        q = np.full(b.n_triangles, 1j * rho * k * c * U)
        pex = rho * c * U * (1j * ka / (1j * ka - 1)) * np.ones(b.p1_dof_count)
        A, r, _ = O.assemble(g, k, q)
        Cm, d = O.chief_rows(g, pts, k, q)
        pr = np.random.default_rng(1).standard_normal(b.p1_dof_count) + 0j
        f = O.field(g, k, pr, q, pts)
        assert rel_l2(Cm @ pr - d[0], f) < 1e-12
        assert np.linalg.norm(Cm @ pex - d[0]) < 1e-2 * np.linalg.norm(d[0])
        p = np.linalg.solve(A, r)
        x, res = O.chief_solve(A, r, Cm, d[0])
        e0, e1 = rel_l2(p, pex), rel_l2(x, pex)
        if gain is None:
            assert abs(e1 - e0) < 0.01 * e0
        else:
            assert e1 < e0 / gain, (e0, e1)
        assert res[0] < 1e-2
