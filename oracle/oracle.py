"""Python front of the CPU oracle (hbem_oracle.c).  TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker / CPU baseline, never by the product path.  The dense solve uses
scipy's LAPACK (zgesv/cgesv, OpenBLAS), standing in for the reference's
Accelerate cgesv (main.swift:7834-7941) and zgesv (main.swift:7101-7185).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
BUILD = HERE / "_build"


class _Geom(C.Structure):
    _fields_ = [
        ("vertices", C.c_void_p),
        ("triangles", C.c_void_p),
        ("l2g", C.c_void_p),
        ("areas", C.c_void_p),
        ("normals", C.c_void_p),
        ("n_vertices", C.c_int32),
        ("n_triangles", C.c_int32),
        ("n_dofs", C.c_int32),
        ("symmetry", C.c_int32),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_libs: dict[int, C.CDLL] = {}


def lib(bits: int = 64) -> C.CDLL:
    if bits not in _libs:
        path = BUILD / f"libhbem_oracle{bits}.so"
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        L.hbo_duffy_rule.restype = C.c_int
        L.hbo_duffy_rule_size.restype = C.c_int
        L.hbo_duffy_pairs.restype = C.c_int
        L.hbo_near_pairs.restype = C.c_int
        L.hbo_apply_duffy.restype = C.c_int64
        L.hbo_apply_near.restype = C.c_int64
        L.hbo_pair_block.restype = C.c_int
        L.hbo_set_threads.restype = C.c_int
        real = C.c_double if bits == 64 else C.c_float
        L.hbo_kernel.argtypes = [real] * 8 + [C.c_void_p]
        L.hbo_assemble_regular.argtypes = [C.c_void_p, real, real] + [C.c_void_p] * 8
        L.hbo_apply_duffy.argtypes = [C.c_void_p, C.c_int, real, real] + [C.c_void_p] * 9
        L.hbo_assemble_subset.argtypes = [C.c_void_p, C.c_void_p, C.c_int, real, real] + [C.c_void_p] * 8
        L.hbo_assemble_subset.restype = C.c_int64
        L.hbo_chief_rows.argtypes = [C.c_void_p, C.c_int, C.c_void_p, real, real, C.c_int] + [C.c_void_p] * 8
        L.hbo_apply_near.argtypes = [C.c_void_p, C.c_int, C.c_double, real, real] + [C.c_void_p] * 8
        L.hbo_field.argtypes = [C.c_void_p, real] + [C.c_void_p] * 4 + [C.c_int] + [C.c_void_p] * 3
        L.hbo_pair_block.argtypes = [C.c_void_p] + [C.c_int] * 6 + [real, real, C.c_void_p, C.c_void_p]
        L.hbo_near_pairs.argtypes = [C.c_void_p, C.c_double, C.c_void_p, C.c_int]
        L.hbo_duffy_pairs.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.hbo_surface_average.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.hbo_duffy_rule.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 5
        L.hbo_rule6.argtypes = [C.c_void_p] * 3
        _libs[bits] = L
    return _libs[bits]


SYMMETRY = {None: 0, "yz": 1, "xz": 2, "yz+xz": 3, "xy": 4}


@dataclass
class Geometry:
    """Mesh arrays for the C side.  Vertices, areas and normals are kept in
    float64 (the caller's float32 buffers convert exactly) and handed to the
    64-bit build as they are, to the 32-bit build rounded to float32."""

    vertices: np.ndarray
    triangles: np.ndarray
    l2g: np.ndarray
    areas: np.ndarray
    normals: np.ndarray
    n_dofs: int
    symmetry: int
    tags: np.ndarray

    @classmethod
    def from_buffers(cls, b, symmetry_plane: str | None = None) -> "Geometry":
        return cls(
            np.ascontiguousarray(b.vertices_3xn_f32, np.float64),
            np.ascontiguousarray(b.triangles_3xm_i32, np.int32),
            np.ascontiguousarray(b.p1_local2global_i32, np.int32),
            np.ascontiguousarray(b.triangle_areas_f32, np.float64),
            np.ascontiguousarray(b.triangle_normals_3xm_f32, np.float64),
            int(b.p1_dof_count),
            SYMMETRY[symmetry_plane],
            np.ascontiguousarray(b.physical_tags_i32, np.int32),
        )

    @classmethod
    def from_vertices(cls, vertices, like: "Geometry") -> "Geometry":
        """Same topology as ``like`` with float64 vertex positions; areas and
        unit normals recomputed in float64 from the (v1 - v0) x (v2 - v0)
        winding (the geometry-buffer convention)."""
        v = np.ascontiguousarray(vertices, np.float64)
        tri = like.triangles
        p0, p1, p2 = v[:, tri[0]], v[:, tri[1]], v[:, tri[2]]
        cr = np.cross((p1 - p0).T, (p2 - p0).T).T
        tw = np.linalg.norm(cr, axis=0)
        return cls(v, like.triangles, like.l2g, np.ascontiguousarray(0.5 * tw),
                   np.ascontiguousarray(cr / tw), like.n_dofs, like.symmetry, like.tags)

    @property
    def n_triangles(self) -> int:
        return int(self.triangles.shape[1])

    def c(self, bits: int = 64) -> _Geom:
        if bits == 64:
            arrs = (self.vertices, self.areas, self.normals)
        else:
            arrs = (np.ascontiguousarray(self.vertices, np.float32),
                    np.ascontiguousarray(self.areas, np.float32),
                    np.ascontiguousarray(self.normals, np.float32))
        self._keep = arrs
        return _Geom(
            arrs[0].ctypes.data, self.triangles.ctypes.data, self.l2g.ctypes.data,
            arrs[1].ctypes.data, arrs[2].ctypes.data, int(self.vertices.shape[1]),
            self.n_triangles, self.n_dofs, self.symmetry,
        )


def _p(a):
    return None if a is None else a.ctypes.data


def _dt(bits):
    return np.float64 if bits == 64 else np.float32


def set_threads(n: int) -> int:
    r = lib(64).hbo_set_threads(int(n))
    lib(32).hbo_set_threads(int(n))
    return r


def rule6(bits=64):
    d = _dt(bits)
    x, y, w = np.zeros(6, d), np.zeros(6, d), np.zeros(6, d)
    lib(bits).hbo_rule6(_p(x), _p(y), _p(w))
    return x, y, w


def duffy_rule(kind: int, order: int = 4, bits=64):
    d = _dt(bits)
    n = lib(bits).hbo_duffy_rule_size(kind, order)
    arrs = [np.zeros(n, d) for _ in range(5)]
    got = lib(bits).hbo_duffy_rule(kind, order, *[_p(a) for a in arrs])
    assert got == n
    return arrs  # test x, test y, trial x, trial y, weight


def kernel(d, n, k, kim=0.0, bits=64):
    out = np.zeros(4, _dt(bits))
    lib(bits).hbo_kernel(*[float(v) for v in d], *[float(v) for v in n], float(k), float(kim), _p(out))
    return complex(out[0], out[1]), complex(out[2], out[3])


def duffy_pairs(geom: Geometry):
    counts = np.zeros(3, np.int64)
    g = geom.c(64)
    n = lib(64).hbo_duffy_pairs(C.byref(g), None, 0, _p(counts))
    pairs = np.zeros((max(n, 1), 9), np.int32)
    lib(64).hbo_duffy_pairs(C.byref(g), _p(pairs), n, _p(counts))
    return pairs[:n], {"coincident": int(counts[0]), "edge": int(counts[1]), "vertex": int(counts[2])}


def near_pairs(geom: Geometry, threshold: float):
    g = geom.c(64)
    n = lib(64).hbo_near_pairs(C.byref(g), float(threshold), None, 0)
    pairs = np.zeros((max(n, 1), 9), np.int32)
    lib(64).hbo_near_pairs(C.byref(g), float(threshold), _p(pairs), n)
    return pairs[:n]


def pair_block(geom: Geometry, test, trial, tm=0, sm=0, mode=0, order=4, k=1.0, kim=0.0, bits=64):
    d = _dt(bits)
    slp, dlp = np.zeros(6, d), np.zeros(18, d)
    g = geom.c(bits)
    rc = lib(bits).hbo_pair_block(C.byref(g), test, trial, tm, sm, mode, order, float(k), float(kim), _p(slp), _p(dlp))
    if rc != 0:
        raise ValueError("pair_block: invalid pair/mode")
    return slp[0::2] + 1j * slp[1::2], (dlp[0::2] + 1j * dlp[1::2]).reshape(3, 3)


def _split(q, n, d):
    if q is None:
        return None, None
    q = np.asarray(q, np.complex128).reshape(n)
    return np.ascontiguousarray(q.real, d), np.ascontiguousarray(q.imag, d)


def assemble(geom: Geometry, k: float, q=None, *, kim: float = 0.0, beta=None,
             duffy: bool = True, duffy_order: int = 4, near_level: int = 0,
             near_threshold: float = 1.5, bits: int = 64):
    """Corrected assembly: regular 6x6 on all pairs + Sauter-Schwab deltas on
    touching pairs (+ optional subdivided near-field deltas).  Returns
    (A (n,n) complex row-major, rhs (n,), stats)."""
    d = _dt(bits)
    n, T = geom.n_dofs, geom.n_triangles
    Are, Aim = np.zeros((n, n), d), np.zeros((n, n), d)
    rre, rim = np.zeros(n, d), np.zeros(n, d)
    qre, qim = _split(q if q is not None else np.zeros(T), T, d)
    bre, bim = _split(beta, T, d)
    g = geom.c(bits)
    L = lib(bits)
    L.hbo_assemble_regular(C.byref(g), k, kim, _p(qre), _p(qim), _p(bre), _p(bim), _p(Are), _p(Aim), _p(rre), _p(rim))
    stats = {}
    if duffy:
        counts = np.zeros(3, np.int64)
        u = L.hbo_apply_duffy(C.byref(g), duffy_order, k, kim, _p(qre), _p(qim), _p(bre), _p(bim),
                              _p(Are), _p(Aim), _p(rre), _p(rim), _p(counts))
        stats["planned_pairs"] = {"coincident": int(counts[0]), "edge": int(counts[1]), "vertex": int(counts[2])}
        stats["unique_triplets"] = int(u)
    if near_level > 0:
        stats["near_pairs"] = int(L.hbo_apply_near(C.byref(g), near_level, near_threshold, k, kim, _p(qre), _p(qim),
                                                   _p(bre), _p(bim), _p(Are), _p(Aim), _p(rre), _p(rim)))
    return Are + 1j * Aim, rre + 1j * rim, stats


def assemble_subset(geom: Geometry, sel, k: float, q=None, *, kim: float = 0.0, beta=None,
                    duffy_order: int = 4, bits: int = 64):
    """Part of the corrected system from pairs with the test or the trial
    triangle selected (``sel`` boolean per triangle): the entries that move
    with the selected triangles' vertices.  Returns (A, rhs) like assemble."""
    d = _dt(bits)
    n, T = geom.n_dofs, geom.n_triangles
    Are, Aim = np.zeros((n, n), d), np.zeros((n, n), d)
    rre, rim = np.zeros(n, d), np.zeros(n, d)
    qre, qim = _split(q if q is not None else np.zeros(T), T, d)
    bre, bim = _split(beta, T, d)
    s = np.ascontiguousarray(np.asarray(sel, bool), np.uint8)
    g = geom.c(bits)
    lib(bits).hbo_assemble_subset(C.byref(g), _p(s), duffy_order, k, kim, _p(qre), _p(qim), _p(bre),
                                  _p(bim), _p(Are), _p(Aim), _p(rre), _p(rim))
    return Are + 1j * Aim, rre + 1j * rim


def chief_rows(geom: Geometry, pts_3xm, k: float, q, *, kim: float = 0.0, beta=None, bits: int = 64):
    """CHIEF interior rows (C (m, n), d (n_src, m)); q is (T,) or (n_src, T)."""
    d_ = _dt(bits)
    n, T = geom.n_dofs, geom.n_triangles
    pts = np.ascontiguousarray(pts_3xm, np.float32)
    m = int(pts.shape[1])
    qq = np.atleast_2d(np.asarray(q, np.complex128))
    ns = qq.shape[0]
    qre, qim = np.ascontiguousarray(qq.real, d_), np.ascontiguousarray(qq.imag, d_)
    bre, bim = _split(beta, T, d_)
    Cre, Cim = np.zeros((m, n), d_), np.zeros((m, n), d_)
    dre, dim = np.zeros((ns, m), d_), np.zeros((ns, m), d_)
    g = geom.c(bits)
    lib(bits).hbo_chief_rows(C.byref(g), m, _p(pts), k, kim, ns, _p(qre), _p(qim), _p(bre), _p(bim),
                             _p(Cre), _p(Cim), _p(dre), _p(dim))
    return Cre + 1j * Cim, dre + 1j * dim


def chief_solve(A, rhs, Cm, d, weight: float = 1.0):
    """Least squares min ||A x - b||^2 + ||s (C x - d)||^2 in FP64 with
    s = weight ||A||_inf / ||C||_inf (solveDenseLeastSquaresZgels,
    MS:7220-7416); rhs (n,) or (n_src, n), d (m,) or (n_src, m).  Returns the
    solutions and the CHIEF residuals ||s (C x - d)||_2 / ||b||_2."""
    a_inf = np.abs(A).sum(axis=1).max()
    c_inf = np.abs(Cm).sum(axis=1).max()
    s = weight * (a_inf / c_inf if c_inf > 0 else 1.0)
    B = np.atleast_2d(rhs)
    D = np.atleast_2d(d)
    M = np.vstack([A, s * Cm])
    X = np.linalg.lstsq(M, np.vstack([B.T, s * D.T]), rcond=None)[0].T
    res = np.array([np.linalg.norm(s * (Cm @ x - dd)) / np.linalg.norm(b) for x, dd, b in zip(X, D, B)])
    return (X[0] if np.ndim(rhs) == 1 else X), res


def field(geom: Geometry, k: float, p, q, obs_3xn, bits: int = 64):
    d = _dt(bits)
    n, T = geom.n_dofs, geom.n_triangles
    pre, pim = _split(p, n, d)
    qre, qim = _split(q, T, d)
    obs = np.ascontiguousarray(obs_3xn, np.float32)
    m = int(obs.shape[1])
    ore, oim = np.zeros(m, d), np.zeros(m, d)
    g = geom.c(bits)
    lib(bits).hbo_field(C.byref(g), k, _p(pre), _p(pim), _p(qre), _p(qim), m, _p(obs), _p(ore), _p(oim))
    return ore + 1j * oim


def surface_average(geom: Geometry, p, tag: int, bits: int = 64):
    d = _dt(bits)
    pre, pim = _split(p, geom.n_dofs, d)
    out = np.zeros(3, np.float64)
    g = geom.c(bits)
    lib(bits).hbo_surface_average(C.byref(g), _p(geom.tags), int(tag), _p(pre), _p(pim), _p(out))
    return complex(out[0], out[1]), float(out[2])


def solve(A, b, bits: int = 64):
    """LAPACK solve: zgesv (FP64) or cgesv (FP32)."""
    from scipy.linalg import lapack

    if bits == 64:
        lu, piv, x, info = lapack.zgesv(np.asarray(A, np.complex128), np.asarray(b, np.complex128))
    else:
        lu, piv, x, info = lapack.cgesv(np.asarray(A, np.complex64), np.asarray(b, np.complex64))
    if info != 0:
        raise np.linalg.LinAlgError(f"gesv info={info}")
    return x
