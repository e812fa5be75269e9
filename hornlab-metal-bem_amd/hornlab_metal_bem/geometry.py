"""Validated geometry buffers handed to libhbem.

Same data contract as the reference's MetalGeometryBuffers
(hornlab_metal_bem/metal/geometry.py:36-213): zero-based int32 indices,
float32 SoA vertices (3, N) with |x| <= 1e-6 snapped to 0, plane-major
triangles (3, M), P1 local-to-global (M, 3), float32 areas and unit normals
computed in float64 from the snapped vertices.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np
from numpy.typing import NDArray

from .config import NATIVE_SYMMETRY_PLANES
from .mesh import open_boundary_edges

# Vertices this close to a coordinate plane are snapped onto it so mirrored
# images match the 1e-6 coordinate keys of the singular-pair search.
_PLANE_SNAP_TOLERANCE = 1.0e-6

SYMMETRY_CODES = {None: 0, "yz": 1, "xz": 2, "yz+xz": 3, "xy": 4}


class GeometryError(ValueError):
    """Grid/space metadata violates the geometry contract."""


# Name used by the reference for the same error.
MetalGeometryError = GeometryError


@dataclass(frozen=True)
class GeometryBuffers:
    vertices_3xn_f32: NDArray[np.float32]
    triangles_3xm_i32: NDArray[np.int32]
    physical_tags_i32: NDArray[np.int32]
    p1_local2global_i32: NDArray[np.int32]
    triangle_areas_f32: NDArray[np.float32]
    triangle_normals_3xm_f32: NDArray[np.float32]
    p1_dof_count: int
    dp0_dof_count: int

    @property
    def n_vertices(self) -> int:
        return int(self.vertices_3xn_f32.shape[1])

    @property
    def n_triangles(self) -> int:
        return int(self.triangles_3xm_i32.shape[1])

    @property
    def triangles_nx3_i32(self) -> NDArray[np.int32]:
        return np.ascontiguousarray(self.triangles_3xm_i32.T)


MetalGeometryBuffers = GeometryBuffers


def _int32(name: str, value: Any) -> NDArray[np.int32]:
    arr = np.asarray(value)
    if not np.issubdtype(arr.dtype, np.integer):
        raise GeometryError(f"{name} must contain integer values")
    if arr.size == 0:
        raise GeometryError(f"{name} must not be empty")
    info = np.iinfo(np.int32)
    if int(arr.min()) < info.min or int(arr.max()) > info.max:
        raise GeometryError(f"{name} values must fit int32")
    return np.asarray(arr, dtype=np.int32)


def _zero_based(name: str, idx: NDArray[np.int32], upper: int | None = None) -> None:
    lo, hi = int(idx.min()), int(idx.max())
    if lo < 0:
        raise GeometryError(f"{name} indices must be zero-based and nonnegative")
    if lo != 0:
        raise GeometryError(
            f"{name} indices must be compact zero-based arrays with minimum 0; got minimum {lo}"
        )
    if upper is not None and hi >= upper:
        raise GeometryError(f"{name} indices must be zero-based and less than {upper}; got max {hi}")


def build_geometry_buffers(grid: Any, physical_tags: Any, p1_space: Any, dp0_space: Any | None = None) -> GeometryBuffers:
    vertices = np.asarray(grid.vertices, dtype=np.float64)
    if vertices.ndim != 2 or vertices.shape[0] != 3 or vertices.shape[1] == 0:
        raise GeometryError("grid.vertices must have shape (3, n_vertices) with n_vertices > 0")
    if not np.all(np.isfinite(vertices)):
        raise GeometryError("grid.vertices must be finite")
    vertices = vertices.copy()
    vertices[np.abs(vertices) <= _PLANE_SNAP_TOLERANCE] = 0.0
    elements = np.asarray(grid.elements)
    if elements.ndim != 2 or elements.shape[0] != 3 or elements.shape[1] == 0:
        raise GeometryError("grid.elements must have shape (3, n_triangles) with n_triangles > 0")
    tris = _int32("grid.elements", elements)
    _zero_based("grid.elements", tris, upper=vertices.shape[1])
    rows = tris.T
    rep = (rows[:, 0] == rows[:, 1]) | (rows[:, 1] == rows[:, 2]) | (rows[:, 0] == rows[:, 2])
    if np.any(rep):
        raise GeometryError(f"grid.elements triangle {int(np.flatnonzero(rep)[0])} repeats a vertex index")
    n_tri = int(tris.shape[1])
    if hasattr(grid, "number_of_elements") and int(grid.number_of_elements) != n_tri:
        raise GeometryError(
            f"grid.number_of_elements={int(grid.number_of_elements)} does not match elements shape "
            f"with {n_tri} triangles"
        )
    tags = np.asarray(physical_tags)
    if tags.ndim == 2 and 1 in tags.shape:
        tags = tags.reshape(-1)
    if tags.shape != (n_tri,):
        raise GeometryError(
            f"physical_tags must have one value per triangle: expected {(n_tri,)}, got {tags.shape}"
        )
    tags = _int32("physical_tags", tags)
    if not hasattr(p1_space, "local2global"):
        raise GeometryError("p1_space must expose a local2global array")
    l2g_raw = np.asarray(p1_space.local2global)
    if l2g_raw.shape != (n_tri, 3):
        raise GeometryError(
            f"p1_space.local2global must have shape (n_triangles, 3): expected {(n_tri, 3)}, "
            f"got {l2g_raw.shape}"
        )
    l2g = _int32("p1_space.local2global", l2g_raw)
    _zero_based("p1_space.local2global", l2g)
    need = int(l2g.max()) + 1
    p1_count = int(getattr(p1_space, "global_dof_count", need))
    if p1_count < need:
        raise GeometryError(
            f"p1_space.global_dof_count is smaller than local2global requires: {p1_count} < {need}"
        )
    dp0_count = n_tri if dp0_space is None else int(getattr(dp0_space, "global_dof_count", n_tri))
    if dp0_count != n_tri:
        raise GeometryError(f"dp0_space.global_dof_count must equal n_triangles: {dp0_count} != {n_tri}")
    v = vertices.T
    p0, p1, p2 = v[rows[:, 0]], v[rows[:, 1]], v[rows[:, 2]]
    cross = np.cross(p1 - p0, p2 - p0)
    twice = np.linalg.norm(cross, axis=1)
    if np.any(twice <= 0.0):
        raise GeometryError(f"grid.elements triangle {int(np.flatnonzero(twice <= 0.0)[0])} has zero area")
    normals = cross / twice[:, None]
    return GeometryBuffers(
        vertices_3xn_f32=np.ascontiguousarray(vertices, dtype=np.float32),
        triangles_3xm_i32=np.ascontiguousarray(tris, dtype=np.int32),
        physical_tags_i32=np.ascontiguousarray(tags, dtype=np.int32),
        p1_local2global_i32=np.ascontiguousarray(l2g, dtype=np.int32),
        triangle_areas_f32=np.ascontiguousarray(0.5 * twice, dtype=np.float32),
        triangle_normals_3xm_f32=np.ascontiguousarray(normals.T, dtype=np.float32),
        p1_dof_count=p1_count,
        dp0_dof_count=dp0_count,
    )


# Reference name of the same builder.
build_metal_geometry_buffers = build_geometry_buffers


def validate_native_symmetry_plane(
    buffers: GeometryBuffers,
    symmetry_plane: str | None,
    *,
    tolerance: float = 1.0e-7,
    check_open_edges: bool = True,
) -> str | None:
    """Reduced-domain contract of a mirror-symmetric solve (reference:
    metal/geometry.py:143-213): mesh on the positive side of each requested
    plane, touching it, no triangle lying in it, and (optionally) every open
    edge on a requested plane."""
    if symmetry_plane is None:
        return None
    plane = str(symmetry_plane).strip().lower()
    if plane not in NATIVE_SYMMETRY_PLANES:
        raise GeometryError("native_symmetry_plane currently supports 'yz', 'xz', 'xy', and 'yz+xz'")
    coords = np.asarray(buffers.vertices_3xn_f32, dtype=np.float64)
    tris = buffers.triangles_nx3_i32
    used = np.unique(tris.reshape(-1))

    def check(component: int, name: str, label: str) -> None:
        vals = coords[component]
        lo = float(np.min(vals[used]))
        if lo < -tolerance:
            raise GeometryError(
                f"native_symmetry_plane={plane!r} requires a positive-{name} reduced-domain mesh; "
                f"minimum {name.upper()} is {lo:.6g}"
            )
        if not np.any(np.abs(vals[used]) <= tolerance):
            raise GeometryError(
                f"native_symmetry_plane={plane!r} requires boundary vertices on {name.upper()}=0"
            )
        flat = np.all(np.abs(vals[tris]) <= tolerance, axis=1)
        if np.any(flat):
            raise GeometryError(
                f"native_symmetry_plane={plane!r} treats {label} as an image plane, not a "
                f"physical boundary; triangle {int(np.flatnonzero(flat)[0])} lies entirely on the plane"
            )

    if "yz" in plane:
        check(0, "x", "X=0")
    if "xz" in plane:
        check(1, "y", "Y=0")
    if plane == "xy":
        check(2, "z", "Z=0")
    if check_open_edges:
        edges = open_boundary_edges(tris)
        if edges.size:
            comps = {"yz": [0], "xz": [1], "xy": [2], "yz+xz": [0, 1]}[plane]
            explained = np.zeros(edges.shape[0], dtype=bool)
            for c in comps:
                explained |= np.all(np.abs(coords[c][edges]) <= tolerance, axis=1)
            if not np.all(explained):
                bad = edges[int(np.flatnonzero(~explained)[0])]
                raise GeometryError(
                    f"native_symmetry_plane={plane!r} requires every open boundary edge on a "
                    f"requested symmetry plane; edge {bad.tolist()} is off-plane (pass "
                    "native_check_open_edges=False for open shells)"
                )
    return plane
