"""Observation frame inference and polar observation arcs.

Restates hornlab_metal_bem/observation.py:76-313 of the reference: the
forward axis comes from the area-summed source-tag normals (PCA fallback),
is projected into the symmetry subspace for reduced meshes, and is oriented
throat -> mouth by where the source sits along it; the mouth centre is the
mean of the vertices within 2% of the forward extreme.
"""
from __future__ import annotations

import logging
from dataclasses import dataclass

import numpy as np
from numpy.typing import NDArray

from .config import ObservationConfig

logger = logging.getLogger(__name__)


@dataclass
class ObservationFrame:
    """axis: throat -> mouth unit vector; origin: measurement origin;
    u/v: horizontal/vertical transverse unit vectors."""

    axis: NDArray[np.float64]
    origin: NDArray[np.float64]
    u: NDArray[np.float64]
    v: NDArray[np.float64]
    mouth_center: NDArray[np.float64]
    source_center: NDArray[np.float64]


def _principal_axis(vertices: NDArray[np.float64], center: NDArray[np.float64]) -> NDArray[np.float64]:
    d = vertices - center[None, :]
    if d.shape[0] == 0:
        return np.array([0.0, 1.0, 0.0])
    w, vecs = np.linalg.eigh(d.T @ d)
    axis = vecs[:, int(np.argmax(w))]
    nrm = float(np.linalg.norm(axis))
    return axis / nrm if nrm > 1e-12 else np.array([0.0, 1.0, 0.0])


def project_to_symmetry_subspace(vector, symmetry_plane: str | None) -> NDArray[np.float64]:
    """Zero the components normal to the requested mirror planes."""
    out = np.array(vector, dtype=np.float64, copy=True)
    if symmetry_plane is None:
        return out
    plane = str(symmetry_plane).strip().lower()
    if plane in ("yz", "yz+xz"):
        out[0] = 0.0
    if plane in ("xz", "yz+xz"):
        out[1] = 0.0
    if plane == "xy":
        out[2] = 0.0
    return out


def infer_frame(
    grid,
    physical_tags: NDArray[np.int32],
    source_tag: int = 2,
    origin_at: str = "mouth",
    symmetry_plane: str | None = None,
) -> ObservationFrame:
    vertices = np.array(grid.vertices.T, dtype=np.float64)
    elements = np.array(grid.elements.T, dtype=np.int64)
    mask = np.asarray(physical_tags) == source_tag
    if not np.any(mask):
        raise ValueError(f"No elements with tag {source_tag} in mesh")
    src = elements[mask]
    src = src[np.all((src >= 0) & (src < vertices.shape[0]), axis=1)]

    normal = None
    source_center = vertices.mean(axis=0)
    if src.shape[0] > 0:
        p0, p1, p2 = vertices[src[:, 0]], vertices[src[:, 1]], vertices[src[:, 2]]
        raw = np.cross(p1 - p0, p2 - p0)
        twice_area = np.linalg.norm(raw, axis=1)
        ok = twice_area > 1e-15
        if np.any(ok):
            raw, twice_area = raw[ok], twice_area[ok]
            centroids = (p0[ok] + p1[ok] + p2[ok]) / 3.0
            # align every normal with the first one so mixed winding adds up
            signs = np.sign(raw @ raw[0])
            signs[signs == 0] = 1.0
            total = (raw * signs[:, None]).sum(axis=0)
            tn = float(np.linalg.norm(total))
            if tn > 1e-12:
                normal = total / tn
                source_center = np.average(centroids, weights=twice_area, axis=0)
    from_tags = normal is not None
    if normal is None:
        normal = _principal_axis(vertices, source_center)
    if symmetry_plane is not None:
        proj = project_to_symmetry_subspace(normal, symmetry_plane)
        pn = float(np.linalg.norm(proj))
        if pn > 1e-12:
            normal = proj / pn

    proj_all = vertices @ normal
    s_proj = float(source_center @ normal)
    hi, lo = float(proj_all.max()), float(proj_all.min())
    span = hi - lo
    if span < 1e-12:
        axis = normal.copy()
    else:
        from_min = abs(s_proj - lo) / span
        from_max = abs(s_proj - hi) / span
        if from_tags and min(from_min, from_max) > 0.25:
            axis = normal.copy()  # enclosed geometry: trust the source normal
        elif from_min < from_max:
            axis = normal.copy()
        else:
            axis = -normal

    along = vertices @ axis
    cut = along.max() - 0.02 * (along.max() - along.min())
    mouth_center = vertices[along >= cut].mean(axis=0)

    ex, ey = np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0])
    ref = ex if abs(float(axis @ ex)) < 0.9 else ey
    u = ref - float(ref @ axis) * axis
    u /= np.linalg.norm(u)
    v = np.cross(axis, u)
    origin = mouth_center.copy() if origin_at == "mouth" else source_center.copy()
    if symmetry_plane is not None:
        origin = project_to_symmetry_subspace(origin, symmetry_plane)
    return ObservationFrame(
        axis=axis, origin=origin, u=u, v=v, mouth_center=mouth_center, source_center=source_center
    )


def build_observation_points(
    frame: ObservationFrame, config: ObservationConfig
) -> tuple[NDArray[np.float64], NDArray[np.float64]]:
    """Polar arcs (P, N, 3) around ``frame.origin`` and their angles (N,).
    theta = 0 is on-axis, 180 behind; ``custom_points`` bypasses the arcs."""
    if config.custom_points is not None:
        planes = []
        for plane in config.planes:
            if plane not in config.custom_points:
                raise ValueError(
                    f"custom_points missing plane {plane!r}; "
                    f"available: {list(config.custom_points.keys())}"
                )
            pts = np.asarray(config.custom_points[plane], dtype=np.float64)
            if pts.ndim != 2 or pts.shape[1] != 3:
                raise ValueError(f"custom_points[{plane!r}] must be (N, 3), got {pts.shape}")
            planes.append(pts)
        counts = [p.shape[0] for p in planes]
        if len(set(counts)) > 1:
            raise ValueError(
                "All custom_points planes must have the same number of points, got "
                f"{dict(zip(config.planes, counts))}"
            )
        angles = np.linspace(config.angle_min_deg, config.angle_max_deg, counts[0])
        return np.stack(planes, axis=0), angles

    angles = np.linspace(config.angle_min_deg, config.angle_max_deg, config.angle_count)
    rad = np.deg2rad(angles)
    planes = []
    for plane in config.planes:
        if plane == "horizontal":
            t = frame.u
        elif plane == "vertical":
            t = frame.v
        elif plane == "diagonal":
            t = (frame.u + frame.v) / np.sqrt(2.0)
        else:
            raise ValueError(f"Unknown plane: {plane!r}")
        planes.append(
            frame.origin[None, :]
            + config.distance_m * np.cos(rad)[:, None] * frame.axis[None, :]
            + config.distance_m * np.sin(rad)[:, None] * t[None, :]
        )
    return np.stack(planes, axis=0), angles
