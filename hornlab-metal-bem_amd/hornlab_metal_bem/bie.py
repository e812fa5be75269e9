"""Neumann data and surface reductions of the standard exterior BIE.

Restates hornlab_metal_bem/bie.py:154-434 of the reference:
  - per-face source profile multipliers (normal, axial piston, taper,
    annulus, explicit, callable);
  - DP0 Neumann coefficients q = i rho omega v_n (velocity mode) or
    q = i rho omega (w / (i omega)) = rho w (acceleration mode);
  - area-weighted P1 surface-pressure averages (host twins of the GPU
    reduction, used by tests).
"""
from __future__ import annotations

import numpy as np
from numpy.typing import NDArray

from .config import (
    AnnularProfile,
    AxialProfile,
    CallableProfile,
    NormalProfile,
    PerFaceProfile,
    SolveConfig,
    SourceMotion,
    TaperProfile,
    VelocityMode,
)


def _unit_axis(axis) -> NDArray[np.float64] | None:
    a = np.asarray(axis, dtype=np.float64).reshape(-1)
    if a.shape[0] != 3:
        return None
    n = float(np.linalg.norm(a))
    if not np.isfinite(n) or n <= 1e-12:
        return None
    return a / n


def _profile_for_tag(config: SolveConfig, tag: int):
    profiles = {int(t): p for t, p in (config.source_velocity_profiles or {}).items()}
    if tag in profiles:
        return profiles[tag]
    return AxialProfile() if config.source_motion == SourceMotion.AXIAL else NormalProfile()


def _axial_projection(raw, twice_area, idx, axis) -> NDArray[np.float64]:
    mags = twice_area[idx]
    unit = raw[idx] / np.where(mags > 1e-15, mags, 1.0)[:, None]
    proj = unit @ axis
    # orient the whole tag outward-positive by its area-weighted vote
    if float(proj @ mags) < 0.0:
        proj = -proj
    return proj


def _normalised_radius(centroids, idx, axis, center) -> NDArray[np.float64]:
    d = centroids[idx] - center[None, :]
    radial = np.linalg.norm(d - np.outer(d @ axis, axis), axis=1)
    rmax = float(radial.max()) if radial.size else 0.0
    if not np.isfinite(rmax) or rmax <= 1e-15:
        return np.zeros_like(radial)
    return np.clip(radial / rmax, 0.0, 1.0)


def _taper(t: NDArray[np.float64], profile: TaperProfile) -> NDArray[np.float64]:
    out = np.ones_like(t)
    sel = t > profile.start
    if np.any(sel):
        x = np.clip((t[sel] - profile.start) / (1.0 - profile.start), 0.0, 1.0)
        out[sel] = (1.0 - x) if profile.kind == "linear" else 0.5 * (1.0 + np.cos(np.pi * x))
    out[t >= 1.0] = 0.0
    return out


def _build_source_face_scale(grid, physical_tags, config: SolveConfig, axis, source_center):
    """Per-face velocity multiplier, or None when every source tag is a
    plain uniform-normal source (reference: bie.py:154-284)."""
    profiles = {int(t): p for t, p in (config.source_velocity_profiles or {}).items()}
    tags = sorted({int(t) for t in config.velocity_sources} | set(profiles))
    if not tags:
        return None
    eff = {t: _profile_for_tag(config, t) for t in tags}
    if all(isinstance(p, NormalProfile) for p in eff.values()):
        return None
    unit_axis = _unit_axis(axis)
    if unit_axis is None and all(isinstance(p, (NormalProfile, AxialProfile)) for p in eff.values()):
        return None
    if unit_axis is None and any(
        isinstance(p, (TaperProfile, AnnularProfile, CallableProfile)) for p in eff.values()
    ):
        raise ValueError("source velocity profiles require a non-degenerate axis")
    center = np.asarray(source_center, dtype=np.float64).reshape(-1)
    if center.shape[0] != 3:
        raise ValueError("source_center must have shape (3,)")
    verts = np.asarray(grid.vertices.T, dtype=np.float64)
    elems = np.asarray(grid.elements.T, dtype=np.int64)
    p0, p1, p2 = verts[elems[:, 0]], verts[elems[:, 1]], verts[elems[:, 2]]
    raw = np.cross(p1 - p0, p2 - p0)
    twice_area = np.linalg.norm(raw, axis=1)
    unit = raw / np.where(twice_area > 1e-15, twice_area, 1.0)[:, None]
    centroids = (p0 + p1 + p2) / 3.0
    tags_arr = np.asarray(physical_tags)
    scale = np.zeros(elems.shape[0], dtype=np.complex128)
    any_source = False
    complex_seen = False
    for tag in tags:
        idx = np.where(tags_arr == tag)[0]
        if idx.size == 0:
            continue
        any_source = True
        prof = eff[tag]
        if isinstance(prof, NormalProfile):
            vals = np.ones(idx.size)
        elif isinstance(prof, AxialProfile):
            vals = np.ones(idx.size) if unit_axis is None else _axial_projection(raw, twice_area, idx, unit_axis)
        elif isinstance(prof, TaperProfile):
            vals = _axial_projection(raw, twice_area, idx, unit_axis) * _taper(
                _normalised_radius(centroids, idx, unit_axis, center), prof
            )
        elif isinstance(prof, AnnularProfile):
            t = _normalised_radius(centroids, idx, unit_axis, center)
            vals = _axial_projection(raw, twice_area, idx, unit_axis) * (
                (t >= prof.r_inner) & (t <= prof.r_outer)
            ).astype(np.float64)
        elif isinstance(prof, PerFaceProfile):
            vals = np.asarray(prof.weights, dtype=np.complex128)
            if vals.ndim != 1 or vals.shape[0] != idx.size:
                raise ValueError(
                    f"PerFaceProfile.weights length must equal the number of faces for tag {tag}"
                )
            complex_seen |= bool(np.any(vals.imag != 0.0))
        elif isinstance(prof, CallableProfile):
            vals = np.asarray(
                prof.callback(centroids[idx], unit[idx], unit_axis.copy(), center.copy()),
                dtype=np.complex128,
            )
            if vals.ndim != 1 or vals.shape[0] != idx.size:
                raise ValueError(
                    f"CallableProfile.callback must return one weight per face for tag {tag}"
                )
            if not np.all(np.isfinite(vals)):
                raise ValueError("CallableProfile.callback returned non-finite weights")
            complex_seen |= bool(np.any(vals.imag != 0.0))
        else:  # pragma: no cover - rejected by SolveConfig validation
            raise ValueError("source_velocity_profiles values must be SourceProfile instances")
        scale[idx] = vals
    if not any_source:
        return None
    return scale if complex_seen else scale.real.copy()


def _build_driver_neumann_coeffs(
    dp0_space,
    physical_tags,
    omega: float,
    config: SolveConfig,
    dtype,
    impedance_tags: set[int] | None = None,
    axial_face_scale=None,
    source_face_scale=None,
) -> NDArray:
    """DP0 Neumann coefficients of the velocity sources at angular frequency
    ``omega`` (reference: bie.py:287-362).  Tags carrying a Robin condition
    are skipped."""
    if source_face_scale is not None:
        if axial_face_scale is not None:
            raise ValueError("pass only one of source_face_scale or axial_face_scale")
        axial_face_scale = source_face_scale
    coeffs = np.zeros(dp0_space.global_dof_count, dtype=dtype)
    rho = config.air_density
    f_hz = float(omega) / (2.0 * np.pi) if omega > 0 else 0.0
    sources = (
        config.velocity_source_callback(f_hz)
        if config.velocity_source_callback is not None
        else config.velocity_sources
    )
    if impedance_tags is None:
        impedance_tags = set(config.impedance_sources.keys())
        if config.impedance_source_callback is not None:
            impedance_tags |= {int(t) for t in config.impedance_source_callback(f_hz).keys()}
    tags_arr = np.asarray(physical_tags)
    for tag, weight in sources.items():
        if tag in impedance_tags:
            continue
        idx = np.where(tags_arr == tag)[0]
        if idx.size == 0:
            continue
        if axial_face_scale is None:
            v_n = weight
            if config.velocity_mode == VelocityMode.ACCELERATION:
                v_n = weight / (1j * omega) if omega > 0 else 0.0
            coeffs[idx] = 1j * rho * omega * v_n
        else:
            v_n = weight * axial_face_scale[idx]
            if config.velocity_mode == VelocityMode.ACCELERATION:
                v_n = v_n / (1j * omega) if omega > 0 else np.zeros_like(v_n)
            coeffs[idx] = 1j * rho * omega * v_n
    return coeffs


def _compute_impedance(grid, p_surface, physical_tags, p1_space, source_tag: int = 2) -> complex:
    """Area-weighted mean of the triangle-averaged P1 pressure on a tag."""
    return compute_surface_pressure_avg(grid, p_surface, physical_tags, p1_space, [source_tag])[
        source_tag
    ]


def compute_surface_pressure_avg(grid, p_surface, physical_tags, p1_space, tags) -> dict[int, complex]:
    areas = np.asarray(grid.volumes)
    coeffs = np.asarray(p_surface.coefficients)
    l2g = np.asarray(p1_space.local2global)
    tags_arr = np.asarray(physical_tags)
    out: dict[int, complex] = {}
    for tag in tags:
        idx = np.where(tags_arr == tag)[0]
        if idx.size == 0:
            out[tag] = 0.0 + 0.0j
            continue
        a = areas[idx]
        total = float(np.sum(a))
        if total < 1e-30:
            out[tag] = 0.0 + 0.0j
            continue
        out[tag] = complex(np.sum(coeffs[l2g[idx]].mean(axis=1) * a) / total)
    return out
