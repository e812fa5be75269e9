"""MI355X (gfx950) BEM solver behind the hornlab_metal_bem API.

Same public names and signatures as the reference package
(hornlab_metal_bem/__init__.py:53-252): ``native_config()``, ``solve()``,
``solve_frequencies()``, ``solve_multi_source()``, ``load_mesh()`` and the
config/result types.  The per-frequency hot path (assembly with Duffy
corrections, complex LU, exterior field, surface reductions) runs in
libhbem on the GPU; there is no CPU fallback.
"""
from __future__ import annotations

from dataclasses import replace as _replace

import numpy as np

from .config import (
    AnnularProfile,
    AxialProfile,
    CallableProfile,
    NormalProfile,
    ObservationConfig,
    PerFaceProfile,
    SolveConfig,
    SourceMotion,
    SourceProfile,
    TaperProfile,
    VelocityMode,
)
from .mesh import LoadedMesh, MeshError, load_mesh, loaded_mesh_from_arrays
from .observation import ObservationFrame, infer_frame
from .result import MeshInfo, SolveResult

__all__ = [
    "native_config",
    "solve",
    "solve_frequencies",
    "solve_multi_source",
    "load_mesh",
    "loaded_mesh_from_arrays",
    "SolveConfig",
    "SolveResult",
    "ObservationConfig",
    "VelocityMode",
    "SourceMotion",
    "SourceProfile",
    "NormalProfile",
    "AxialProfile",
    "TaperProfile",
    "AnnularProfile",
    "PerFaceProfile",
    "CallableProfile",
    "LoadedMesh",
    "MeshInfo",
    "MeshError",
    "ObservationFrame",
]


def native_config(**overrides) -> SolveConfig:
    """The supported configuration of the GPU path (corrected assembly)."""
    values = {"metal_native_assembly_mode": "corrected"}
    values.update(overrides)
    return SolveConfig(**values)


def _resolve_mesh(mesh, config: SolveConfig) -> LoadedMesh:
    if isinstance(mesh, LoadedMesh):
        return mesh
    return load_mesh(
        mesh,
        scale=config.mesh_scale,
        validate=config.mesh_validate,
        merge_tol=config.mesh_merge_tol,
        repair_normals=config.mesh_repair_normals,
        native_symmetry_plane=config.native_symmetry_plane,
        aperture_tag=config.aperture_tag,
    )


def _resolve_frame(loaded: LoadedMesh, config: SolveConfig) -> ObservationFrame:
    if config.frame_override is not None:
        return config.frame_override
    return infer_frame(
        loaded.grid,
        loaded.physical_tags,
        source_tag=min(config.velocity_sources.keys(), default=2),
        origin_at=config.observation.origin,
        symmetry_plane=config.native_symmetry_plane,
    )


def solve(mesh, config: SolveConfig | None = None) -> SolveResult:
    """Frequency sweep of a mesh path or ``LoadedMesh``."""
    if config is None:
        config = native_config()
    from .sweep import _build_frequency_grid, run_sweep_native_metal, should_route_native_metal

    should_route_native_metal(config)
    loaded = _resolve_mesh(mesh, config)
    frame = _resolve_frame(loaded, config)
    return run_sweep_native_metal(loaded, _build_frequency_grid(config), frame, config)


def solve_frequencies(mesh, frequencies_hz, config: SolveConfig | None = None) -> SolveResult:
    """Solve at caller-ordered frequencies."""
    if config is None:
        config = native_config()
    from .sweep import run_sweep_native_metal, should_route_native_metal

    should_route_native_metal(config)
    loaded = _resolve_mesh(mesh, config)
    frame = _resolve_frame(loaded, config)
    return run_sweep_native_metal(loaded, np.asarray(frequencies_hz, np.float64), frame, config)


def solve_multi_source(mesh, sources, config: SolveConfig | None = None, frequencies_hz=None):
    """One SolveResult per velocity-source dict; the frame comes from the
    first source's lowest tag unless ``frame_override`` is set."""
    if config is None:
        config = native_config()
    if not sources:
        raise ValueError("sources must contain at least one velocity dict")
    from .sweep import (
        _build_frequency_grid,
        run_sweep_native_metal_multi_source,
        should_route_native_metal,
    )

    should_route_native_metal(config)
    loaded = _resolve_mesh(mesh, config)
    frame = _resolve_frame(loaded, _replace(config, velocity_sources=dict(sources[0])))
    freqs = _build_frequency_grid(config) if frequencies_hz is None else np.asarray(frequencies_hz, np.float64)
    return run_sweep_native_metal_multi_source(loaded, freqs, frame, config, sources)


def solve_circsym(*args, **kwargs):
    """The axisymmetric solver of the reference is not part of the GPU path."""
    raise NotImplementedError("solve_circsym is outside the HIP per-frequency path")


solve_circsym_frequencies = solve_circsym
