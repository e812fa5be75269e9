"""Frequency sweep on the HIP path (reference: hornlab_metal_bem/sweep.py).

Builds the frequency grid, wavenumbers, observation points and DP0 Neumann
rows, runs the resident GPU batch (assemble -> LU -> field -> reductions) and
assembles the ``SolveResult`` with the reference's fields, directivity and
diagnostic policies.
"""
from __future__ import annotations

import logging
import math
import time
from typing import Any

import numpy as np
from numpy.typing import NDArray

from ._constants import REFERENCE_PRESSURE, SPEED_OF_SOUND
from .backends import AssemblyBackendUnavailable
from .bie import _build_driver_neumann_coeffs, _build_source_face_scale
from .config import NATIVE_SYMMETRY_PLANES, BIEFormulation, SolveConfig
from .geometry import build_geometry_buffers
from .mesh import LoadedMesh, make_pure_function_spaces
from .observation import ObservationFrame, build_observation_points
from .result import SolveResult

logger = logging.getLogger(__name__)


def _build_frequency_grid(config: SolveConfig) -> NDArray[np.float64]:
    if config.freq_spacing == "log":
        return np.geomspace(config.freq_min_hz, config.freq_max_hz, config.freq_count)
    return np.linspace(config.freq_min_hz, config.freq_max_hz, config.freq_count)


def should_route_native_metal(config: SolveConfig) -> bool:
    """Reject configurations the HIP path does not implement."""
    if config.native_symmetry_plane is not None and config.native_symmetry_plane not in NATIVE_SYMMETRY_PLANES:
        raise AssemblyBackendUnavailable(
            "native_symmetry_plane must be None or one of "
            + ", ".join(repr(p) for p in NATIVE_SYMMETRY_PLANES)
        )
    if config.aperture_tag is not None:
        raise AssemblyBackendUnavailable("aperture_tag (coupled infinite baffle) is not built on the HIP path")
    if config.metal_native_assembly_mode == "parity":
        raise AssemblyBackendUnavailable(
            "metal_native_assembly_mode='parity' cross-checks the reference's CPU assembly inside "
            "its helper; the HIP path has no CPU assembly to compare against"
        )
    return True


should_route_native = should_route_native_metal


def _k_values_for_native(frequencies, config: SolveConfig):
    """k = 2 pi f / c in float32; complex_k adds k_im = shift * k
    (reference: sweep.py:150-161)."""
    k_real = (2.0 * np.pi * np.asarray(frequencies, np.float64) / SPEED_OF_SOUND).astype(np.float32)
    if config.formulation == BIEFormulation.COMPLEX_K:
        k_imag = (k_real.astype(np.float64) * config.complex_k_shift).astype(np.float32)
    else:
        k_imag = np.zeros_like(k_real)
    return k_real, k_imag


def _active_impedance_sources(physical_tags, config: SolveConfig) -> dict[int, complex]:
    """Static Robin walls present in the mesh (reference: sweep.py:164-182)."""
    active: dict[int, complex] = {}
    if not config.impedance_sources:
        return active
    mesh_tags = {int(t) for t in np.unique(physical_tags)}
    for tag, beta in config.impedance_sources.items():
        if int(tag) not in mesh_tags:
            logger.warning("impedance_sources references tag %d but mesh has no elements with that tag; skipping", int(tag))
            continue
        active[int(tag)] = complex(beta)
    return active


def _impedance_sources_for_frequencies(physical_tags, frequencies, config: SolveConfig):
    """Per-frequency {tag: beta} dicts with the callback, else one static
    dict; the callback overrides and extends the static walls, passivity
    Re(beta) >= 0 is enforced (reference: sweep.py:185-235)."""
    static = _active_impedance_sources(physical_tags, config)
    if config.impedance_source_callback is None:
        return static
    mesh_tags = {int(t) for t in np.unique(physical_tags)}
    per_case = []
    for freq in frequencies:
        f = float(freq)
        merged = dict(static)
        for tag, beta in config.impedance_source_callback(f).items():
            t = int(tag)
            if t not in mesh_tags:
                logger.warning("impedance_source_callback returned tag %d not in mesh; skipping", t)
                continue
            b = complex(beta)
            if not (math.isfinite(b.real) and math.isfinite(b.imag)):
                raise ValueError(f"impedance_source_callback({f:.3f}) returned non-finite beta for tag {t}")
            if b.real < 0.0:
                raise ValueError(
                    f"impedance_source_callback({f:.3f}) returned Re(beta)={b.real:.4g} < 0 for tag {t}; "
                    "admittance must be passive (Re(beta) >= 0)"
                )
            merged[t] = b
        per_case.append(merged)
    return per_case


def _build_neumann_rows(dp0_space, physical_tags, frequencies, config, source_face_scale=None,
                        impedance=None):
    """q rows [F][T]; Robin tags carry no prescribed velocity (reference:
    sweep.py:238-285)."""
    rows = []
    for i, f in enumerate(frequencies):
        payload = impedance[i] if isinstance(impedance, list) else (impedance or {})
        rows.append(
            _build_driver_neumann_coeffs(
                dp0_space, physical_tags, 2.0 * np.pi * float(f), config, np.complex64,
                impedance_tags={int(t) for t in payload}, source_face_scale=source_face_scale,
            )
        )
    return np.stack(rows, axis=0)


def _field_points_3xn(obs_points: NDArray[np.float64]) -> NDArray[np.float32]:
    P, N, _ = obs_points.shape
    return np.ascontiguousarray(obs_points.reshape(P * N, 3).T, dtype=np.float32)


def _mesh_max_edge_m(mesh: LoadedMesh) -> float:
    V = np.asarray(mesh.grid.vertices, np.float64).T
    E = np.asarray(mesh.grid.elements, np.int64).T
    if E.size == 0:
        return 0.0
    p0, p1, p2 = V[E[:, 0]], V[E[:, 1]], V[E[:, 2]]
    return float(max(np.linalg.norm(p1 - p0, axis=1).max(), np.linalg.norm(p2 - p1, axis=1).max(),
                     np.linalg.norm(p0 - p2, axis=1).max()))


def _directivity_from_pressure(pressure: NDArray[np.complex128], on_axis_idx: int) -> NDArray[np.float64]:
    """20 log10(|p| / p_ref) with a -120 dB floor, minus the on-axis column
    (reference: sweep.py:523-532)."""
    floor = REFERENCE_PRESSURE * 10.0 ** (-120.0 / 20.0)
    spl = 20.0 * np.log10(np.maximum(np.abs(pressure), floor) / REFERENCE_PRESSURE)
    return spl - spl[:, on_axis_idx][:, None]


def _system_field(system, n_planes: int, n_angles: int, n_sphere: int):
    """One source's field row split into the arc pressure (n_planes,
    n_angles) and the observation-sphere values (or None)."""
    flat = np.asarray(system.field, np.complex128).reshape(-1)
    arc = flat[: n_planes * n_angles].reshape(n_planes, n_angles)
    sphere = flat[n_planes * n_angles : n_planes * n_angles + n_sphere] if n_sphere else None
    return arc, sphere


def new_result_rows(source_tags) -> dict:
    """Per-source accumulators of a sweep (filled by _append_system_result)."""
    return dict(pressure=[], spl=[], imp=[], surf=[], diag=[], log=[],
                pavg={int(t): [] for t in source_tags})


def _append_system_result(rows: dict, *, frequency_hz: float, system, case, diagnostics: dict,
                          source_tags, n_planes: int, n_angles: int, n_sphere: int,
                          on_axis_idx: int, keep_surface_pressure: bool):
    """Appends one source's result at one frequency to ``rows`` and returns
    (solver-log entry, arc pressure, directivity) (reference
    _append_system_result, sweep.py:535-609).  ``case`` is the frequency's
    result (operator-level times, lapack_info), ``system`` the source's own
    (field, impedance, tag averages, surface pressure; the case itself for
    the first source)."""
    pressure, sphere = _system_field(system, n_planes, n_angles, n_sphere)
    directivity = _directivity_from_pressure(pressure, on_axis_idx)
    entry = {
        "frequency_hz": float(frequency_hz),
        "iterations": None,
        "timing_s": case.assembly_s + case.dense_solve_s + case.field_s,
        "backend": "hip_gfx950_resident_assembly_lu_field_batch",
        "assembly_s": case.assembly_s,
        "dense_solve_s": case.dense_solve_s,
        "field_s": system.field_s,
        "lapack_info": case.lapack_info,
        "impedance": system.impedance,
        "native_diagnostics": diagnostics,
        "observation_sphere_pressure_complex": sphere,
    }
    for t in source_tags:
        rows["pavg"][int(t)].append((system.surface_pressure_avg or {}).get(int(t), 0j))
    rows["pressure"].append(pressure)
    rows["spl"].append(directivity)
    rows["imp"].append(system.impedance)
    if keep_surface_pressure:
        rows["surf"].append(np.asarray(system.pressure, np.complex128))
    rows["diag"].append(diagnostics)
    rows["log"].append(entry)
    return entry, pressure, directivity


def _apply_dense_solve_policy(diagnostics: dict, *, threshold: float) -> None:
    """rcond below the threshold marks the frequency suspect (reference:
    sweep.py:383-412)."""
    diagnostics["dense_solve_rcond_warning_threshold"] = float(threshold)
    diagnostics["dense_solve_suspect"] = False
    if diagnostics.get("dense_solve_dtype") == "float64" and diagnostics.get("dense_solve_refine_stalled"):
        # the FP64-accurate solve stopped above its residual target: the
        # answer is not the FP64 solution the setting asks for
        diagnostics["dense_solve_suspect"] = True
        diagnostics["dense_solve_recommendation"] = (
            "The float64 dense solve did not reach its residual target (refinement stalled); "
            "treat this frequency as suspect and compare against nearby points."
        )
    if threshold <= 0.0:
        return
    raw = diagnostics.get("dense_solve_rcond")
    if raw is None:
        diagnostics["dense_solve_policy_available"] = False
        return
    diagnostics["dense_solve_policy_available"] = True
    rcond = float(raw)
    if math.isfinite(rcond) and rcond < threshold:
        diagnostics["dense_solve_suspect"] = True
        diagnostics["dense_solve_recommendation"] = (
            "Treat this frequency as suspect; nudge or densify the frequency grid around the "
            "resonance and compare against nearby points. This flag detects dense-solve "
            "conditioning only."
        )


def _apply_mesh_resolution_policy(diagnostics: dict, *, frequency_hz: float, mesh_max_edge_m: float,
                                  elements_per_wavelength_min: float) -> None:
    """Elements-per-wavelength check (reference: sweep.py:415-440)."""
    diagnostics["mesh_max_edge_m"] = float(mesh_max_edge_m)
    diagnostics["mesh_elements_per_wavelength_min"] = float(elements_per_wavelength_min)
    if mesh_max_edge_m <= 0.0:
        diagnostics["mesh_elements_per_wavelength"] = math.inf
        diagnostics["mesh_max_valid_frequency_hz"] = math.inf
        diagnostics["mesh_resolution_suspect"] = False
        return
    epw = (SPEED_OF_SOUND / float(frequency_hz)) / float(mesh_max_edge_m)
    diagnostics["mesh_elements_per_wavelength"] = float(epw)
    diagnostics["mesh_max_valid_frequency_hz"] = float(
        SPEED_OF_SOUND / (float(elements_per_wavelength_min) * float(mesh_max_edge_m))
    )
    diagnostics["mesh_resolution_suspect"] = bool(epw < elements_per_wavelength_min)


def _prepare(mesh: LoadedMesh, frame: ObservationFrame, config: SolveConfig):
    mesh_tags = {int(t) for t in np.unique(mesh.physical_tags)}
    missing = sorted(set(config.velocity_sources) - mesh_tags)
    if missing:
        raise ValueError(
            f"velocity_sources tags {missing} are not present in the mesh; "
            f"available physical tags: {sorted(mesh_tags)}"
        )
    obs_points, angles = build_observation_points(frame, config.observation)
    p1, dp0 = make_pure_function_spaces(mesh.grid)
    buffers = build_geometry_buffers(mesh.grid, mesh.physical_tags, p1, dp0)
    face_scale = _build_source_face_scale(
        mesh.grid, mesh.physical_tags, config, frame.axis, frame.source_center
    )
    return obs_points, angles, p1, dp0, buffers, face_scale


def _native_sweep(
    mesh: LoadedMesh, frequencies, frame: ObservationFrame, config: SolveConfig,
    sources: list[dict] | None,
) -> list[SolveResult]:
    """Standard-Neumann sweep on the GPU (reference: sweep.py:612-926, and
    the multi-RHS sweep sweep.py:957-1301 when ``sources`` is given: one
    assembled and factored operator per frequency, one right-hand side per
    source dict).  Returns one SolveResult per source."""
    from dataclasses import replace

    from .native import HipBemSession

    should_route_native_metal(config)
    freqs = np.asarray(frequencies, dtype=np.float64).reshape(-1)
    if freqs.size == 0:
        raise ValueError("frequencies must contain at least one value")
    t_total = time.time()
    obs_points, angles, p1, dp0, buffers, face_scale = _prepare(mesh, frame, config)
    src_dicts = [dict(config.velocity_sources)] if sources is None else [dict(s) for s in sources]
    S = len(src_dicts)
    multi = sources is not None
    # reductions: each source's own tags, or the union of every source's tags
    # in every result (multi-source, reference sweep.py:972-975)
    union = sorted({int(t) for d in src_dicts for t in d})
    source_tags = union if multi else list(config.velocity_sources.keys())
    imp_tags = [min((int(t) for t in d), default=2) for d in src_dicts]
    on_axis = int(np.argmin(np.abs(angles)))
    P, N, _ = obs_points.shape
    pts = _field_points_3xn(obs_points)
    n_sphere = 0
    if config.observation.sphere_points is not None:
        sph = np.ascontiguousarray(np.asarray(config.observation.sphere_points, np.float64).T, np.float32)
        pts = np.concatenate([pts, sph], axis=1)
        n_sphere = int(sph.shape[1])
    max_edge = _mesh_max_edge_m(mesh)
    k_real, k_imag = _k_values_for_native(freqs, config)
    impedance = _impedance_sources_for_frequencies(mesh.physical_tags, freqs, config)
    q_rows = [
        _build_neumann_rows(dp0, mesh.physical_tags, freqs, replace(config, velocity_sources=d),
                            face_scale, impedance)
        for d in src_dicts
    ]
    per_src = [new_result_rows(source_tags) for _ in range(S)]
    done: list[float] = []
    stream = config.on_frequency_result is not None
    chunk = (config.hip_batch_frequencies or 8) if stream else freqs.size
    stop = False
    with HipBemSession.create_session(
        buffers, symmetry_plane=config.native_symmetry_plane,
        check_open_edges=config.native_check_open_edges,
    ) as session:
        session.set_assembly_mode(config.metal_native_assembly_mode)
        for c0 in range(0, freqs.size, chunk):
            sl = slice(c0, min(freqs.size, c0 + chunk))
            imp_sl = impedance[sl] if isinstance(impedance, list) else impedance
            results = session.assemble_solve_evaluate_standard_neumann_batch(
                freqs[sl], k_real[sl], q_rows[0][sl], pts,
                k_imag_f32=k_imag[sl], field_k_real=k_real[sl], source_tags=source_tags,
                impedance_source_tag=imp_tags[0],
                write_surface_pressure=config.return_surface_pressure,
                solve_mode=config.effective_hip_solve_mode,
                refine_max_iterations=config.hip_refine_max_iterations,
                refine_tolerance=config.hip_refine_tolerance,
                batch_frequencies=config.hip_batch_frequencies,
                estimate_rcond=config.hip_estimate_rcond,
                impedance_sources=imp_sl if imp_sl else None,
                extra_neumann_dp0=np.stack([q[sl] for q in q_rows[1:]], axis=0) if S > 1 else None,
                extra_impedance_source_tags=imp_tags[1:] if S > 1 else None,
                chief_points=(None if config.chief_points is None else
                              np.asarray(config.chief_points, np.float64).reshape(-1, 3).T),
                chief_weight=config.chief_weight,
                precise_geometry=config.hip_precise_geometry,
            )
            for j, r in enumerate(results):
                i = c0 + j
                f = float(freqs[i])
                diag = dict(r.diagnostics)
                _apply_dense_solve_policy(diag, threshold=config.dense_solve_rcond_warning_threshold)
                _apply_mesh_resolution_policy(
                    diag, frequency_hz=f, mesh_max_edge_m=max_edge,
                    elements_per_wavelength_min=config.mesh_elements_per_wavelength_min,
                )
                entries = []
                for sidx in range(S):
                    rr = r if sidx == 0 else r.extra_sources[sidx - 1]
                    entries.append(_append_system_result(
                        per_src[sidx], frequency_hz=f, system=rr, case=r, diagnostics=diag,
                        source_tags=source_tags, n_planes=P, n_angles=N, n_sphere=n_sphere,
                        on_axis_idx=on_axis, keep_surface_pressure=config.return_surface_pressure,
                    ))
                done.append(f)
                if config.progress_callback is not None:
                    config.progress_callback(i, freqs.size, f)
                if config.on_frequency_result is not None:
                    entry, pressure, directivity = entries[0]
                    cb = {
                        **entry,
                        "observation_pressure_complex": pressure,
                        "observation_directivity_db": directivity,
                        "observation_angles_deg": angles,
                        "observation_planes": config.observation.planes,
                    }
                    if multi:
                        cb["source_results"] = [e for e, _, _ in entries]
                    if config.on_frequency_result(i, f, cb) is False:
                        logger.info("Early stop requested after %.1f Hz", f)
                        stop = True
                        break
            if stop:
                break
    out = []
    for sidx in range(S):
        acc = per_src[sidx]
        log = acc["log"]
        timings = {
            "solve_s": sum(e["timing_s"] for e in log),
            "assembly_s": sum(e["assembly_s"] for e in log),
            "dense_solve_s": sum(e["dense_solve_s"] for e in log),
            "directivity_s": sum(e["field_s"] for e in log),
            "total_s": time.time() - t_total,
        }
        sp_avg = {t: np.array(v, dtype=np.complex128) for t, v in acc["pavg"].items()}
        cfg = replace(config, velocity_sources=src_dicts[sidx]) if multi else config
        out.append(
            SolveResult(
                frequencies_hz=np.array(done, dtype=np.float64),
                pressure_complex=np.stack(acc["pressure"], axis=0),
                directivity_db=np.stack(acc["spl"], axis=0),
                impedance=np.array(acc["imp"], dtype=np.complex128),
                observation_angles_deg=angles,
                observation_points=obs_points,
                observation_planes=config.observation.planes,
                config=cfg,
                mesh_info=mesh.info,
                timings=timings,
                solver_log=log,
                surface_pressure_avg=sp_avg if sp_avg else None,
                surface_pressure_complex=(
                    np.stack(acc["surf"], axis=0) if config.return_surface_pressure else None
                ),
                native_diagnostics=acc["diag"],
            )
        )
    return out


def run_sweep_native_metal(
    mesh: LoadedMesh, frequencies: NDArray[np.float64], frame: ObservationFrame, config: SolveConfig
) -> SolveResult:
    """Standard-Neumann sweep on the GPU (reference: sweep.py:612-926)."""
    return _native_sweep(mesh, frequencies, frame, config, None)[0]


run_sweep = run_sweep_native_metal


def run_sweep_native_metal_multi_source(
    mesh: LoadedMesh, frequencies, frame: ObservationFrame, config: SolveConfig, sources: list[dict]
) -> list[SolveResult]:
    """One SolveResult per velocity-source dict, all sources sharing each
    frequency's assembled and factored operator (reference: sweep.py:957-1301).
    ``surface_pressure_avg`` covers the union of every source's tags."""
    if not sources:
        raise ValueError("sources must contain at least one velocity dict")
    if config.velocity_source_callback is not None:
        raise ValueError("multi-source solves do not support velocity_source_callback")
    mesh_tags = {int(t) for t in np.unique(mesh.physical_tags)}
    for index, source in enumerate(sources):
        missing = sorted(set(int(t) for t in source) - mesh_tags)
        if missing:
            raise ValueError(
                f"sources[{index}] tags {missing} are not present in the mesh; "
                f"available physical tags: {sorted(mesh_tags)}"
            )
    if len(sources) == 1:
        from dataclasses import replace

        return [run_sweep_native_metal(mesh, frequencies, frame, replace(config, velocity_sources=dict(sources[0])))]
    return _native_sweep(mesh, frequencies, frame, config, sources)

