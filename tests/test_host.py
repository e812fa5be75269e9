"""Host-side logic (config, geometry buffers, frames, Neumann data, result
assembly) and the C ABI surface.  CPU only; restates the reference's
pure-Python test tiers (test_config.py, test_metal_geometry.py,
test_bie_unit.py, test_observation.py, test_sweep_results.py)."""
from __future__ import annotations

import ctypes
import math
import re
from pathlib import Path
from types import SimpleNamespace

import numpy as np
import pytest

from hornlab_metal_bem import ObservationConfig, SolveConfig, native_config
from hornlab_metal_bem.backends import AssemblyBackendUnavailable
from hornlab_metal_bem.bie import _build_driver_neumann_coeffs, _build_source_face_scale
from hornlab_metal_bem.config import AxialProfile, TaperProfile, VelocityMode
from hornlab_metal_bem.geometry import GeometryError, build_geometry_buffers, validate_native_symmetry_plane
from hornlab_metal_bem.mesh import (
    loaded_mesh_from_arrays,
    make_pure_function_spaces,
    open_boundary_edges,
    signed_volume_indicator,
)
from hornlab_metal_bem.observation import build_observation_points, infer_frame
from hornlab_metal_bem.procedural import HornSpec, horn_body, icosphere, mirror_quarter_to_full
from hornlab_metal_bem.sweep import (
    _apply_dense_solve_policy,
    _apply_mesh_resolution_policy,
    _build_frequency_grid,
    _directivity_from_pressure,
    _k_values_for_native,
)
from meshes import tiny_two_triangles, tiny_yz_xz_quarter

ROOT = Path(__file__).resolve().parents[1]


# -- config -----------------------------------------------------------------

def test_native_config_defaults():
    c = native_config()
    assert c.metal_native_assembly_mode == "corrected"
    assert c.velocity_mode == VelocityMode.ACCELERATION
    assert c.air_density == pytest.approx(1.2041)
    assert c.velocity_sources == {2: 1.0}
    assert c.effective_hip_solve_mode == "fp32"
    assert native_config(dense_solve_dtype="float64").effective_hip_solve_mode == "fp64"


@pytest.mark.parametrize(
    "kwargs,msg",
    [
        ({"freq_count": 0}, "freq_count"),
        ({"freq_min_hz": -1.0}, "freq_min_hz"),
        ({"freq_min_hz": 100.0, "freq_max_hz": 50.0}, "freq_max_hz"),
        ({"freq_spacing": "cubic"}, "freq_spacing"),
        ({"native_symmetry_plane": "zz"}, "native_symmetry_plane"),
        ({"dense_solve_dtype": "float16"}, "dense_solve_dtype"),
        ({"velocity_mode": "x"}, "velocity_mode"),
        ({"hip_solve_mode": "fp8"}, "hip_solve_mode"),
        ({"air_density": 0.0}, "air_density"),
        ({"source_velocity_profiles": {2: TaperProfile(start=1.5)}}, "TaperProfile.start"),
    ],
)
def test_config_validation(kwargs, msg):
    with pytest.raises(ValueError, match=re.escape(msg)):
        SolveConfig(**kwargs)


def test_frequency_grid_and_wavenumbers():
    c = native_config(freq_min_hz=100, freq_max_hz=2000, freq_count=10, freq_spacing="linear")
    f = _build_frequency_grid(c)
    assert np.allclose(f, np.linspace(100, 2000, 10))
    k, kim = _k_values_for_native(f, c)
    assert k.dtype == np.float32 and np.allclose(k, 2 * np.pi * f / 343.0, rtol=1e-7)
    assert np.all(kim == 0)
    ck = native_config(formulation="complex_k", complex_k_shift=0.01)
    _, kim2 = _k_values_for_native(f, ck)
    assert np.allclose(kim2, 0.01 * k, rtol=1e-6)


# -- geometry buffers ---------------------------------------------------------

def test_geometry_buffers_layout_and_snapping():
    b = tiny_two_triangles()
    assert b.vertices_3xn_f32.shape == (3, 4) and b.vertices_3xn_f32.dtype == np.float32
    assert b.triangles_3xm_i32.shape == (3, 2) and b.triangles_3xm_i32.dtype == np.int32
    assert b.p1_local2global_i32.shape == (2, 3)
    assert np.allclose(b.triangle_areas_f32, [0.5, 0.5])
    assert np.allclose(np.linalg.norm(b.triangle_normals_3xm_f32, axis=0), 1.0)
    grid = SimpleNamespace(vertices=np.array([[5e-7, 1.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0]]),
                           elements=np.array([[0], [1], [2]]))
    p1 = SimpleNamespace(local2global=np.array([[0, 1, 2]]), global_dof_count=3)
    snapped = build_geometry_buffers(grid, np.array([2]), p1)
    assert snapped.vertices_3xn_f32[0, 0] == 0.0


def test_geometry_rejects_bad_indices():
    grid = SimpleNamespace(vertices=np.eye(3), elements=np.array([[1], [2], [3]]))
    p1 = SimpleNamespace(local2global=np.array([[0, 1, 2]]), global_dof_count=3)
    with pytest.raises(GeometryError, match="zero-based"):
        build_geometry_buffers(grid, np.array([2]), p1)
    grid2 = SimpleNamespace(vertices=np.eye(3), elements=np.array([[0], [0], [2]]))
    with pytest.raises(GeometryError, match="repeats a vertex"):
        build_geometry_buffers(grid2, np.array([2]), p1)


def test_symmetry_plane_validation():
    assert validate_native_symmetry_plane(tiny_yz_xz_quarter(), "yz+xz", check_open_edges=False) == "yz+xz"
    with pytest.raises(GeometryError, match="positive"):
        b = tiny_two_triangles()
        shifted = b.__class__(**{**b.__dict__, "vertices_3xn_f32": b.vertices_3xn_f32 - 0.5})
        validate_native_symmetry_plane(shifted, "yz")
    Vq, Fq, Tq = horn_body(HornSpec(), edge_length=0.04, quarter=True)
    from meshes import mesh_buffers

    assert validate_native_symmetry_plane(mesh_buffers(Vq, Fq, Tq), "yz+xz") == "yz+xz"


# -- procedural meshes ----------------------------------------------------------

def test_icosphere_counts_and_winding():
    V, F = icosphere(0.1, 3)
    assert V.shape == (642, 3) and F.shape == (1280, 3)
    assert signed_volume_indicator(V, F) > 0
    assert open_boundary_edges(F).shape[0] == 0
    assert np.allclose(np.linalg.norm(V, axis=1), 0.1)


def test_horn_body_closed_outward_and_quarter_rim_on_planes():
    V, F, T = horn_body(HornSpec(), edge_length=0.02, quarter=False)
    assert open_boundary_edges(F).shape[0] == 0
    assert signed_volume_indicator(V, F) > 0
    assert set(np.unique(T)) == {1, 2}
    Vq, Fq, Tq = horn_body(HornSpec(), edge_length=0.02, quarter=True)
    e = open_boundary_edges(Fq)
    on = np.all((np.abs(Vq[e][:, :, 0]) < 1e-12) | (np.abs(Vq[e][:, :, 1]) < 1e-12), axis=1)
    assert on.all()
    Vf, Ff, Tf = mirror_quarter_to_full(Vq, Fq, Tq)
    assert open_boundary_edges(Ff).shape[0] == 0 and signed_volume_indicator(Vf, Ff) > 0


# -- frame and observation -------------------------------------------------------

def test_horn_frame_axis_and_mouth_origin():
    V, F, T = horn_body(HornSpec(), edge_length=0.02, quarter=True)
    mesh = loaded_mesh_from_arrays(V, F, T, native_symmetry_plane="yz+xz")
    fr = infer_frame(mesh.grid, mesh.physical_tags, source_tag=2, symmetry_plane="yz+xz")
    assert np.allclose(fr.axis, [0, 0, 1], atol=1e-9)
    assert np.allclose(fr.origin[:2], 0.0) and fr.origin[2] == pytest.approx(0.150, abs=2e-3)
    pts, ang = build_observation_points(fr, ObservationConfig(planes=["horizontal", "vertical", "diagonal"]))
    assert pts.shape == (3, 37, 3) and ang[0] == 0.0 and ang[-1] == 180.0
    assert np.allclose(np.linalg.norm(pts - fr.origin, axis=2), 2.0)
    assert np.allclose(pts[:, 0], fr.origin + 2.0 * fr.axis)


def test_custom_points_validation():
    fr = SimpleNamespace(axis=np.array([0, 0, 1.0]), origin=np.zeros(3), u=np.array([1.0, 0, 0]), v=np.array([0, 1.0, 0]))
    cfg = ObservationConfig(planes=["horizontal"], custom_points={"horizontal": np.zeros((5, 3))})
    pts, ang = build_observation_points(fr, cfg)
    assert pts.shape == (1, 5, 3) and ang.shape == (5,)
    with pytest.raises(ValueError, match="missing plane"):
        build_observation_points(fr, ObservationConfig(planes=["vertical"], custom_points={"horizontal": np.zeros((5, 3))}))


# -- Neumann data ------------------------------------------------------------------

def test_neumann_known_answers():
    """q = i rho omega v (velocity), q = rho w (acceleration); default rho
    1.2041 (test_bie_unit.py:25-129)."""
    dp0 = SimpleNamespace(global_dof_count=2)
    tags = np.array([1, 2], np.int32)
    w = 2 * np.pi * 500.0
    qv = _build_driver_neumann_coeffs(dp0, tags, w, native_config(velocity_mode="velocity"), np.complex128)
    assert qv[0] == 0 and qv[1] == pytest.approx(1j * 1.2041 * w)
    qa = _build_driver_neumann_coeffs(dp0, tags, w, native_config(), np.complex128)
    assert qa[1] == pytest.approx(1.2041 + 0j)
    qs = _build_driver_neumann_coeffs(dp0, tags, w, native_config(velocity_sources={2: 0.5}, air_density=1.21,
                                                                   velocity_mode="velocity"), np.complex128)
    assert qs[1] == pytest.approx(1j * 1.21 * w * 0.5)


def test_axial_profile_on_flat_disc_equals_normal():
    V, F, T = horn_body(HornSpec(), edge_length=0.02, quarter=False)
    mesh = loaded_mesh_from_arrays(V, F, T)
    cfg = native_config(source_motion="axial")
    scale = _build_source_face_scale(mesh.grid, mesh.physical_tags, cfg, np.array([0, 0, 1.0]), np.zeros(3))
    assert scale is not None
    assert np.allclose(scale[T == 2], 1.0, atol=1e-12) and np.all(scale[T == 1] == 0)
    assert _build_source_face_scale(mesh.grid, mesh.physical_tags, native_config(), np.array([0, 0, 1.0]), np.zeros(3)) is None


# -- result policies -------------------------------------------------------------------

def test_directivity_and_policies():
    p = np.array([[2.0, 1.0, 0.0]], complex)
    d = _directivity_from_pressure(p, 0)
    assert d[0, 0] == 0 and d[0, 1] == pytest.approx(-20 * math.log10(2))
    assert np.isfinite(d[0, 2])
    diag = {"dense_solve_rcond": 1e-8}
    _apply_dense_solve_policy(diag, threshold=1e-6)
    assert diag["dense_solve_suspect"] is True
    diag2 = {}
    _apply_dense_solve_policy(diag2, threshold=1e-6)
    assert diag2["dense_solve_policy_available"] is False
    m = {}
    _apply_mesh_resolution_policy(m, frequency_hz=20000.0, mesh_max_edge_m=0.01, elements_per_wavelength_min=6.0)
    assert m["mesh_resolution_suspect"] is True
    assert m["mesh_max_valid_frequency_hz"] == pytest.approx(343.0 / 0.06)


# -- C ABI surface ------------------------------------------------------------------------

def _declared_symbols():
    text = (ROOT / "include" / "hbem.h").read_text()
    return sorted(set(re.findall(r"\b(hbem_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from hornlab_metal_bem.native import EXPORTED_SYMBOLS, load_library

    lib = load_library()
    declared = _declared_symbols()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(EXPORTED_SYMBOLS) == declared
    from hornlab_metal_bem.native import HBEM_ABI_VERSION

    header = (ROOT / "include" / "hbem.h").read_text()
    assert f"#define HBEM_ABI_VERSION {HBEM_ABI_VERSION}" in header
    assert lib.hbem_abi_version() == HBEM_ABI_VERSION


def test_no_gpu_fails_loudly():
    from hornlab_metal_bem.native import HipBemSession, discover_native_runtime

    if discover_native_runtime().available:
        pytest.skip("a GPU is visible")
    with pytest.raises(AssemblyBackendUnavailable):
        HipBemSession.create_session(tiny_two_triangles())


def test_near_quadrature_env_parsing():
    """HORNLAB_METAL_BEM_NATIVE_NEAR_QUADRATURE as the reference helper reads
    it (main.swift:5024-5055; test_metal_native.py:1952-1977 for junk)."""
    from hornlab_metal_bem.native import NEAR_QUADRATURE_ENV, near_quadrature_from_env as parse

    assert parse({}) is None
    assert parse({NEAR_QUADRATURE_ENV: ""}) is None
    assert parse({NEAR_QUADRATURE_ENV: "0"}) is None
    assert parse({NEAR_QUADRATURE_ENV: "1"}) == (1, 1.5)
    assert parse({NEAR_QUADRATURE_ENV: "2"}) == (2, 1.5)
    assert parse({NEAR_QUADRATURE_ENV: "2:0.75"}) == (2, 0.75)
    for junk in ("junk", "3", "1:0", "1:-2", "1:inf", "1:nan", "0:1.5", "1:", "1:2:3", " 1"):
        with pytest.raises(RuntimeError, match=NEAR_QUADRATURE_ENV):
            parse({NEAR_QUADRATURE_ENV: junk})
