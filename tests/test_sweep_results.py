"""Result assembly of the sweep, on the CPU with a fake library session
(the reference's tests/test_sweep_results.py:87-155 and the fake-helper
tier of tests/test_metal_native.py:1234-1326): per-case rows into
SolveResult, directivity, impedance, tag averages, solver log, the
dense-solve and mesh-resolution policies, streaming early stop, the
multi-source split, and the arguments the sweep hands the session."""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest

from hornlab_metal_bem import ObservationConfig, native_config, solve, solve_multi_source
from hornlab_metal_bem import sweep
from hornlab_metal_bem.backends import AssemblyBackendUnavailable
from hornlab_metal_bem.procedural import sphere_mesh

P_REF = 20e-6


def test_system_field_splits_arc_and_sphere():
    flat = np.arange(2 * 3 + 4, dtype=np.complex128)
    arc, sphere = sweep._system_field(SimpleNamespace(field=flat), 2, 3, 4)
    np.testing.assert_array_equal(arc, flat[:6].reshape(2, 3))
    np.testing.assert_array_equal(sphere, flat[6:])
    arc, none = sweep._system_field(SimpleNamespace(field=flat[:6]), 2, 3, 0)
    assert none is None and arc.shape == (2, 3)


def test_append_system_result_keeps_field_surface_pressure_and_diagnostics():
    system = SimpleNamespace(
        field=np.array([1.0 + 0.5j, 2.0 - 0.5j], np.complex64),
        pressure=np.array([3.0 + 1.0j, 4.0 + 1.5j, 5.0 + 2.0j], np.complex64),
        impedance=1.0 + 2.0j, surface_pressure_avg={2: 3.0 + 4.0j},
        assembly_s=0.1, dense_solve_s=0.2, field_s=0.3, lapack_info=0,
    )
    rows = sweep.new_result_rows([2])
    diag = {"assembly_implementation": "test_assembly"}
    entry, arc, directivity = sweep._append_system_result(
        rows, frequency_hz=1000.0, system=system, case=system, diagnostics=diag, source_tags=[2],
        n_planes=1, n_angles=2, n_sphere=0, on_axis_idx=0, keep_surface_pressure=True)
    np.testing.assert_allclose(rows["pressure"][0], [[1.0 + 0.5j, 2.0 - 0.5j]])
    np.testing.assert_allclose(rows["surf"][0], [3.0 + 1.0j, 4.0 + 1.5j, 5.0 + 2.0j])
    assert rows["pavg"][2] == [3.0 + 4.0j]
    assert rows["imp"] == [1.0 + 2.0j]
    assert entry["lapack_info"] == 0
    assert entry["timing_s"] == pytest.approx(0.6)
    assert entry["native_diagnostics"]["assembly_implementation"] == "test_assembly"
    assert rows["diag"] == [entry["native_diagnostics"]]
    # directivity: 20 log10(|p| / p_ref) minus the on-axis column
    expect = 20 * np.log10(np.abs([2.0 - 0.5j]) / np.abs(1.0 + 0.5j))
    np.testing.assert_allclose(directivity[0], [0.0, expect[0]], atol=1e-5)


def test_dense_solve_policy_rcond_and_float64_stall():
    d = {"dense_solve_rcond": 1e-8}
    sweep._apply_dense_solve_policy(d, threshold=1e-6)
    assert d["dense_solve_suspect"] is True and d["dense_solve_policy_available"] is True
    d = {"dense_solve_rcond": 1e-3, "dense_solve_dtype": "float64", "dense_solve_refine_stalled": True}
    sweep._apply_dense_solve_policy(d, threshold=1e-6)
    assert d["dense_solve_suspect"] is True and "stalled" in d["dense_solve_recommendation"]
    d = {"dense_solve_rcond": 1e-3, "dense_solve_dtype": "mixed_fp16", "dense_solve_refine_stalled": True}
    sweep._apply_dense_solve_policy(d, threshold=1e-6)
    assert d["dense_solve_suspect"] is False
    d = {}
    sweep._apply_dense_solve_policy(d, threshold=1e-6)
    assert d["dense_solve_policy_available"] is False


class _FakeSession:
    """Stands in for HipBemSession: records the arguments of each batch call
    and returns synthetic per-case results of the library's shape."""

    calls: list = []
    modes: list = []

    def __init__(self, buffers):
        self.buffers = buffers

    @classmethod
    def create_session(cls, buffers, *, symmetry_plane=None, check_open_edges=True):
        return cls(buffers)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return None

    def set_assembly_mode(self, mode):
        _FakeSession.modes.append(mode)

    def assemble_solve_evaluate_standard_neumann_batch(self, freqs, k, q, pts, **kw):
        from hornlab_metal_bem.native import DenseSolveFieldResult, ExtraSourceSolveResult

        _FakeSession.calls.append(dict(kw, freqs=np.asarray(freqs), q=np.asarray(q)))
        n_obs = pts.shape[1]
        n = self.buffers.p1_dof_count
        out = []
        for i, f in enumerate(freqs):
            # field magnitude grows with the observation index: known directivity
            field = (np.arange(n_obs) + 1.0) * (1.0 + 1j) * float(f) * 1e-3
            extras = []
            for e in range(0 if kw.get("extra_neumann_dp0") is None else len(kw["extra_neumann_dp0"])):
                extras.append(ExtraSourceSolveResult(
                    source_index=e, pressure=np.full(n, 2.0 + e, np.complex64),
                    field=(field * (2.0 + e)).astype(np.complex64), field_s=0.01,
                    impedance=complex(2.0 + e, 0.0), surface_pressure_avg={2: complex(2.0 + e)}))
            out.append(DenseSolveFieldResult(
                frequency_hz=float(f), pressure=np.full(n, 1.0 + 0.0j, np.complex64),
                field=field.astype(np.complex64), assembly_s=0.1, dense_solve_s=0.2, field_s=0.05,
                lapack_info=0, impedance=complex(f, -f), surface_pressure_avg={2: complex(1.0, f)},
                diagnostics={"dense_solve_rcond": 1e-8 if i == 0 else 0.1,
                             "dense_solve_dtype": "float32"},
                extra_sources=tuple(extras)))
        return out


@pytest.fixture
def fake_session(monkeypatch):
    import hornlab_metal_bem.native as native

    _FakeSession.calls = []
    _FakeSession.modes = []
    monkeypatch.setattr(native, "HipBemSession", _FakeSession)
    return _FakeSession


def _config(**kw):
    return native_config(
        freq_min_hz=500.0, freq_max_hz=1500.0, freq_count=3, freq_spacing="linear",
        velocity_mode="velocity", air_density=1.21,
        observation=ObservationConfig(planes=["horizontal", "vertical"], distance_m=2.0,
                                      angle_count=5),
        **kw,
    )


def test_solve_assembles_solve_result_from_session_cases(fake_session):
    mesh = sphere_mesh(0.1, 1)
    res = solve(mesh, _config(return_surface_pressure=True))
    np.testing.assert_allclose(res.frequencies_hz, [500.0, 1000.0, 1500.0])
    assert res.pressure_complex.shape == (3, 2, 5)
    for i, f in enumerate(res.frequencies_hz):
        field = (np.arange(10) + 1.0) * (1.0 + 1j) * f * 1e-3
        np.testing.assert_allclose(res.pressure_complex[i], field.reshape(2, 5), rtol=1e-6)
        spl = 20 * np.log10(np.abs(field.reshape(2, 5)) / P_REF)
        on_axis = int(np.argmin(np.abs(res.observation_angles_deg)))
        np.testing.assert_allclose(res.directivity_db[i], spl - spl[:, on_axis][:, None], atol=1e-4)
    np.testing.assert_allclose(res.impedance, [500 - 500j, 1000 - 1000j, 1500 - 1500j])
    np.testing.assert_allclose(res.surface_pressure_avg[2], [1 + 500j, 1 + 1000j, 1 + 1500j])
    assert res.surface_pressure_complex.shape == (3, mesh.info.n_vertices)
    assert len(res.solver_log) == 3 and len(res.native_diagnostics) == 3
    # rcond 1e-8 at the first frequency is below the 1e-6 policy threshold
    assert [d["dense_solve_suspect"] for d in res.native_diagnostics] == [True, False, False]
    assert all("mesh_elements_per_wavelength" in d for d in res.native_diagnostics)
    assert res.timings["dense_solve_s"] == pytest.approx(0.6)
    # what the sweep asked of the session
    call = fake_session.calls[0]
    assert call["solve_mode"] == "fp32" and call["estimate_rcond"] is True
    assert call["source_tags"] == [2] and call["impedance_source_tag"] == 2
    # q = i rho omega v_n (velocity mode, unit normal velocity on tag 2)
    np.testing.assert_allclose(call["q"][1], 1j * 1.21 * 2 * np.pi * 1000.0, rtol=1e-6)
    assert fake_session.modes == ["corrected"]


def test_solve_float64_and_modes_reach_the_session(fake_session):
    mesh = sphere_mesh(0.1, 1)
    solve(mesh, _config(dense_solve_dtype="float64", metal_native_assembly_mode="optimized"))
    assert fake_session.calls[-1]["solve_mode"] == "fp64"
    assert fake_session.modes[-1] == "optimized"
    with pytest.raises(AssemblyBackendUnavailable, match="parity"):
        solve(mesh, _config(metal_native_assembly_mode="parity"))


def test_streaming_callback_stops_early_and_reports_each_case(fake_session):
    mesh = sphere_mesh(0.1, 1)
    seen = []

    def on_result(i, f, payload):
        seen.append((i, f, payload["observation_pressure_complex"].shape))
        return False

    res = solve(mesh, _config(on_frequency_result=on_result, hip_batch_frequencies=1))
    assert seen == [(0, 500.0, (2, 5))]
    np.testing.assert_allclose(res.frequencies_hz, [500.0])
    assert len(fake_session.calls) == 1


def test_multi_source_splits_results_per_source(fake_session):
    mesh = sphere_mesh(0.1, 1)
    tags = np.asarray(mesh.physical_tags).copy()
    tags[: tags.size // 2] = 3
    from hornlab_metal_bem import loaded_mesh_from_arrays

    V = np.asarray(mesh.grid.vertices).T
    F = np.asarray(mesh.grid.elements).T
    mesh2 = loaded_mesh_from_arrays(V, F, tags)
    out = solve_multi_source(mesh2, [{2: 1.0}, {3: 1.0}], _config())
    assert len(out) == 2
    np.testing.assert_allclose(out[1].pressure_complex, 2.0 * out[0].pressure_complex, rtol=1e-6)
    np.testing.assert_allclose(out[1].impedance, [2.0, 2.0, 2.0])
    assert out[0].config.velocity_sources == {2: 1.0}
    assert out[1].config.velocity_sources == {3: 1.0}
    call = fake_session.calls[0]
    assert call["extra_neumann_dp0"].shape[0] == 1
    assert sorted(call["source_tags"]) == [2, 3]
