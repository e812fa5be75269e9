"""ctypes boundary to libhbem (include/hbem.h).

``HipBemSession`` takes the place of the reference's
``MetalNativeStandardSession`` (hornlab_metal_bem/metal/native.py:233-1879):
same operation names and argument meaning, but an in-process C ABI instead of
a helper subprocess exchanging JSON manifests and binary files.  Results are
returned as arrays instead of file descriptors.

There is no CPU fallback: if the library or a gfx950 device is missing every
entry point raises ``AssemblyBackendUnavailable``.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Callable

import numpy as np
from numpy.typing import NDArray

from .backends import AssemblyBackendUnavailable
from .geometry import SYMMETRY_CODES, GeometryBuffers, validate_native_symmetry_plane

HBEM_ABI_VERSION = 6
_LIB_NAME = "libhbem.so"

SOLVE_MODES = {"fp32": 0, "fp32_ir": 1, "mixed_fp16": 2, "mixed_bf16": 3, "fp64": 4}

_SOLVE_METHODS = {
    "fp32": "fp32_lu",
    "fp32_ir": "fp32_lu_fp64_residual_refinement",
    "mixed_fp16": "fp16_trailing_update_lu_gmres_ir",
    "mixed_bf16": "bf16_trailing_update_lu_gmres_ir",
    "fp64": "fp32_lu_gmres_ir_to_fp64_accuracy",
}

# metal_native_assembly_mode -> quadrature rule of the HIP kernels
# (include/hbem.h HBEM_ASSEMBLY_*).  "reference" is the reference's CPU
# quadrature: the 6x6 rule on every pair, plus Sauter-Schwab corrections when
# the case has complex k or Robin walls (main.swift:8139-8186); it runs here
# on the same GPU kernels.  "parity" compares the CPU reference with the GPU
# kernels inside the helper and has no counterpart without that CPU path.
ASSEMBLY_CORRECTED, ASSEMBLY_REGULAR = 0, 1
ASSEMBLY_MODES = ("corrected", "optimized", "reference")
ASSEMBLY_MODE_ENV = "HORNLAB_METAL_BEM_NATIVE_ASSEMBLY_MODE"
DENSE_SOLVE_DTYPE_ENV = "HORNLAB_METAL_BEM_NATIVE_DENSE_SOLVE_DTYPE"


class _Geometry(C.Structure):
    _fields_ = [
        ("n_vertices", C.c_int32),
        ("n_triangles", C.c_int32),
        ("p1_dof_count", C.c_int32),
        ("symmetry", C.c_int32),
        ("vertices_3xn", C.POINTER(C.c_float)),
        ("triangles_3xm", C.POINTER(C.c_int32)),
        ("p1_local2global", C.POINTER(C.c_int32)),
        ("areas", C.POINTER(C.c_float)),
        ("normals_3xm", C.POINTER(C.c_float)),
        ("physical_tags", C.POINTER(C.c_int32)),
    ]


class _SessionInfo(C.Structure):
    _fields_ = [
        ("duffy_coincident", C.c_int64),
        ("duffy_edge", C.c_int64),
        ("duffy_vertex", C.c_int64),
        ("duffy_image_pairs", C.c_int64),
        ("duffy_unique_slots", C.c_int64),
        ("n_triangles", C.c_int32),
        ("p1_dof_count", C.c_int32),
        ("symmetry", C.c_int32),
        ("device_cu_count", C.c_int32),
        ("setup_seconds", C.c_double),
        ("device_name", C.c_char * 64),
    ]


class _SolveOptions(C.Structure):
    _fields_ = [
        ("solve_mode", C.c_int32),
        ("refine_max_iterations", C.c_int32),
        ("refine_tolerance", C.c_double),
        ("batch_frequencies", C.c_int32),
        ("estimate_rcond", C.c_int32),
        ("n_tags", C.c_int32),
        ("tags", C.POINTER(C.c_int32)),
        ("impedance_tag", C.c_int32),
        ("robin_beta", C.c_void_p),
        ("n_sources", C.c_int32),
        ("lu_graph", C.c_int32),
        ("chief_points", C.c_void_p),
        ("n_chief", C.c_int32),
        ("precise_geometry", C.c_int32),
        ("chief_weight", C.c_double),
        ("solve_trans", C.c_int32),
        ("reserved", C.c_int32 * 3),
    ]


class _CaseDiag(C.Structure):
    _fields_ = [
        ("assembly_s", C.c_double),
        ("dense_solve_s", C.c_double),
        ("field_s", C.c_double),
        ("lapack_info", C.c_int32),
        ("refine_iterations", C.c_int32),
        ("dense_solve_rcond", C.c_double),
        ("dense_solve_condition_1norm", C.c_double),
        ("refine_residual_rel", C.c_double),
        ("regular_assembly_s", C.c_double),
        ("singular_assembly_s", C.c_double),
        ("batch_index", C.c_int32),
        ("batch_size", C.c_int32),
        ("near_assembly_s", C.c_double),
        ("near_pairs", C.c_int64),
        ("chief_residual_rel", C.c_double),
        ("refine_stalled", C.c_int32),
        ("reserved_i", C.c_int32),
        ("rcond_s", C.c_double),
    ]


class _GroupCase(C.Structure):
    _fields_ = [
        ("n_freq", C.c_int32),
        ("k_real", C.c_void_p),
        ("field_k", C.c_void_p),
        ("q", C.c_void_p),
        ("n_obs", C.c_int32),
        ("obs_3xn", C.c_void_p),
        ("field", C.c_void_p),
        ("p", C.c_void_p),
        ("tag_avg", C.c_void_p),
    ]


# Every symbol the header declares (checked by the CPU test suite).
EXPORTED_SYMBOLS = (
    "hbem_abi_version",
    "hbem_last_error",
    "hbem_device_count",
    "hbem_probe_transcendental",
    "hbem_session_create",
    "hbem_session_destroy",
    "hbem_session_get_info",
    "hbem_session_set_near_field",
    "hbem_session_set_assembly_mode",
    "hbem_assemble",
    "hbem_solve_batch",
    "hbem_solve_batch_device",
    "hbem_evaluate_field",
    "hbem_dense_solve",
    "hbem_group_create",
    "hbem_group_destroy",
    "hbem_group_solve_device",
    "hbem_group_autotune",
    "hbem_dense_solve_noalloc",
    "hbem_adjoint_forward",
    "hbem_adjoint_backward",
)


def diag_to_dict(d: "_CaseDiag") -> dict[str, Any]:
    return {
        "assembly_seconds": float(d.assembly_s),
        "regular_assembly_seconds": float(d.regular_assembly_s),
        "singular_assembly_seconds": float(d.singular_assembly_s),
        "dense_solve_seconds": float(d.dense_solve_s),
        "field_seconds": float(d.field_s),
        "lapack_info": int(d.lapack_info),
        "dense_solve_rcond": float(d.dense_solve_rcond),
        "dense_solve_condition_1norm": float(d.dense_solve_condition_1norm),
        "dense_solve_refine_iterations": int(d.refine_iterations),
        "dense_solve_refine_residual_rel": float(d.refine_residual_rel),
        "batch_index": int(d.batch_index),
        "batch_size": int(d.batch_size),
        "near_assembly_seconds": float(d.near_assembly_s),
        "dense_solve_refine_stalled": bool(d.refine_stalled),
        "dense_solve_rcond_seconds": float(d.rcond_s),
    }

_lock = threading.Lock()
_lib: C.CDLL | None = None


def probe_transcendental_rate() -> float:
    """Sustained sin/cos/rsq operations per second of the device
    (hbem_probe_transcendental)."""
    out = C.c_double(0.0)
    _check(load_library().hbem_probe_transcendental(C.byref(out)), "probe_transcendental")
    return float(out.value)


def library_path() -> Path:
    override = os.environ.get("HBEM_LIBRARY")
    if override:
        return Path(override)
    return Path(__file__).resolve().parent / "_lib" / _LIB_NAME


def _f32p(a: NDArray | None):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _i32p(a: NDArray | None):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int32))


def _f64p(a: NDArray | None):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def load_library() -> C.CDLL:
    """Load libhbem.so (once).  Raises AssemblyBackendUnavailable if absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = library_path()
        if not path.exists():
            raise AssemblyBackendUnavailable(
                f"libhbem is not built ({path} missing); run __graft_entry__.build() or make"
            )
        # One HIP runtime per process: when PyTorch is installed its bundled
        # libamdhip64 is loaded first, so libhbem binds to it and device
        # pointers move between torch tensors and the library.  Loading the
        # system runtime first would leave torch unable to initialise the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(str(path))
        lib.hbem_abi_version.restype = C.c_int32
        lib.hbem_last_error.restype = C.c_char_p
        lib.hbem_device_count.restype = C.c_int32
        lib.hbem_probe_transcendental.argtypes = [C.POINTER(C.c_double)]
        lib.hbem_probe_transcendental.restype = C.c_int32
        lib.hbem_session_create.argtypes = [C.POINTER(_Geometry), C.POINTER(C.c_void_p)]
        lib.hbem_session_create.restype = C.c_int32
        lib.hbem_session_destroy.argtypes = [C.c_void_p]
        lib.hbem_session_destroy.restype = C.c_int32
        lib.hbem_session_get_info.argtypes = [C.c_void_p, C.POINTER(_SessionInfo)]
        lib.hbem_session_get_info.restype = C.c_int32
        lib.hbem_session_set_near_field.argtypes = [
            C.c_void_p, C.c_int32, C.c_double, C.POINTER(C.c_int64)]
        lib.hbem_session_set_near_field.restype = C.c_int32
        lib.hbem_session_set_assembly_mode.argtypes = [C.c_void_p, C.c_int32]
        lib.hbem_session_set_assembly_mode.restype = C.c_int32
        fp = C.POINTER(C.c_float)
        lib.hbem_assemble.argtypes = [C.c_void_p, C.c_float, C.c_float, fp, fp, fp, fp, fp, fp, fp, fp]
        lib.hbem_assemble.restype = C.c_int32
        lib.hbem_solve_batch.argtypes = [
            C.c_void_p, C.c_int32, fp, fp, fp, fp, fp, C.c_int32, fp,
            C.POINTER(_SolveOptions), fp, fp, fp, fp,
            C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(_CaseDiag),
        ]
        lib.hbem_solve_batch.restype = C.c_int32
        vp = C.c_void_p
        lib.hbem_solve_batch_device.argtypes = [
            C.c_void_p, C.c_int32, vp, vp, vp, vp, C.c_int32, vp, C.POINTER(_SolveOptions),
            vp, vp, vp, C.POINTER(_CaseDiag),
        ]
        lib.hbem_solve_batch_device.restype = C.c_int32
        lib.hbem_adjoint_forward.argtypes = [
            C.c_void_p, C.c_int32, vp, vp, vp, vp, C.c_int32, vp, C.POINTER(_SolveOptions),
            vp, vp, vp, C.POINTER(_CaseDiag),
        ]
        lib.hbem_adjoint_forward.restype = C.c_int32
        lib.hbem_adjoint_backward.argtypes = [C.c_void_p, vp, vp, vp, C.c_int32]
        lib.hbem_adjoint_backward.restype = C.c_int32
        lib.hbem_evaluate_field.argtypes = [
            C.c_void_p, C.c_int32, fp, fp, fp, fp, fp, C.c_int32, fp, fp, fp,
        ]
        lib.hbem_evaluate_field.restype = C.c_int32
        lib.hbem_dense_solve.argtypes = [
            C.c_int32, C.c_int32, fp, fp, fp, fp, C.POINTER(_SolveOptions), fp, fp,
            C.POINTER(_CaseDiag),
        ]
        lib.hbem_dense_solve.restype = C.c_int32
        lib.hbem_group_create.argtypes = [C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]
        lib.hbem_group_create.restype = C.c_int32
        lib.hbem_group_destroy.argtypes = [C.c_void_p]
        lib.hbem_group_destroy.restype = C.c_int32
        lib.hbem_group_solve_device.argtypes = [
            C.c_void_p, C.POINTER(_GroupCase), C.POINTER(_SolveOptions), C.POINTER(_CaseDiag),
            C.POINTER(C.c_double), C.POINTER(C.c_double),
        ]
        lib.hbem_group_solve_device.restype = C.c_int32
        lib.hbem_group_autotune.argtypes = [
            C.c_void_p, C.POINTER(_GroupCase), C.POINTER(_SolveOptions), C.POINTER(C.c_int32),
            C.POINTER(C.c_double),
        ]
        lib.hbem_group_autotune.restype = C.c_int32
        lib.hbem_dense_solve_noalloc.argtypes = [
            C.c_int32, C.c_int32, vp, vp, vp, vp, C.POINTER(_SolveOptions), vp, vp, vp, vp,
        ]
        lib.hbem_dense_solve_noalloc.restype = C.c_int32
        if lib.hbem_abi_version() != HBEM_ABI_VERSION:
            raise AssemblyBackendUnavailable(
                f"libhbem ABI {lib.hbem_abi_version()} != expected {HBEM_ABI_VERSION}; rebuild"
            )
        _lib = lib
        return lib


NEAR_QUADRATURE_ENV = "HORNLAB_METAL_BEM_NATIVE_NEAR_QUADRATURE"


def near_quadrature_from_env(environ: Any = None) -> tuple[int, float] | None:
    """(level, threshold) of the near-field quadrature requested through
    HORNLAB_METAL_BEM_NATIVE_NEAR_QUADRATURE, or None when it is unset, empty
    or "0" (requestedNearQuadratureConfig, main.swift:5024-5055): "1"/"2"
    give that level at threshold 1.5, "<level>:<threshold>" a finite positive
    threshold; anything else is an error naming the variable."""
    env = os.environ if environ is None else environ
    raw = env.get(NEAR_QUADRATURE_ENV)
    if raw is None or raw == "" or raw == "0":
        return None

    def level_of(text: str) -> int | None:
        try:
            level = int(text)
        except ValueError:
            return None
        return level if 1 <= level <= 2 and text.strip() == text else None

    level = level_of(raw)
    if level is not None:
        return level, 1.5
    parts = raw.split(":")
    if len(parts) == 2:
        level = level_of(parts[0])
        try:
            threshold = float(parts[1])
        except ValueError:
            threshold = float("nan")
        if level is not None and np.isfinite(threshold) and threshold > 0.0:
            return level, threshold
    raise RuntimeError(
        f"{NEAR_QUADRATURE_ENV} must be unset, '0', '1', '2', "
        "or '<level>:<positive threshold>' with level 1...2"
    )


def _check(rc: int, op: str) -> None:
    if rc == 0:
        return
    lib = load_library()
    msg = (lib.hbem_last_error() or b"").decode("utf-8", "replace")
    if rc == -2:
        raise AssemblyBackendUnavailable(f"HIP BEM backend failed during {op}: {msg}")
    raise RuntimeError(f"HIP BEM backend failed during {op}: {msg}")


@dataclass(frozen=True)
class NativeRuntimeStatus:
    library_path: Path | None
    library_loaded: bool
    device_count: int
    unavailable_reasons: tuple[str, ...] = ()

    @property
    def available(self) -> bool:
        return self.library_loaded and self.device_count > 0


def discover_native_runtime(run_smoke_test: bool = False) -> NativeRuntimeStatus:
    """Library + device probe (stands in for native.py:95-179)."""
    path = library_path()
    try:
        lib = load_library()
    except (AssemblyBackendUnavailable, OSError) as exc:
        return NativeRuntimeStatus(path, False, 0, (str(exc),))
    count = int(lib.hbem_device_count())
    reasons = () if count > 0 else ("no gfx950 (MI355X) device is visible",)
    return NativeRuntimeStatus(path, True, count, reasons)


def _complex_out(re: NDArray[np.float32], im: NDArray[np.float32]) -> NDArray[np.complex64]:
    return (re + 1j * im).astype(np.complex64)


@dataclass
class ExtraSourceSolveResult:
    """Outputs of one additional right-hand side of a multi-source case,
    sharing the case's assembled and factored operator (the reference's
    ExtraSourceSolveResult, hornlab_metal_bem/metal/session.py:329-351, as
    in-memory arrays)."""

    source_index: int
    pressure: NDArray[np.complex64] | None
    field: NDArray[np.complex64]
    field_s: float = 0.0
    field_row_index: int | None = None
    impedance: complex | None = None
    surface_pressure_avg: dict[int, complex] | None = None


@dataclass
class DenseSolveFieldResult:
    """Per-frequency solve + field result (the fields of the reference's
    DenseSolveFieldResult, hornlab_metal_bem/metal/session.py:353-376, as
    in-memory arrays)."""

    frequency_hz: float
    pressure: NDArray[np.complex64] | None
    field: NDArray[np.complex64]
    assembly_s: float
    dense_solve_s: float
    field_s: float
    lapack_info: int
    impedance: complex | None
    surface_pressure_avg: dict[int, complex] | None
    diagnostics: dict[str, Any] = field(default_factory=dict)
    field_row_index: int | None = None
    extra_sources: tuple[ExtraSourceSolveResult, ...] = ()


class HipBemSession:
    """Resident GPU session: geometry, singular-pair plan and workspaces stay
    on the device across calls."""

    def __init__(self, handle: C.c_void_p, buffers: GeometryBuffers, symmetry_plane: str | None,
                 extra_env: dict[str, str] | None = None):
        self._handle = handle
        self.geometry_buffers = buffers
        self.symmetry_plane = symmetry_plane
        self._closed = False
        self._near: tuple[int, float] | None = None
        self._near_pairs = 0
        self._env = dict(extra_env or {})
        self.assembly_mode = "corrected"
        self._rule = ASSEMBLY_CORRECTED
        mode = self._setting(ASSEMBLY_MODE_ENV)
        if mode is not None:
            self.set_assembly_mode(mode)

    def _setting(self, key: str) -> str | None:
        """A helper setting: the session's extra_env first, then the process
        environment (the reference passes both to its helper subprocess)."""
        if key in self._env:
            return self._env[key]
        return os.environ.get(key)

    @classmethod
    def create_session(
        cls,
        grid: Any | None = None,
        physical_tags: Any | None = None,
        p1_space: Any | None = None,
        dp0_space: Any | None = None,
        *,
        geometry_buffers: GeometryBuffers | None = None,
        runtime_config: Any | None = None,
        work_dir: Any | None = None,
        session_id: str | None = None,
        keep_artifacts: bool = False,
        regular_triangle_order: int = 4,
        duffy_1d_order: int = 4,
        precision: str = "complex64",
        symmetry_plane: str | None = None,
        aperture_tag: int | None = None,
        velocity_source_tags: Any | None = None,
        check_open_edges: bool = True,
        runtime_status: Any | None = None,
        extra_env: dict[str, str] | None = None,
    ) -> "HipBemSession":
        """Resident session for one mesh (reference create_session,
        native.py:252-377).  ``geometry_buffers`` may be passed positionally
        in place of ``grid``.  work_dir, session_id, keep_artifacts,
        runtime_config and runtime_status belong to the reference's file IPC
        and helper discovery and change nothing here; ``extra_env`` carries
        the helper settings this path honours (assembly mode, dense-solve
        dtype, near-field quadrature)."""
        if isinstance(grid, GeometryBuffers):
            if geometry_buffers is not None:
                raise ValueError("pass geometry buffers once")
            geometry_buffers, grid = grid, None
        if regular_triangle_order != 4 or duffy_1d_order != 4:
            raise ValueError(
                "regular_triangle_order and duffy_1d_order are fixed at 4; "
                "the HIP kernels do not take quadrature overrides"
            )
        if precision != "complex64":
            raise ValueError("precision must be 'complex64' (the FP32 assembly of the HIP path)")
        if aperture_tag is not None:
            raise AssemblyBackendUnavailable(
                "aperture_tag (coupled infinite baffle) is not supported by the HIP path"
            )
        if geometry_buffers is None:
            if grid is None or physical_tags is None:
                raise ValueError("create_session needs geometry_buffers or (grid, physical_tags)")
            from .geometry import build_geometry_buffers
            from .mesh import make_pure_function_spaces

            if p1_space is None or dp0_space is None:
                p1_space, dp0_space = make_pure_function_spaces(grid)
            geometry_buffers = build_geometry_buffers(grid, physical_tags, p1_space, dp0_space)
        plane = validate_native_symmetry_plane(
            geometry_buffers, symmetry_plane, check_open_edges=check_open_edges
        )
        lib = load_library()
        b = geometry_buffers
        keep = (
            np.ascontiguousarray(b.vertices_3xn_f32, dtype=np.float32),
            np.ascontiguousarray(b.triangles_3xm_i32, dtype=np.int32),
            np.ascontiguousarray(b.p1_local2global_i32, dtype=np.int32),
            np.ascontiguousarray(b.triangle_areas_f32, dtype=np.float32),
            np.ascontiguousarray(b.triangle_normals_3xm_f32, dtype=np.float32),
            np.ascontiguousarray(b.physical_tags_i32, dtype=np.int32),
        )
        g = _Geometry(
            b.n_vertices, b.n_triangles, b.p1_dof_count, SYMMETRY_CODES[plane],
            _f32p(keep[0]), _i32p(keep[1]), _i32p(keep[2]), _f32p(keep[3]), _f32p(keep[4]),
            _i32p(keep[5]),
        )
        handle = C.c_void_p()
        _check(lib.hbem_session_create(C.byref(g), C.byref(handle)), "create_session")
        try:
            return cls(handle, b, plane, extra_env)
        except BaseException:
            lib.hbem_session_destroy(handle)
            raise

    def set_assembly_mode(self, mode: str) -> None:
        """metal_native_assembly_mode of later operations: "corrected" (6x6
        rule + Sauter-Schwab on touching pairs), "optimized" (6x6 rule on
        every pair) or "reference" (the 6x6 rule, with the Sauter-Schwab
        corrections for complex-k and Robin cases, main.swift:8139-8186).
        "parity" needs the reference's CPU assembly and is refused."""
        if mode == "parity":
            raise AssemblyBackendUnavailable(
                "metal_native_assembly_mode='parity' cross-checks the reference's CPU assembly "
                "inside its helper; the HIP path has no CPU assembly to compare against"
            )
        if mode not in ASSEMBLY_MODES:
            raise ValueError(
                "metal_native_assembly_mode must be 'corrected', 'optimized', 'reference' or 'parity'"
            )
        self.assembly_mode = mode

    def _apply_assembly_rule(self, corrections: bool) -> str:
        """Selects the kernels' quadrature for an operation; ``corrections``:
        the operation has complex k or Robin walls.  Returns the rule name
        reported in the diagnostics."""
        self._ensure_open()
        mode = self._setting(ASSEMBLY_MODE_ENV)
        if mode is not None and mode != self.assembly_mode:
            self.set_assembly_mode(mode)
        if self.assembly_mode == "corrected" or (self.assembly_mode == "reference" and corrections):
            rule, name = ASSEMBLY_CORRECTED, "corrected"
        else:
            rule, name = ASSEMBLY_REGULAR, "regular_6x6"
        if rule != self._rule:
            _check(load_library().hbem_session_set_assembly_mode(self._handle, rule),
                   "set_assembly_mode")
            self._rule = rule
        return name

    # -- lifecycle ---------------------------------------------------------
    def close(self) -> None:
        if not self._closed and self._handle:
            load_library().hbem_session_destroy(self._handle)
        self._closed = True

    def __enter__(self) -> "HipBemSession":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def _ensure_open(self) -> None:
        if self._closed:
            raise RuntimeError("HIP BEM session is closed")

    def info(self) -> dict[str, Any]:
        self._ensure_open()
        out = _SessionInfo()
        _check(load_library().hbem_session_get_info(self._handle, C.byref(out)), "get_info")
        return {
            "planned_pairs": {
                "coincident": int(out.duffy_coincident),
                "edge": int(out.duffy_edge),
                "vertex": int(out.duffy_vertex),
                "total": int(out.duffy_coincident + out.duffy_edge + out.duffy_vertex),
            },
            "image_adjacent_pairs": int(out.duffy_image_pairs),
            "unique_triplets": int(out.duffy_unique_slots),
            "raw_triplets_if_expanded": 9
            * int(out.duffy_coincident + out.duffy_edge + out.duffy_vertex),
            "n_triangles": int(out.n_triangles),
            "p1_dof_count": int(out.p1_dof_count),
            "symmetry_code": int(out.symmetry),
            "device_cu_count": int(out.device_cu_count),
            "device_name": out.device_name.decode("utf-8", "replace"),
            "resident_context_seconds": float(out.setup_seconds),
        }

    # -- operations ---------------------------------------------------------
    def set_near_quadrature(self, level: int, threshold: float = 1.5) -> int:
        """Turns the near-field correction on (level 1 or 2) or off (level 0)
        for later assemblies; returns the number of near pairs."""
        self._ensure_open()
        count = C.c_int64(0)
        _check(
            load_library().hbem_session_set_near_field(
                self._handle, int(level), float(threshold), C.byref(count)),
            "set_near_quadrature",
        )
        self._near = (int(level), float(threshold)) if level > 0 else None
        self._near_pairs = int(count.value)
        return self._near_pairs

    def _sync_near_quadrature(self) -> None:
        """Applies the environment's near-field setting, read per operation as
        the reference helper does; with the variable unset the session keeps
        the setting of set_near_quadrature (off by default).  The session's
        extra_env takes precedence over the process environment."""
        if NEAR_QUADRATURE_ENV in self._env:
            want = near_quadrature_from_env(self._env)
        elif NEAR_QUADRATURE_ENV in os.environ:
            want = near_quadrature_from_env()
        else:
            return
        if want != self._near:
            if want is None:
                self.set_near_quadrature(0)
            else:
                self.set_near_quadrature(*want)

    def near_quadrature_diagnostics(self, seconds: float = 0.0) -> dict[str, Any] | None:
        if self._near is None:
            return None
        return {"level": self._near[0], "threshold": self._near[1],
                "pair_count": self._near_pairs, "seconds": float(seconds)}

    def _beta_rows(self, payloads: list[dict[int, complex] | None]) -> NDArray[np.complex64] | None:
        """Per-triangle Robin admittance rows [len(payloads), T] from {tag:
        beta} dicts (the reference's impedance payload, native.py:870-902);
        None when no case has a Robin wall."""
        if not any(payloads):
            return None
        tags = np.asarray(self.geometry_buffers.physical_tags_i32, np.int64)
        rows = np.zeros((len(payloads), tags.size), np.complex64)
        for i, d in enumerate(payloads):
            for tag, beta in (d or {}).items():
                b = complex(beta)
                if int(tag) < 0:
                    raise ValueError("impedance_sources tags must be non-negative")
                if not (np.isfinite(b.real) and np.isfinite(b.imag)):
                    raise ValueError("impedance_sources values must be finite")
                rows[i, tags == int(tag)] = np.complex64(b)
        return rows

    def assemble_standard_neumann(
        self,
        frequency_hz: float,
        k_real: float,
        neumann_dp0: NDArray[Any],
        *,
        k_imag: float = 0.0,
        impedance_sources: dict[int, complex] | None = None,
        operation_id: str | None = None,
    ) -> tuple[NDArray[np.complex64], NDArray[np.complex64]]:
        """Corrected system at one wavenumber: (A (n, n) row-major, rhs (n,)).
        ``impedance_sources`` {tag: beta} adds the Robin fold."""
        self._ensure_open()
        self._sync_near_quadrature()
        q = np.asarray(neumann_dp0)
        T, n = self.geometry_buffers.n_triangles, self.geometry_buffers.p1_dof_count
        if q.shape != (T,):
            raise ValueError(f"neumann_dp0 must have shape {(T,)}, got {q.shape}")
        if not np.all(np.isfinite(q)):
            raise ValueError("neumann_dp0 must contain only finite values")
        qre = np.ascontiguousarray(q.real, dtype=np.float32)
        qim = np.ascontiguousarray(q.imag, dtype=np.float32)
        beta = self._beta_rows([impedance_sources])
        self._apply_assembly_rule(float(k_imag) != 0.0 or beta is not None)
        bre = np.ascontiguousarray(beta[0].real, np.float32) if beta is not None else None
        bim = np.ascontiguousarray(beta[0].imag, np.float32) if beta is not None else None
        are = np.empty((n, n), np.float32)
        aim = np.empty((n, n), np.float32)
        rre = np.empty(n, np.float32)
        rim = np.empty(n, np.float32)
        _check(
            load_library().hbem_assemble(
                self._handle, float(np.float32(k_real)), float(np.float32(k_imag)), _f32p(qre),
                _f32p(qim), _f32p(bre), _f32p(bim), _f32p(are), _f32p(aim), _f32p(rre), _f32p(rim),
            ),
            "assemble_standard_neumann",
        )
        return _complex_out(are, aim), _complex_out(rre, rim)

    def assemble_solve_evaluate_standard_neumann_batch(
        self,
        frequency_hz: NDArray[Any],
        k_real: NDArray[Any],
        neumann_dp0: NDArray[Any],
        observation_points: NDArray[Any],
        *,
        k_imag_f32: NDArray[Any] | None = None,
        field_k_real: NDArray[Any] | None = None,
        source_tags: list[int] | tuple[int, ...] | None = None,
        impedance_source_tag: int | None = None,
        write_surface_pressure: bool = True,
        solve_mode: str = "fp32",
        refine_max_iterations: int = 8,
        refine_tolerance: float = 1e-7,
        batch_frequencies: int | None = None,
        estimate_rcond: bool = True,
        on_case_result: Callable[[int, DenseSolveFieldResult], bool | None] | None = None,
        dense_solve_dtype: str | None = None,
        impedance_sources: dict[int, complex] | list[dict[int, complex]] | None = None,
        extra_neumann_dp0: NDArray[Any] | None = None,
        extra_impedance_source_tags: list[int | None] | tuple[int | None, ...] | None = None,
        lu_graph: bool = False,
        chief_points: NDArray[Any] | None = None,
        chief_weight: float = 1.0,
        precise_geometry: bool = False,
        batch_id: str = "batch",
        operation_id: str = "assembly-solve-field-batch",
        write_batched_field: bool = False,
    ) -> list[DenseSolveFieldResult]:
        """Assemble, solve and evaluate every frequency; returns one result per
        frequency in order.  ``on_case_result(i, result)`` is called in order;
        returning exactly False stops after that case.  ``impedance_sources``
        ({tag: beta}, or one dict per frequency) adds Robin walls;
        ``extra_neumann_dp0`` (n_extra, F, T) adds right-hand sides that share
        each case's operator and LU (multi-RHS, native.py:775-783).
        ``dense_solve_dtype="float64"`` (or the session's
        HORNLAB_METAL_BEM_NATIVE_DENSE_SOLVE_DTYPE) solves each system to FP64
        accuracy (FP32 LU + GMRES-IR on the FP64 residual) in place of the
        reference's zgesv.  batch_id, operation_id and write_batched_field name
        the reference's IPC files and do not change results here (fields are
        always returned as arrays)."""
        self._ensure_open()
        self._sync_near_quadrature()
        freqs = np.asarray(frequency_hz, dtype=np.float64).reshape(-1)
        k = np.ascontiguousarray(np.asarray(k_real, dtype=np.float32).reshape(-1))
        F = int(freqs.size)
        if F == 0:
            raise ValueError("frequency_hz must be a non-empty 1D array")
        if k.shape != freqs.shape:
            raise ValueError("k_real must have the same shape as frequency_hz")
        kim = (
            np.zeros(F, np.float32)
            if k_imag_f32 is None
            else np.ascontiguousarray(np.asarray(k_imag_f32, dtype=np.float32).reshape(-1))
        )
        if kim.shape != freqs.shape:
            raise ValueError("k_imag_f32 must have the same shape as frequency_hz")
        if np.any(kim < 0):
            raise ValueError("k_imag_f32 must be non-negative")
        fk = k if field_k_real is None else np.ascontiguousarray(np.asarray(field_k_real, np.float32))
        T, n = self.geometry_buffers.n_triangles, self.geometry_buffers.p1_dof_count
        q = np.asarray(neumann_dp0)
        if q.shape != (F, T):
            raise ValueError(f"neumann_dp0 must have shape {(F, T)}, got {q.shape}")
        if not np.iscomplexobj(q):
            raise ValueError("neumann_dp0 must be complex")
        if not np.all(np.isfinite(q)):
            raise ValueError("neumann_dp0 must contain only finite values")
        n_extra = 0
        if extra_neumann_dp0 is not None:
            xq = np.asarray(extra_neumann_dp0)
            if xq.ndim != 3 or xq.shape[1:] != (F, T):
                raise ValueError(f"extra_neumann_dp0 must have shape (n_extra, {F}, {T}), got {xq.shape}")
            if xq.shape[0] == 0:
                raise ValueError("extra_neumann_dp0 must be None or non-empty")
            if not np.iscomplexobj(xq):
                raise ValueError("extra_neumann_dp0 must be complex")
            if not np.all(np.isfinite(xq)):
                raise ValueError("extra_neumann_dp0 must contain only finite values")
            n_extra = int(xq.shape[0])
            q = np.concatenate([q[:, None, :], np.transpose(xq, (1, 0, 2))], axis=1)  # (F, S, T)
        else:
            q = q[:, None, :]
        if extra_impedance_source_tags is not None:
            if n_extra == 0:
                raise ValueError("extra_impedance_source_tags requires extra_neumann_dp0")
            if len(extra_impedance_source_tags) != n_extra:
                raise ValueError("extra_impedance_source_tags must have one entry per extra source")
        S = 1 + n_extra
        if isinstance(impedance_sources, list):
            if len(impedance_sources) != F:
                raise ValueError(f"impedance_sources list must have one dict per frequency ({F})")
            payloads = list(impedance_sources)
        else:
            payloads = [impedance_sources] * F
        beta = self._beta_rows(payloads)
        obs = np.asarray(observation_points, dtype=np.float32)
        if obs.ndim != 2 or 3 not in obs.shape:
            raise ValueError("observation_points must have shape (3, n_obs) or (n_obs, 3)")
        if obs.shape[0] != 3:
            obs = obs.T
        obs = np.ascontiguousarray(obs, dtype=np.float32)
        n_obs = int(obs.shape[1])
        if dense_solve_dtype is None:
            dense_solve_dtype = self._setting(DENSE_SOLVE_DTYPE_ENV) or "float32"
        if dense_solve_dtype not in ("float32", "float64"):
            raise ValueError("dense_solve_dtype must be 'float32' or 'float64'")
        if dense_solve_dtype == "float64" and solve_mode in ("fp32", "fp32_ir"):
            solve_mode = "fp64"
        if solve_mode not in SOLVE_MODES:
            raise ValueError(f"solve_mode must be one of {sorted(SOLVE_MODES)}")
        if solve_mode == "fp64" and int(refine_max_iterations) <= 0:
            raise ValueError("the float64 dense solve needs refine_max_iterations > 0")
        rule = self._apply_assembly_rule(bool(np.any(kim != 0)) or beta is not None)
        user_tags = [int(t) for t in (source_tags or [])]
        imp_tag = int(impedance_source_tag) if impedance_source_tag is not None else (
            int(min(user_tags)) if user_tags else 2
        )
        extra_tags = [int(t) for t in (extra_impedance_source_tags or []) if t is not None]
        all_tags = list(dict.fromkeys(user_tags + extra_tags))
        tags = np.ascontiguousarray(all_tags, dtype=np.int32)
        opts = _SolveOptions()
        opts.solve_mode = SOLVE_MODES[solve_mode]
        opts.refine_max_iterations = int(refine_max_iterations)
        opts.refine_tolerance = float(refine_tolerance)
        opts.batch_frequencies = int(batch_frequencies or 0)
        opts.estimate_rcond = 1 if estimate_rcond else 0
        opts.n_tags = int(tags.size)
        opts.tags = _i32p(tags) if tags.size else None
        opts.impedance_tag = imp_tag
        opts.n_sources = S
        opts.lu_graph = 1 if lu_graph else 0
        opts.precise_geometry = 1 if precise_geometry else 0
        chief = None
        if chief_points is not None:
            chief = np.asarray(chief_points, dtype=np.float32)
            if chief.ndim != 2 or chief.shape[0] != 3:
                raise ValueError("chief_points must have shape (3, m)")
            if chief.shape[1] == 0:
                raise ValueError("chief_points must be non-empty when set")
            if not np.all(np.isfinite(chief)):
                raise ValueError("chief_points must be finite")
            if not (np.isfinite(chief_weight) and float(chief_weight) > 0):
                raise ValueError("chief_weight must be finite and positive")
            chief = np.ascontiguousarray(chief)
            opts.chief_points = chief.ctypes.data
            opts.n_chief = int(chief.shape[1])
            opts.chief_weight = float(chief_weight)
            if solve_mode == "fp32":
                solve_mode = "fp32_ir"
                opts.solve_mode = SOLVE_MODES[solve_mode]
        beta_f = None
        if beta is not None:
            beta_f = np.ascontiguousarray(beta.astype(np.complex64)).view(np.float32)
            opts.robin_beta = beta_f.ctypes.data
        qre = np.ascontiguousarray(q.real, dtype=np.float32)
        qim = np.ascontiguousarray(q.imag, dtype=np.float32)
        fre = np.zeros((F, S, n_obs), np.float32)
        fim = np.zeros((F, S, n_obs), np.float32)
        pre = np.zeros((F, S, n), np.float32)
        pim = np.zeros((F, S, n), np.float32)
        imp = np.zeros((F, S, 2), np.float64)
        pav = np.zeros((F, S, max(1, tags.size), 2), np.float64)
        # With a streaming callback the sweep runs in resident chunks and the
        # callback sees each chunk's cases before the next chunk is solved
        # (per-case streaming and early stop, native.py:1362-1447).
        chunk = F if on_case_result is None else max(1, int(batch_frequencies or 8))
        self._last_rule = rule
        results: list[DenseSolveFieldResult] = []
        for f0 in range(0, F, chunk):
            f1 = min(F, f0 + chunk)
            if not self._solve_chunk(f0, f1, k, kim, fk, qre, qim, obs, opts, beta_f, fre, fim,
                                     pre, pim, imp, pav, tags, freqs, solve_mode, payloads, n_extra,
                                     S, n, n_obs, user_tags, extra_impedance_source_tags,
                                     write_surface_pressure, on_case_result, results):
                break
        return results

    def _solve_chunk(self, f0, f1, k, kim, fk, qre, qim, obs, opts, beta_f, fre, fim, pre, pim,
                     imp, pav, tags, freqs, solve_mode, payloads, n_extra, S, n, n_obs, user_tags,
                     extra_impedance_source_tags, write_surface_pressure, on_case_result,
                     results) -> bool:
        """Solves frequencies f0..f1-1 (one library call) and appends their
        results; False when the callback asked to stop."""
        Fc = f1 - f0
        if beta_f is not None:
            opts.robin_beta = beta_f[f0:f1].ctypes.data
        diag = (_CaseDiag * Fc)()
        kc, kimc = np.ascontiguousarray(k[f0:f1]), np.ascontiguousarray(kim[f0:f1])
        _check(
            load_library().hbem_solve_batch(
                self._handle, Fc, _f32p(kc), _f32p(kimc) if np.any(kimc != 0) else None,
                _f32p(np.ascontiguousarray(fk[f0:f1])), _f32p(qre[f0:f1]), _f32p(qim[f0:f1]), n_obs,
                _f32p(obs), C.byref(opts), _f32p(fre[f0:f1]), _f32p(fim[f0:f1]), _f32p(pre[f0:f1]),
                _f32p(pim[f0:f1]), _f64p(imp[f0:f1]), _f64p(pav[f0:f1]) if tags.size else None, diag,
            ),
            "assemble_solve_evaluate_standard_neumann_batch",
        )
        for i in range(f0, f1):
            d = diag[i - f0]
            diagnostics: dict[str, Any] = {
                "implementation": "hip_gfx950_resident_assembly_lu_field_batch",
                "assembly_mode": self.assembly_mode,
                "assembly_rule": self._last_rule,
                "lapack_info": int(d.lapack_info),
                "dense_solve_condition_1norm": float(d.dense_solve_condition_1norm),
                "dense_solve_dtype": ("float32" if solve_mode == "fp32" else
                                      "float64" if solve_mode == "fp64" else solve_mode),
                "dense_solve_method": _SOLVE_METHODS[solve_mode],
                "dense_solve_refine_iterations": int(d.refine_iterations),
                "field_k_real_f32": float(fk[i]),
                "assembly_k_imag_f32": float(kim[i]),
                "complex_k": bool(kim[i] != 0),
                "symmetry_plane": self.symmetry_plane,
                "pressure_shape": [n],
                "field_shape": [n_obs],
                "robin_boundary": bool(payloads[i]),
                "multi_source": n_extra > 0,
                "n_sources": S,
            }
            diagnostics["regular_assembly_seconds"] = float(d.regular_assembly_s)
            diagnostics["singular_assembly_seconds"] = float(d.singular_assembly_s)
            near = self.near_quadrature_diagnostics(float(d.near_assembly_s))
            if near is not None:
                diagnostics["near_quadrature"] = near
            diagnostics["resident_batch"] = {"index": int(d.batch_index), "size": int(d.batch_size)}
            if opts.n_chief > 0:
                diagnostics["chief_points"] = True
                diagnostics["chief_points_count"] = int(opts.n_chief)
                diagnostics["chief_residual_rel"] = float(d.chief_residual_rel)
            if d.dense_solve_rcond >= 0:
                diagnostics["dense_solve_rcond"] = float(d.dense_solve_rcond)
            if d.refine_residual_rel >= 0:
                diagnostics["dense_solve_refine_residual_rel"] = float(d.refine_residual_rel)
                diagnostics["dense_solve_refine_stalled"] = bool(d.refine_stalled)
            if d.dense_solve_rcond >= 0:
                diagnostics["dense_solve_rcond_seconds"] = float(d.rcond_s)
            def avg(src: int) -> dict[int, complex]:
                return {int(t): complex(pav[i, src, j, 0], pav[i, src, j, 1]) for j, t in enumerate(tags)}

            extras = []
            for e in range(n_extra):
                etag = None
                if extra_impedance_source_tags is not None:
                    etag = extra_impedance_source_tags[e]
                e_imp = (
                    complex(imp[i, 1 + e, 0], imp[i, 1 + e, 1])
                    if etag is None
                    else avg(1 + e).get(int(etag))
                )
                extras.append(
                    ExtraSourceSolveResult(
                        source_index=e,
                        pressure=_complex_out(pre[i, 1 + e], pim[i, 1 + e]) if write_surface_pressure else None,
                        field=_complex_out(fre[i, 1 + e], fim[i, 1 + e]),
                        field_s=float(d.field_s) / S,
                        field_row_index=i,
                        impedance=e_imp,
                        surface_pressure_avg={t: v for t, v in avg(1 + e).items() if t in user_tags} or None,
                    )
                )
            res = DenseSolveFieldResult(
                frequency_hz=float(freqs[i]),
                pressure=_complex_out(pre[i, 0], pim[i, 0]) if write_surface_pressure else None,
                field=_complex_out(fre[i, 0], fim[i, 0]),
                assembly_s=float(d.assembly_s),
                dense_solve_s=float(d.dense_solve_s),
                field_s=float(d.field_s),
                lapack_info=int(d.lapack_info),
                impedance=complex(imp[i, 0, 0], imp[i, 0, 1]),
                surface_pressure_avg=(
                    {t: v for t, v in avg(0).items() if t in user_tags} if user_tags else None
                ),
                diagnostics=diagnostics,
                field_row_index=i,
                extra_sources=tuple(extras),
            )
            results.append(res)
            if on_case_result is not None and on_case_result(i, res) is False:
                return False
        return True

    def solve_batch_device(
        self,
        n_freq: int,
        k_real_ptr: int,
        field_k_ptr: int,
        q_ptr: int,
        n_obs: int,
        obs_ptr: int,
        field_ptr: int,
        tag_avg_ptr: int,
        *,
        k_imag_ptr: int | None = None,
        p_ptr: int | None = None,
        source_tags: list[int] | tuple[int, ...] = (),
        impedance_source_tag: int = 2,
        solve_mode: str = "fp32",
        refine_max_iterations: int = 8,
        refine_tolerance: float = 1e-7,
        batch_frequencies: int | None = None,
        estimate_rcond: bool = False,
        robin_beta_ptr: int | None = None,
        n_sources: int = 1,
        lu_graph: bool = False,
        precise_geometry: bool = False,
    ) -> list[dict[str, Any]]:
        """Device-resident batch (hbem_solve_batch_device): every pointer is
        a device address on the session's GPU (e.g. ``tensor.data_ptr()``).
        q is [F][n_sources][T]; robin_beta_ptr an interleaved complex [F][T]
        admittance or None.  Returns the per-case diagnostics."""
        self._ensure_open()
        self._sync_near_quadrature()
        tags = np.ascontiguousarray(list(source_tags), dtype=np.int32)
        opts = _SolveOptions()
        opts.solve_mode = SOLVE_MODES[solve_mode]
        opts.refine_max_iterations = int(refine_max_iterations)
        opts.refine_tolerance = float(refine_tolerance)
        opts.batch_frequencies = int(batch_frequencies or 0)
        opts.estimate_rcond = 1 if estimate_rcond else 0
        opts.n_tags = int(tags.size)
        opts.tags = _i32p(tags) if tags.size else None
        opts.impedance_tag = int(impedance_source_tag)
        opts.robin_beta = robin_beta_ptr
        opts.n_sources = int(n_sources)
        opts.lu_graph = 1 if lu_graph else 0
        opts.precise_geometry = 1 if precise_geometry else 0
        diag = (_CaseDiag * int(n_freq))()
        _check(
            load_library().hbem_solve_batch_device(
                self._handle, int(n_freq), k_real_ptr, k_imag_ptr, field_k_ptr, q_ptr, int(n_obs),
                obs_ptr, C.byref(opts), field_ptr, p_ptr, tag_avg_ptr, diag,
            ),
            "solve_batch_device",
        )
        return [diag_to_dict(d) for d in diag]


    def _device_options(self, solve_mode, refine_max_iterations, refine_tolerance, estimate_rcond,
                        source_tags, impedance_source_tag, robin_beta_ptr, n_sources, keep):
        tags = np.ascontiguousarray(list(source_tags), dtype=np.int32)
        keep.append(tags)
        opts = _SolveOptions()
        opts.solve_mode = SOLVE_MODES[solve_mode]
        opts.refine_max_iterations = int(refine_max_iterations)
        opts.refine_tolerance = float(refine_tolerance)
        opts.batch_frequencies = 0
        opts.estimate_rcond = 1 if estimate_rcond else 0
        opts.n_tags = int(tags.size)
        opts.tags = _i32p(tags) if tags.size else None
        opts.impedance_tag = int(impedance_source_tag)
        opts.robin_beta = robin_beta_ptr
        opts.n_sources = int(n_sources)
        return opts

    def adjoint_forward_device(
        self,
        n_freq: int,
        k_real_ptr: int,
        field_k_ptr: int,
        q_ptr: int,
        n_obs: int,
        obs_ptr: int,
        field_ptr: int,
        tag_avg_ptr: int,
        *,
        k_imag_ptr: int | None = None,
        p_ptr: int | None = None,
        robin_beta_ptr: int | None = None,
        source_tags: list[int] | tuple[int, ...] = (),
        impedance_source_tag: int = 0,
        solve_mode: str = "fp32_ir",
        refine_max_iterations: int = 8,
        refine_tolerance: float = 1e-7,
    ) -> list[dict[str, Any]]:
        """hbem_adjoint_forward: solve_batch_device that keeps every
        frequency's operator, LU and solution resident for
        adjoint_backward_device."""
        self._ensure_open()
        keep: list[Any] = []
        opts = self._device_options(solve_mode, refine_max_iterations, refine_tolerance, False,
                                    source_tags, impedance_source_tag, robin_beta_ptr, 1, keep)
        diag = (_CaseDiag * int(n_freq))()
        _check(
            load_library().hbem_adjoint_forward(
                self._handle, int(n_freq), k_real_ptr, k_imag_ptr, field_k_ptr, q_ptr, int(n_obs),
                obs_ptr, C.byref(opts), field_ptr, p_ptr, tag_avg_ptr, diag,
            ),
            "adjoint_forward",
        )
        return [diag_to_dict(d) for d in diag]

    def adjoint_backward_device(self, grad_field_ptr: int, grad_vertices_ptr: int,
                                grad_q_ptr: int | None = None, refine_iterations: int = 3) -> None:
        """hbem_adjoint_backward: vertex (and Neumann) gradients of a real
        loss whose field cotangent is at grad_field_ptr."""
        self._ensure_open()
        _check(
            load_library().hbem_adjoint_backward(
                self._handle, grad_field_ptr, grad_vertices_ptr, grad_q_ptr, int(refine_iterations)),
            "adjoint_backward",
        )
    def evaluate_standard_exterior_batch(
        self,
        frequency_hz: NDArray[Any],
        k_real: NDArray[Any],
        pressure_p1: NDArray[Any],
        neumann_dp0: NDArray[Any],
        observation_points: NDArray[Any],
        *,
        batch_id: str = "batch",
        operation_id: str = "field-batch",
    ) -> NDArray[np.complex64]:
        """Field (F, n_obs) from given surface pressure and Neumann data
        (batch_id and operation_id name the reference's IPC files only)."""
        self._ensure_open()
        k = np.ascontiguousarray(np.asarray(k_real, dtype=np.float32).reshape(-1))
        F = int(k.size)
        T, n = self.geometry_buffers.n_triangles, self.geometry_buffers.p1_dof_count
        p = np.asarray(pressure_p1).reshape(F, n)
        q = np.asarray(neumann_dp0).reshape(F, T)
        obs = np.asarray(observation_points, dtype=np.float32)
        if obs.shape[0] != 3:
            obs = obs.T
        obs = np.ascontiguousarray(obs)
        n_obs = int(obs.shape[1])
        fre = np.zeros((F, n_obs), np.float32)
        fim = np.zeros((F, n_obs), np.float32)
        _check(
            load_library().hbem_evaluate_field(
                self._handle, F, _f32p(k),
                _f32p(np.ascontiguousarray(p.real, np.float32)),
                _f32p(np.ascontiguousarray(p.imag, np.float32)),
                _f32p(np.ascontiguousarray(q.real, np.float32)),
                _f32p(np.ascontiguousarray(q.imag, np.float32)),
                n_obs, _f32p(obs), _f32p(fre), _f32p(fim),
            ),
            "evaluate_standard_exterior_batch",
        )
        return _complex_out(fre, fim)


# The reference's class name for the same role.
MetalNativeStandardSession = HipBemSession


@dataclass
class GroupCase:
    """One member's device arrays for HipBemGroup.solve_device (device
    addresses, e.g. ``tensor.data_ptr()``): k_real, field_k [F]; q [F][T]
    interleaved complex; obs [3][n_obs]; field [F][n_obs] complex; p
    [F][n] complex or None; tag_avg [F][1 + n_tags][2] double."""

    n_freq: int
    k_real: int
    field_k: int
    q: int
    n_obs: int
    obs: int
    field: int
    tag_avg: int
    p: int | None = None


class HipBemGroup:
    """Several sessions solved together (hbem_group_*, the batched uneven
    sweep): one assembly launch over all members, size-classed batched LU on
    parallel streams, the whole sweep replayed from one hipGraph.  Borrows
    the sessions: close the group before them."""

    def __init__(self, sessions: list[HipBemSession]):
        if not sessions:
            raise ValueError("a group needs at least one session")
        for s in sessions:
            s._ensure_open()
        self.sessions = list(sessions)
        arr = (C.c_void_p * len(sessions))(*[s._handle for s in sessions])
        h = C.c_void_p()
        _check(load_library().hbem_group_create(len(sessions), arr, C.byref(h)), "group_create")
        self._handle = h
        self.granularity = 512
        self.tuning: dict[int, float] = {}

    def close(self) -> None:
        if self._handle:
            load_library().hbem_group_destroy(self._handle)
            self._handle = None

    def __enter__(self) -> "HipBemGroup":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def _args(self, cases, source_tags, impedance_source_tag, keep):
        if len(cases) != len(self.sessions):
            raise ValueError("one case per member session")
        arr = (_GroupCase * len(cases))()
        for i, c in enumerate(cases):
            arr[i] = _GroupCase(int(c.n_freq), c.k_real, c.field_k, c.q, int(c.n_obs), c.obs, c.field,
                                c.p, c.tag_avg)
        tags = np.ascontiguousarray(list(source_tags), dtype=np.int32)
        keep.append(tags)
        opts = _SolveOptions()
        opts.solve_mode = SOLVE_MODES["fp32"]
        opts.n_tags = int(tags.size)
        opts.tags = _i32p(tags) if tags.size else None
        opts.impedance_tag = int(impedance_source_tag)
        return arr, opts

    def autotune(self, cases: list[GroupCase], *, source_tags=(), impedance_source_tag: int = 2) -> dict:
        """Times one sweep per size-class step and keeps the fastest."""
        keep: list = []
        arr, opts = self._args(cases, source_tags, impedance_source_tag, keep)
        gran = C.c_int32(0)
        secs = (C.c_double * 8)()
        _check(load_library().hbem_group_autotune(self._handle, arr, C.byref(opts), C.byref(gran), secs),
               "group_autotune")
        self.granularity = abs(int(gran.value))
        self.split_assembly = int(gran.value) < 0
        steps = (64, 256, 512, 1024)
        self.tuning = {f"{g}{'/split' if i >= 4 else '/one-launch'}": float(secs[i])
                       for i, g in enumerate(steps + steps)}
        return {"granularity": self.granularity, "split_assembly": self.split_assembly,
                "seconds": self.tuning}

    def solve_device(self, cases: list[GroupCase], *, source_tags=(), impedance_source_tag: int = 2):
        """One sweep of every member.  Returns (per-member case diagnostics,
        member latencies in s, sweep seconds)."""
        keep: list = []
        arr, opts = self._args(cases, source_tags, impedance_source_tag, keep)
        total = sum(int(c.n_freq) for c in cases)
        diag = (_CaseDiag * total)()
        lat = (C.c_double * len(cases))()
        sweep = C.c_double(0.0)
        _check(load_library().hbem_group_solve_device(self._handle, arr, C.byref(opts), diag, lat,
                                                        C.byref(sweep)), "group_solve_device")
        out, i = [], 0
        for c in cases:
            out.append([diag_to_dict(diag[i + f]) for f in range(int(c.n_freq))])
            i += int(c.n_freq)
        return out, [float(x) for x in lat], float(sweep.value)


def dense_solve(
    A: NDArray[Any],
    b: NDArray[Any],
    *,
    solve_mode: str = "fp32",
    refine_max_iterations: int = 8,
    refine_tolerance: float = 1e-7,
    estimate_rcond: bool = True,
    trans: str = "N",
) -> tuple[NDArray[np.complex64], list[dict[str, Any]]]:
    """GPU LU solve of stacked systems A (S, n, n) x = b (S, n); trans "C"
    solves A^H x = b (FP32 LU, no refinement)."""
    if trans not in ("N", "C"):
        raise ValueError("trans must be 'N' or 'C'")
    A = np.asarray(A)
    b = np.asarray(b)
    if A.ndim == 2:
        A, b = A[None], b[None]
    S, n, n2 = A.shape
    if n != n2 or b.shape != (S, n):
        raise ValueError("A must be (S, n, n) and b (S, n)")
    opts = _SolveOptions()
    opts.solve_mode = SOLVE_MODES[solve_mode]
    opts.refine_max_iterations = int(refine_max_iterations)
    opts.refine_tolerance = float(refine_tolerance)
    opts.estimate_rcond = 1 if estimate_rcond else 0
    opts.solve_trans = 2 if trans == "C" else 0
    xr = np.zeros((S, n), np.float32)
    xi = np.zeros((S, n), np.float32)
    diag = (_CaseDiag * S)()
    _check(
        load_library().hbem_dense_solve(
            S, n,
            _f32p(np.ascontiguousarray(A.real, np.float32)),
            _f32p(np.ascontiguousarray(A.imag, np.float32)),
            _f32p(np.ascontiguousarray(b.real, np.float32)),
            _f32p(np.ascontiguousarray(b.imag, np.float32)),
            C.byref(opts), _f32p(xr), _f32p(xi), diag,
        ),
        "dense_solve",
    )
    info = [
        {
            "lapack_info": int(d.lapack_info),
            "dense_solve_rcond": float(d.dense_solve_rcond),
            "dense_solve_condition_1norm": float(d.dense_solve_condition_1norm),
            "dense_solve_refine_iterations": int(d.refine_iterations),
            "dense_solve_refine_residual_rel": float(d.refine_residual_rel),
            "dense_solve_refine_stalled": bool(d.refine_stalled),
            "dense_solve_s": float(d.dense_solve_s),
            "dense_solve_rcond_seconds": float(d.rcond_s),
        }
        for d in diag
    ]
    return _complex_out(xr, xi), info
