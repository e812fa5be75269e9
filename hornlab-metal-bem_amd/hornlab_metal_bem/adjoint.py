"""Waveguide shape gradients: the exterior field as a differentiable function
of the mesh vertices (and of the Neumann data) through torch.autograd.

The forward pass is the HIP pipeline (assembly, LU, field) with every
frequency's operator, LU factors and surface solution kept resident on the
GPU; the backward pass runs the hand-written adjoint kernels: the field
cotangent g gives lambda = A^{-H}(E^H g + R^H F^H g) through the same LU
(FP64-residual refinement), then the derivatives of the field and of
Re(lambda^H (V q - A p)) with respect to every quadrature point, Jacobian and
normal of the regular pairs, the Sauter-Schwab classes and the mass term.
The reference solver has no derivatives; the loss helpers follow its
directivity convention (sweep.py:523-532).

PyTorch here only holds device memory and the autograd graph; there is no
CPU fallback.
"""
from __future__ import annotations

from dataclasses import dataclass, field as dc_field
from types import SimpleNamespace
from typing import Any

import numpy as np
import torch

from ._constants import REFERENCE_PRESSURE
from .geometry import GeometryBuffers, build_geometry_buffers
from .native import HipBemSession


@dataclass
class FieldProblem:
    """Everything but the vertex positions and the Neumann data: topology,
    wavenumbers (float32 as the sweep forms them), observation points and
    solver settings."""

    triangles_3xm: np.ndarray
    physical_tags: np.ndarray
    p1_local2global: np.ndarray
    p1_dof_count: int
    k_real: np.ndarray
    observation_points: np.ndarray
    symmetry_plane: str | None = None
    k_imag: np.ndarray | None = None
    field_k_real: np.ndarray | None = None
    robin_beta: np.ndarray | None = None  # (F, T) complex admittance
    source_tags: tuple[int, ...] = ()
    impedance_source_tag: int = 0
    solve_mode: str = "fp32_ir"
    refine_max_iterations: int = 8
    refine_tolerance: float = 1e-7
    adjoint_refine_iterations: int = 3
    check_open_edges: bool = True
    last_diagnostics: list[dict[str, Any]] = dc_field(default_factory=list)

    @classmethod
    def from_buffers(cls, buffers: GeometryBuffers, k_real, observation_points, **kw) -> "FieldProblem":
        return cls(
            triangles_3xm=np.asarray(buffers.triangles_3xm_i32, np.int32),
            physical_tags=np.asarray(buffers.physical_tags_i32, np.int32),
            p1_local2global=np.asarray(buffers.p1_local2global_i32, np.int32).reshape(-1, 3),
            p1_dof_count=int(buffers.p1_dof_count),
            k_real=np.asarray(k_real, np.float32).reshape(-1),
            observation_points=np.asarray(observation_points, np.float32),
            **kw,
        )

    def buffers(self, vertices_3xn: np.ndarray) -> GeometryBuffers:
        grid = SimpleNamespace(vertices=np.asarray(vertices_3xn, np.float64),
                               elements=self.triangles_3xm,
                               number_of_elements=int(self.triangles_3xm.shape[1]))
        p1 = SimpleNamespace(local2global=self.p1_local2global, global_dof_count=self.p1_dof_count)
        return build_geometry_buffers(grid, self.physical_tags, p1)


def _f32(x, device) -> torch.Tensor:
    return torch.as_tensor(np.ascontiguousarray(x, np.float32), device=device)


class _ExteriorField(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vertices: torch.Tensor, q: torch.Tensor, problem: FieldProblem) -> torch.Tensor:
        if not q.is_cuda:
            raise RuntimeError("exterior_field needs the Neumann data on a HIP device")
        device = q.device
        F = int(problem.k_real.size)
        T = int(problem.triangles_3xm.shape[1])
        if tuple(q.shape) != (F, T):
            raise ValueError(f"q must have shape {(F, T)}, got {tuple(q.shape)}")
        buffers = problem.buffers(vertices.detach().to("cpu", torch.float64).numpy())
        session = HipBemSession.create_session(buffers, symmetry_plane=problem.symmetry_plane,
                                               check_open_edges=problem.check_open_edges)
        obs = _f32(problem.observation_points, device).contiguous()
        n_obs = int(obs.shape[1])
        k = _f32(problem.k_real, device)
        fk = k if problem.field_k_real is None else _f32(problem.field_k_real, device)
        kim = None if problem.k_imag is None else _f32(problem.k_imag, device)
        qd = torch.view_as_real(q.detach().to(torch.complex64).contiguous()).contiguous()
        beta = None
        if problem.robin_beta is not None:
            beta = torch.view_as_real(torch.as_tensor(
                np.ascontiguousarray(problem.robin_beta, np.complex64), device=device)).contiguous()
        out = torch.zeros((F, n_obs, 2), dtype=torch.float32, device=device)
        tag_avg = torch.zeros((F, 1 + len(problem.source_tags), 2), dtype=torch.float64, device=device)
        try:
            problem.last_diagnostics = session.adjoint_forward_device(
                F, k.data_ptr(), fk.data_ptr(), qd.data_ptr(), n_obs, obs.data_ptr(),
                out.data_ptr(), tag_avg.data_ptr(),
                k_imag_ptr=None if kim is None else kim.data_ptr(),
                robin_beta_ptr=None if beta is None else beta.data_ptr(),
                source_tags=tuple(problem.source_tags),
                impedance_source_tag=int(problem.impedance_source_tag),
                solve_mode=problem.solve_mode,
                refine_max_iterations=problem.refine_max_iterations,
                refine_tolerance=problem.refine_tolerance,
            )
        except Exception:
            session.close()
            raise
        ctx.session = session
        ctx.keep = (k, fk, kim, qd, beta, obs)
        ctx.shape = (F, T, int(buffers.n_vertices))
        ctx.vdtype = vertices.dtype
        ctx.vdevice = vertices.device
        ctx.refine = int(problem.adjoint_refine_iterations)
        return torch.view_as_complex(out)

    @staticmethod
    def backward(ctx, grad_field: torch.Tensor):
        F, T, nv = ctx.shape
        session = ctx.session
        try:
            g = torch.view_as_real(grad_field.detach().to(torch.complex64).contiguous()).contiguous()
            gv = torch.empty((3, nv), dtype=torch.float64, device=g.device)
            gq = None
            if ctx.needs_input_grad[1]:
                gq = torch.empty((F, T, 2), dtype=torch.float32, device=g.device)
            session.adjoint_backward_device(g.data_ptr(), gv.data_ptr(),
                                            None if gq is None else gq.data_ptr(), ctx.refine)
        finally:
            session.close()
            ctx.session = None
        grad_v = gv.to(device=ctx.vdevice, dtype=ctx.vdtype) if ctx.needs_input_grad[0] else None
        grad_q = torch.view_as_complex(gq) if gq is not None else None
        return grad_v, grad_q, None


def exterior_field(vertices: torch.Tensor, q: torch.Tensor, problem: FieldProblem) -> torch.Tensor:
    """Field [F, n_obs] (complex64) at the problem's observation points for
    vertex positions ``vertices`` (3, n_vertices) and Neumann data ``q``
    (F, n_triangles); differentiable in both."""
    return _ExteriorField.apply(vertices, q, problem)


def spl_db(field: torch.Tensor) -> torch.Tensor:
    """20 log10(|p| / p_ref) with the -120 dB floor of the sweep."""
    floor = REFERENCE_PRESSURE * 10.0 ** (-120.0 / 20.0)
    mag = torch.sqrt(field.real ** 2 + field.imag ** 2)
    return 20.0 * torch.log10(torch.clamp(mag, min=floor) / REFERENCE_PRESSURE)


def directivity_ripple_loss(field: torch.Tensor, n_planes: int, n_angles: int,
                            on_axis_index: int = 0, ripple_weight: float = 1.0) -> torch.Tensor:
    """Mean squared deviation of the normalised directivity (dB, each plane
    minus its on-axis value) from a flat 0 dB target, plus ``ripple_weight``
    times the variance of the on-axis SPL over frequency.  ``field`` is
    [F, n_planes * n_angles], plane-major."""
    spl = spl_db(field).reshape(field.shape[0], n_planes, n_angles)
    directivity = spl - spl[:, :, on_axis_index:on_axis_index + 1]
    axis = spl[:, 0, on_axis_index]
    return torch.mean(directivity ** 2) + ripple_weight * torch.mean((axis - axis.mean()) ** 2)
