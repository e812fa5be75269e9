"""Shape gradients through the HIP adjoint against FP64 central finite
differences of the oracle (assembly + zgesv + field), the check the
adjoint shape-gradient workload asks for, at a size the oracle re-solves in
seconds.  Vertices on the symmetry planes are left out: moving them off a
plane changes which image pairs touch."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle as O
from meshes import horn_buffers, rel_l2

pytestmark = pytest.mark.gpu

C0 = 343.0
RHO = 1.2041


def _obs(n_angles=7):
    ang = np.deg2rad(np.linspace(0, 90, n_angles))
    z = 2 * np.cos(ang)
    return np.concatenate([np.stack([2 * np.sin(ang), 0 * ang, z]),
                           np.stack([0 * ang, 2 * np.sin(ang), z])], axis=1).astype(np.float32)


def _loss_np(field, n_planes, n_angles):
    floor = 20e-6 * 10 ** (-6)
    spl = 20 * np.log10(np.maximum(np.abs(field), floor) / 20e-6).reshape(field.shape[0], n_planes, n_angles)
    d = spl - spl[:, :, :1]
    axis = spl[:, 0, 0]
    return float(np.mean(d ** 2) + np.mean((axis - axis.mean()) ** 2))


def _oracle_field(g, k, kim, q, beta, obs, l2g):
    out = []
    for i in range(len(k)):
        bt = None if beta is None else beta[i]
        A, rhs, _ = O.assemble(g, float(k[i]), q[i], kim=float(kim[i]), beta=bt)
        p = np.linalg.solve(A, rhs)
        qt = q[i].copy()
        if bt is not None:
            qt = qt + (1j * k[i] - kim[i]) * bt * p[l2g].mean(axis=1)
        out.append(O.field(g, float(k[i]), p, qt, obs))
    return np.stack(out)


def _case(robin):
    from hornlab_metal_bem.adjoint import FieldProblem

    b = horn_buffers(edge_length=0.04, quarter=True)
    freqs = np.array([900.0, 2600.0])
    k = (2 * np.pi * freqs / C0).astype(np.float32)
    kim = np.zeros_like(k) if not robin else np.array([0.05, 0.1], np.float32)
    tags = np.asarray(b.physical_tags_i32)
    q = np.stack([np.where(tags == 2, 1j * RHO * 2 * np.pi * f, 0.0) for f in freqs]).astype(np.complex64)
    beta = None
    if robin:
        beta = np.stack([np.where(tags == 1, 0.3 + 0.2j, 0.0)] * freqs.size).astype(np.complex64)
    obs = _obs()
    prob = FieldProblem.from_buffers(
        b, k, obs, symmetry_plane="yz+xz", k_imag=kim if robin else None, robin_beta=beta,
        solve_mode="fp32_ir", refine_tolerance=1e-9, adjoint_refine_iterations=4)
    return b, k, kim, q, beta, obs, prob


@pytest.mark.parametrize("robin", [False, True])
def test_adjoint_shape_gradient_matches_fp64_finite_differences(require_gpu, robin):
    """Directivity (dB, 2 planes x 7 angles) plus on-axis ripple loss over two
    frequencies; gradient at 8 seeded off-plane vertices (24 components) and
    at three source triangles' Neumann data.  Robin walls with complex k in
    the second case (lambda then carries R^H F^H g)."""
    import torch
    from hornlab_metal_bem.adjoint import directivity_ripple_loss, exterior_field

    b, k, kim, q, beta, obs, prob = _case(robin)
    V = np.asarray(b.vertices_3xn_f32, np.float64)
    dev = torch.device("cuda:0")
    vt = torch.tensor(V, dtype=torch.float64, device=dev, requires_grad=True)
    qt = torch.tensor(q, device=dev, requires_grad=True)
    f = exterior_field(vt, qt, prob)
    loss = directivity_ripple_loss(f, 2, 7)
    loss.backward()
    gv = vt.grad.cpu().numpy()
    gq = qt.grad.cpu().numpy()

    g0 = O.Geometry.from_buffers(b, "yz+xz")
    l2g = np.asarray(b.p1_local2global_i32).reshape(-1, 3)
    qd = q.astype(np.complex128)
    bd = None if beta is None else beta.astype(np.complex128)
    f0 = _oracle_field(g0, k, kim, qd, bd, obs, l2g)
    assert rel_l2(f.detach().cpu().numpy(), f0) < 1e-4
    L = lambda geom, qq: _loss_np(_oracle_field(geom, k, kim, qq, bd, obs, l2g), 2, 7)
    rng = np.random.default_rng(5)
    cand = np.flatnonzero((np.abs(V[0]) > 1e-3) & (np.abs(V[1]) > 1e-3))
    pick = rng.choice(cand, 8, replace=False)
    h = 1e-6
    fd = np.zeros((3, pick.size))
    for j, v in enumerate(pick):
        for d in range(3):
            Vp, Vm = V.copy(), V.copy()
            Vp[d, v] += h
            Vm[d, v] -= h
            fd[d, j] = (L(O.Geometry.from_vertices(Vp, g0), qd) - L(O.Geometry.from_vertices(Vm, g0), qd)) / (2 * h)
    err = rel_l2(gv[:, pick], fd)
    print("vertex gradient rel-L2 vs FD:", err)
    assert err < 1e-4, (err, gv[:, pick].T, fd.T)
    src = np.flatnonzero(np.asarray(b.physical_tags_i32) == 2)[:3]
    for t in src:
        e = np.zeros_like(qd)
        e[:, t] = 1e-3 * np.abs(qd[:, t]).max()
        for i in range(len(k)):
            ei = np.zeros_like(qd)
            ei[i, t] = e[i, t]
            dr = (L(g0, qd + ei) - L(g0, qd - ei)) / (2 * ei[i, t].real)
            di = (L(g0, qd + 1j * ei) - L(g0, qd - 1j * ei)) / (2 * ei[i, t].real)
            ref = complex(dr, di)
            assert abs(gq[i, t] - ref) <= 1e-4 * abs(ref) + 1e-9, (i, t, gq[i, t], ref)


def test_adjoint_backward_needs_a_resident_forward(require_gpu):
    """The backward entry point refuses to run without its forward and after
    another call has reused the session's workspaces."""
    import ctypes as C

    import torch
    from hornlab_metal_bem.native import HipBemSession

    b = horn_buffers(edge_length=0.05, quarter=True)
    with HipBemSession.create_session(b, symmetry_plane="yz+xz") as s:
        g = torch.zeros((1, 4, 2), dtype=torch.float32, device="cuda:0")
        gv = torch.zeros((3, b.n_vertices), dtype=torch.float64, device="cuda:0")
        with pytest.raises(RuntimeError, match="no adjoint forward"):
            s.adjoint_backward_device(g.data_ptr(), gv.data_ptr())
