"""GPU tests of the boundary contract beyond the plain solve: conjugate-
transpose solves and the device rcond estimate on both solve paths (one
block per vector up to order 2048, stepped above), the FP64-accurate dense
solve (dense_solve_dtype="float64") and the refinement stall flag, and the
assembly modes ("optimized" and "reference" against the oracle's regular
assembly).  All calls go through the C ABI."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from meshes import horn_buffers, rel_l2, tiny_two_triangles

pytestmark = pytest.mark.gpu

C0 = 343.0


def _cond_matrix(rng, n, cond):
    """Complex n x n with singular values geomspace(1, 1/cond)."""
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    W, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return (Q * np.geomspace(1.0, 1.0 / cond, n)) @ W.conj().T


def _exact_rcond(A64):
    return 1.0 / (np.linalg.norm(A64, 1) * np.linalg.norm(np.linalg.inv(A64), 1))


@pytest.mark.parametrize("n,count", [(300, 2), (2240, 2), (4000, 1)])
def test_transpose_solve_and_rcond_match_numpy(require_gpu, n, count):
    """A x = b and A^H x = b against numpy on the one-block solve (n = 300)
    and the stepped solve (n = 2240: tournament panels; n = 4000: sixteen
    outer blocks), with the device clacn2 rcond estimate a lower bound of
    1/(|A|_1 |A^-1|_1) within 3x."""
    from hornlab_metal_bem.native import dense_solve

    print(f"[contract] transpose/rcond n={n} count={count}", flush=True)
    rng = np.random.default_rng(100 + n)
    A = ((rng.standard_normal((count, n, n)) + 1j * rng.standard_normal((count, n, n)))
         / np.sqrt(n)).astype(np.complex64)
    A += (1.5 * np.eye(n, dtype=np.complex64))[None]
    b = (rng.standard_normal((count, n)) + 1j * rng.standard_normal((count, n))).astype(np.complex64)
    xn, info_n = dense_solve(A, b, estimate_rcond=True)
    xc, info_c = dense_solve(A, b, estimate_rcond=False, trans="C")
    for s in range(count):
        A64 = A[s].astype(np.complex128)
        b64 = b[s].astype(np.complex128)
        ref_n = np.linalg.solve(A64, b64)
        ref_c = np.linalg.solve(A64.conj().T, b64)
        assert rel_l2(xn[s], ref_n) < 2e-5, (n, s, rel_l2(xn[s], ref_n))
        assert rel_l2(xc[s], ref_c) < 2e-5, (n, s, rel_l2(xc[s], ref_c))
        rc = _exact_rcond(A64)
        est = info_n[s]["dense_solve_rcond"]
        assert rc <= est * 1.01 and est < 3.0 * rc, (n, s, rc, est)
        assert info_n[s]["dense_solve_rcond_seconds"] > 0.0


def test_rcond_flags_near_singular_system(require_gpu):
    """cond = 1e7: the estimate lands within 3x of the exact value, far
    below the 1e-6 suspect threshold of the sweep policy."""
    from hornlab_metal_bem.native import dense_solve

    rng = np.random.default_rng(8)
    n = 500
    A = _cond_matrix(rng, n, 1e7).astype(np.complex64)[None]
    b = (rng.standard_normal((1, n)) + 1j * rng.standard_normal((1, n))).astype(np.complex64)
    _, info = dense_solve(A, b, estimate_rcond=True)
    rc = _exact_rcond(A[0].astype(np.complex128))
    est = info[0]["dense_solve_rcond"]
    # FP32 solves inside the estimator: a lower bound up to their rounding
    assert 0.5 * rc <= est < 3.0 * rc, (rc, est)
    assert est < 1e-6


def test_fp64_solve_reaches_zgesv_accuracy(require_gpu):
    """cond = 1e6: the FP32 LU alone leaves ~1e-2 error; the float64 mode
    (FP32 LU + GMRES-IR on the FP64 residual) matches zgesv of the same FP32
    matrix to the FP32 rounding of the returned vector and reports no stall,
    while classical refinement at cond = 1e9 stalls and says so."""
    from hornlab_metal_bem.native import dense_solve

    rng = np.random.default_rng(12)
    n = 400
    A = _cond_matrix(rng, n, 1e6).astype(np.complex64)[None]
    b = (rng.standard_normal((1, n)) + 1j * rng.standard_normal((1, n))).astype(np.complex64)
    ref = O.solve(A[0].astype(np.complex128), b[0].astype(np.complex128))
    x32, _ = dense_solve(A, b, solve_mode="fp32")
    x, info = dense_solve(A, b, solve_mode="fp64", refine_max_iterations=12)
    assert rel_l2(x32[0], ref) > 1e-5
    # complex64 output: its rounding (eps32 / 2 ~ 6e-8 per entry) is the floor
    assert rel_l2(x[0], ref) < 1e-7, (rel_l2(x[0], ref), info[0])
    assert info[0]["dense_solve_refine_residual_rel"] < 1e-10, info[0]
    assert info[0]["dense_solve_refine_stalled"] is False

    A9 = _cond_matrix(rng, n, 1e9).astype(np.complex64)[None]
    _, info9 = dense_solve(A9, b, solve_mode="fp32_ir", refine_max_iterations=8, refine_tolerance=1e-12)
    assert info9[0]["dense_solve_refine_stalled"] is True, info9[0]


def test_dense_solve_rejects_options_it_cannot_honour(require_gpu):
    from hornlab_metal_bem.native import SOLVE_MODES, _SolveOptions, _check, _f32p, load_library

    lib = load_library()
    n = 8
    A = np.eye(n, dtype=np.float32)
    z = np.zeros(n, np.float32)
    x = np.zeros(n, np.float32)
    for field, value in (("solve_mode", 9), ("n_sources", 2), ("solve_trans", 1)):
        opts = _SolveOptions()
        opts.solve_mode = SOLVE_MODES["fp32"]
        setattr(opts, field, value)
        rc = lib.hbem_dense_solve(1, n, _f32p(A), _f32p(A), _f32p(z), _f32p(z), opts, _f32p(x),
                                  _f32p(x), None)
        assert rc == -1, (field, rc)
        with pytest.raises(RuntimeError):
            _check(rc, "dense_solve")


@pytest.mark.parametrize("mode", ["optimized", "reference"])
def test_regular_assembly_modes_match_oracle(require_gpu, mode):
    """metal_native_assembly_mode "optimized" (and "reference" without
    complex k or Robin walls): the 6x6 rule on every pair, touching pairs
    included, against the oracle's regular assembly (no Sauter-Schwab
    corrections); the tiny fixture and the yz+xz quarter horn with its
    image pairs."""
    from hornlab_metal_bem.native import HipBemSession

    for buffers, plane, f in ((tiny_two_triangles(), None, 100.0),
                              (horn_buffers(edge_length=0.03, quarter=True), "yz+xz", 3000.0)):
        k = float(np.float32(2 * np.pi * f / C0))
        T = buffers.n_triangles
        q = (np.linspace(0.5, 1.5, T) + 0.25j).astype(np.complex64)
        g = O.Geometry.from_buffers(buffers, plane)
        A_ref, rhs_ref, _ = O.assemble(g, k, q, duffy=False)
        A_cor, _, _ = O.assemble(g, k, q)
        with HipBemSession.create_session(buffers, symmetry_plane=plane) as s:
            s.set_assembly_mode(mode)
            A, rhs = s.assemble_standard_neumann(f, k, q)
            s.set_assembly_mode("corrected")
            A2, _ = s.assemble_standard_neumann(f, k, q)
        assert rel_l2(A, A_ref) < 1e-5, (mode, plane, rel_l2(A, A_ref))
        assert rel_l2(rhs, rhs_ref) < 1e-5, (mode, plane, rel_l2(rhs, rhs_ref))
        # the modes differ where the touching pairs are (the Duffy corrections)
        assert rel_l2(A_cor, A_ref) > 1e-3
        assert rel_l2(A2, A_cor) < 1e-5


def test_reference_mode_keeps_corrections_for_complex_k(require_gpu):
    """"reference" mode with complex k runs the corrected quadrature, as the
    reference's CPU path adds its Duffy corrections for complex-k and Robin
    cases (main.swift:8150-8170)."""
    from hornlab_metal_bem.native import HipBemSession

    buffers = horn_buffers(edge_length=0.03, quarter=True)
    f = 2000.0
    k = float(np.float32(2 * np.pi * f / C0))
    kim = float(np.float32(0.005 * k))
    q = np.full(buffers.n_triangles, 1.0 + 0.5j, np.complex64)
    g = O.Geometry.from_buffers(buffers, "yz+xz")
    A_ref, _, _ = O.assemble(g, k, q, kim=kim)
    with HipBemSession.create_session(buffers, symmetry_plane="yz+xz") as s:
        s.set_assembly_mode("reference")
        A, _ = s.assemble_standard_neumann(f, k, q, k_imag=kim)
    assert rel_l2(A, A_ref) < 1e-5, rel_l2(A, A_ref)


def test_optimized_mode_in_batch_diagnostics(require_gpu):
    """The batch path honours the mode and reports it; the solution matches
    the oracle's regular-only system solved in FP64."""
    from hornlab_metal_bem.native import HipBemSession

    buffers = horn_buffers(edge_length=0.03, quarter=True)
    freqs = np.array([1500.0])
    k = (2 * np.pi * freqs / C0).astype(np.float32)
    T = buffers.n_triangles
    tags = np.asarray(buffers.physical_tags_i32)
    q = np.where(tags == 2, 1j * 1.21 * 2 * np.pi * freqs[0], 0).astype(np.complex64)[None]
    obs = np.array([[0.0], [0.0], [2.0]], np.float32)
    with HipBemSession.create_session(buffers, symmetry_plane="yz+xz",
                                      extra_env={"HORNLAB_METAL_BEM_NATIVE_ASSEMBLY_MODE": "optimized"}) as s:
        res = s.assemble_solve_evaluate_standard_neumann_batch(freqs, k, q, obs, source_tags=[2])
    d = res[0].diagnostics
    assert d["assembly_mode"] == "optimized" and d["assembly_rule"] == "regular_6x6"
    g = O.Geometry.from_buffers(buffers, "yz+xz")
    A_ref, rhs_ref, _ = O.assemble(g, float(k[0]), q[0], duffy=False)
    p_ref = O.solve(A_ref, rhs_ref)
    assert rel_l2(res[0].pressure, p_ref) < 1e-4, rel_l2(res[0].pressure, p_ref)
    assert T == q.shape[1]
