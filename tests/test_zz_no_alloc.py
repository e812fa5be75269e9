"""hbem_dense_solve_noalloc (device arrays, working storage in the library's
fixed device array) against hbem_dense_solve on the same systems, through
the C ABI: two calls in a row, the second smaller, then a conjugate-
transpose call; and the refusal of sizes beyond the array."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _systems(rng, count, n):
    A = ((rng.standard_normal((count, n, n)) + 1j * rng.standard_normal((count, n, n)))
         / np.sqrt(n)).astype(np.complex64)
    A += (1.5 * np.eye(n, dtype=np.complex64))[None]
    b = (rng.standard_normal((count, n)) + 1j * rng.standard_normal((count, n))).astype(np.complex64)
    return A, b


def _noalloc(A, b, trans=0):
    import torch

    from hornlab_metal_bem.native import SOLVE_MODES, _SolveOptions, load_library

    lib = load_library()
    dev = torch.device("cuda:0")
    count, n = b.shape

    def put(a):
        return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)

    Are, Aim, bre, bim = put(A.real), put(A.imag), put(b.real), put(b.imag)
    xre = torch.full((count, n), np.nan, dtype=torch.float32, device=dev)
    xim = torch.full((count, n), np.nan, dtype=torch.float32, device=dev)
    info = torch.full((count,), -7, dtype=torch.int32, device=dev)
    opts = _SolveOptions()
    opts.solve_mode = SOLVE_MODES["fp32"]
    opts.solve_trans = trans
    torch.cuda.synchronize()
    rc = lib.hbem_dense_solve_noalloc(count, n, Are.data_ptr(), Aim.data_ptr(), bre.data_ptr(),
                                      bim.data_ptr(), C.byref(opts), xre.data_ptr(),
                                      xim.data_ptr(), info.data_ptr(), None)
    assert rc == 0, lib.hbem_last_error()
    torch.cuda.synchronize()
    x = xre.cpu().numpy() + 1j * xim.cpu().numpy()
    return x, info.cpu().numpy()


def _check_against_dense_solve(A, b, trans):
    from hornlab_metal_bem.native import dense_solve

    x_na, info = _noalloc(A, b, trans=2 if trans == "C" else 0)
    x_ref, _ = dense_solve(A, b, estimate_rcond=False, trans=trans)
    assert np.all(info == 0), info
    # the same kernels on the same inputs: equal to FP32 rounding
    scale = np.abs(x_ref).max()
    assert np.abs(x_na - x_ref).max() <= 1e-6 * scale, np.abs(x_na - x_ref).max() / scale
    for s in range(b.shape[0]):
        A64 = A[s].astype(np.complex128)
        M = A64 if trans == "N" else A64.conj().T
        ref = np.linalg.solve(M, b[s].astype(np.complex128))
        assert np.linalg.norm(x_na[s] - ref) / np.linalg.norm(ref) < 2e-5


def test_noalloc_matches_dense_solve_twice_then_smaller(require_gpu):
    rng = np.random.default_rng(41)
    print("[noalloc] 3 systems n=300", flush=True)
    A, b = _systems(rng, 3, 300)
    _check_against_dense_solve(A, b, "N")
    print("[noalloc] 2 systems n=130", flush=True)
    A2, b2 = _systems(rng, 2, 130)
    _check_against_dense_solve(A2, b2, "N")
    print("[noalloc] 1 system n=1100 (tournament panels), A^H", flush=True)
    A3, b3 = _systems(rng, 1, 1100)
    _check_against_dense_solve(A3, b3, "C")


def test_noalloc_refuses_sizes_beyond_its_array(require_gpu):
    from hornlab_metal_bem.native import load_library

    lib = load_library()
    fake = C.c_void_p(16)  # never dereferenced: the size check comes first
    for count, n in ((65, 64), (1, 2049), (2, 2048), (5, 1024)):
        rc = lib.hbem_dense_solve_noalloc(count, n, fake, fake, fake, fake, None, fake, fake,
                                          None, None)
        assert rc == -1, (count, n, rc)
        assert b"scratch limits" in lib.hbem_last_error()
