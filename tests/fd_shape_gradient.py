"""FP64 central finite differences of the directivity/ripple loss with
respect to vertex positions, for checking the HIP adjoint at full size.

The oracle assembles each frequency's system once (FP64 corrected
assembly), factors it (LAPACK zgetrf) and solves it.  A perturbed vertex only
moves the pairs that involve its triangles, so each perturbed system is the
base system plus (subset assembly at the moved vertex) - (subset assembly at
the base vertex), solved by refinement on the base factors; the field is
re-evaluated in full.  Run as a script (CPU only; about an hour on 8 cores
for the 3000-dof, 24-frequency workload):

    python tests/fd_shape_gradient.py --fd fd_c5.npz [gpu_gradient.npz]

computes (or loads) the differences at the 50 seeded vertices x 3 (1e-6 m
steps) and, given the gradient that bench.py --workload c5
--save-gradient wrote, prints its relative L2 and worst component error.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np
import scipy.linalg as sla

ROOT = Path(__file__).resolve().parents[1]
sys.path[:0] = [str(ROOT), str(ROOT / "oracle"), str(ROOT / "hornlab-metal-bem_amd")]

import oracle as O  # noqa: E402


def loss_np(field: np.ndarray, n_planes: int, n_angles: int) -> float:
    """Same loss as hornlab_metal_bem.adjoint.directivity_ripple_loss."""
    floor = 20e-6 * 10 ** (-6)
    spl = 20 * np.log10(np.maximum(np.abs(field), floor) / 20e-6).reshape(field.shape[0], n_planes, n_angles)
    d = spl - spl[:, :, :1]
    axis = spl[:, 0, 0]
    return float(np.mean(d ** 2) + np.mean((axis - axis.mean()) ** 2))


class LocalFD:
    def __init__(self, buffers, sym, k, q, obs, n_planes, n_angles):
        like = O.Geometry.from_buffers(buffers, sym)
        self.V = np.asarray(buffers.vertices_3xn_f32, np.float64)
        self.g0 = O.Geometry.from_vertices(self.V, like)
        self.k, self.q, self.obs = np.asarray(k, np.float64), np.asarray(q, np.complex128), obs
        self.shape = (n_planes, n_angles)
        self.A, self.rhs, self.lu, self.p = [], [], [], []
        for i, kk in enumerate(self.k):
            A, r, _ = O.assemble(self.g0, float(kk), self.q[i])
            lu = sla.lu_factor(A, check_finite=False)
            x = sla.lu_solve(lu, r, check_finite=False)
            for _ in range(2):
                x = x + sla.lu_solve(lu, r - A @ x, check_finite=False)
            self.A.append(A)
            self.rhs.append(r)
            self.lu.append(lu)
            self.p.append(x)
        self.base_field = np.stack([O.field(self.g0, float(kk), self.p[i], self.q[i], obs)
                                    for i, kk in enumerate(self.k)])

    def loss_at(self, v: int, V: np.ndarray, base_parts) -> float:
        g = O.Geometry.from_vertices(V, self.g0)
        sel = np.any(self.g0.triangles == v, axis=0)
        fields = []
        for i, kk in enumerate(self.k):
            As, rs = O.assemble_subset(g, sel, float(kk), self.q[i])
            dA = As - base_parts[i][0]
            b = self.rhs[i] + (rs - base_parts[i][1])
            x = self.p[i].copy()
            for _ in range(3):
                x = x + sla.lu_solve(self.lu[i], b - self.A[i] @ x - dA @ x, check_finite=False)
            fields.append(O.field(g, float(kk), x, self.q[i], self.obs))
        return loss_np(np.stack(fields), *self.shape)

    def gradient(self, vertices, h=1e-6, log=None):
        out = np.zeros((3, len(vertices)))
        for j, v in enumerate(vertices):
            sel = np.any(self.g0.triangles == v, axis=0)
            base = [O.assemble_subset(self.g0, sel, float(kk), self.q[i]) for i, kk in enumerate(self.k)]
            for d in range(3):
                Vp, Vm = self.V.copy(), self.V.copy()
                Vp[d, v] += h
                Vm[d, v] -= h
                out[d, j] = (self.loss_at(v, Vp, base) - self.loss_at(v, Vm, base)) / (2 * h)
            if log:
                log(j, v, out[:, j])
        return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("gradient", nargs="?", default=None,
                    help=".npz written by bench.py --workload c5 --save-gradient")
    ap.add_argument("--fd", default=None,
                    help="finite differences (.npz): loaded if present, else computed and saved")
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    import bench

    O.set_threads(args.threads)
    freqs, k, q, pts, buffers, pick = bench.c5_problem()
    t0 = time.perf_counter()
    if args.fd and os.path.exists(args.fd):
        saved = np.load(args.fd)
        ref, base_loss = saved["fd"], float(saved["loss"])
    else:
        fd = LocalFD(buffers, "yz+xz", k.astype(np.float32).astype(np.float64),
                     q.astype(np.complex64), pts, 3, pts.shape[1] // 3)
        base_loss = loss_np(fd.base_field, 3, pts.shape[1] // 3)
        print(f"base systems: {time.perf_counter() - t0:.1f} s, loss {base_loss}", flush=True)
        ref = fd.gradient(pick, log=lambda j, v, g: print(j, v, g, flush=True))
        if args.fd:
            np.savez(args.fd, vertices=pick, fd=ref, loss=base_loss)
    result = {"vertices": int(len(pick)), "components": int(ref.size), "step_m": 1e-6,
              "loss_fp64": base_loss, "seconds": round(time.perf_counter() - t0, 1)}
    if args.gradient:
        data = np.load(args.gradient)
        assert np.array_equal(pick, data["vertices"]), "vertex sample differs from the saved gradient"
        ad = data["grad"]
        result.update({
            "loss_gpu": float(data["loss"]),
            "rel_l2": float(np.linalg.norm(ad - ref) / np.linalg.norm(ref)),
            "max_abs_err_over_max_abs": float(np.max(np.abs(ad - ref)) / np.max(np.abs(ref))),
        })
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
