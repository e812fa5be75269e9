#!/usr/bin/env python3
"""Frequencies solved per second of the per-frequency BEM hot path.

Workload (one step): the 40-frequency sweep of a yz+xz quarter exponential
horn (throat 12.7 mm, mouth 150 mm, length 150 mm, 20 mm rear shell, ~1000
P1 dofs, +-5% seeded jitter), 500 Hz - 20 kHz log-spaced, complex64, field on
3 planes x 37 angles at 2 m: corrected assembly (regular + Sauter-Schwab),
batched complex LU with FP16 MFMA trailing updates and GMRES-IR against the
FP64 residual to 1e-6, the rcond estimate, exterior field and surface
reductions, all on one GPU with inputs resident in HBM.  The plain FP32 path
(one frequency per launch sequence) and the batched FP32 path are timed
beside it (the line's "modes").

Prints one JSON line.  The CPU baseline times the oracle's FP32 port of the
reference CPU path (regular + Duffy + level-1 near-field assembly, LAPACK
cgesv, field) on a sample of the same sweep, and the accuracy fields compare
sampled frequencies against the FP64 oracle (FP64 assembly + zgesv).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "hornlab-metal-bem_amd"))

PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector, published
FLOPS_PER_EVAL = 80  # G + dG/dn evaluation incl. block accumulation (fixed count)
TRANS_PER_EVAL = 4  # sqrt/rcp, sin, cos, (exp) per evaluation (fixed count)
REFERENCE_FREQ_PER_S = 20.0  # reference README: 40 freqs in ~2 s on Apple M-series
# HBM bytes per k_assemble_regular dispatch of one sweep of the default mesh
# of each workload: rocprofv3 FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (factors
# from a 512 MiB copy kernel), profiles/pmc_traffic_c2_regular_assembly.md
REGULAR_TRAFFIC_BYTES: dict[str, float] = {"c2": 3.208e9}


WORKLOADS = {
    # quarter horn, yz+xz, ~1000 dofs, 40 log frequencies 500 Hz - 20 kHz, 3 x 37 points
    "c2": dict(quarter=True, symmetry="yz+xz", dofs=1080, fmin=500.0, fmax=20000.0, nf=40,
               angles=37),
    # full horn, no symmetry, ~8000 dofs, 64 log frequencies 200 Hz - 16 kHz, 3 x 181 points
    "c3": dict(quarter=False, symmetry=None, dofs=8000, fmin=200.0, fmax=16000.0, nf=64,
               angles=181),
    # adjoint shape gradient: quarter exponential horn ~3000 dofs, 24 log frequencies 1-12 kHz
    "c5": dict(quarter=True, symmetry="yz+xz", dofs=3000, fmin=1000.0, fmax=12000.0, nf=24,
               angles=37),
    # 16 seeded quarter-horn variants, n ~ U(600, 4000), 32 log frequencies 300 Hz - 18 kHz
    "c4": dict(quarter=True, symmetry="yz+xz", dofs=None, fmin=300.0, fmax=18000.0, nf=32,
               angles=37, variants=16),
}


def c4_variants(seed: int = 4):
    """The batched uneven sweep: 16 horn variants (throat 10-25 mm, mouth
    80-200 mm, length 80-250 mm, conical / exponential / tractrix-like
    flares), dof targets uniform in 600-4000, 32 log frequencies each."""
    from hornlab_metal_bem import ObservationConfig, native_config
    from hornlab_metal_bem.mesh import loaded_mesh_from_arrays
    from hornlab_metal_bem.procedural import HornSpec, edge_length_for_dofs, horn_body

    w = WORKLOADS["c4"]
    rng = np.random.default_rng(seed)
    flares = ("conical", "exponential", "tractrix")
    out = []
    for i in range(w["variants"]):
        spec = HornSpec(float(rng.uniform(0.010, 0.025)), float(rng.uniform(0.080, 0.200)),
                        float(rng.uniform(0.080, 0.250)), 0.020, flares[i % 3])
        target = int(rng.integers(600, 4001))
        # n ~ h^-2: one coarse probe brackets the bisection
        V0, _, _ = horn_body(spec, edge_length=0.03, quarter=True, jitter=0.0, seed=i)
        h0 = 0.03 * math.sqrt(V0.shape[0] / target)
        h = edge_length_for_dofs(spec, target, quarter=True, seed=i, rtol=0.02,
                                 bracket=(h0 / 1.5, h0 * 1.5))
        V, F, T = horn_body(spec, edge_length=h, quarter=True, jitter=0.05, seed=i)
        mesh = loaded_mesh_from_arrays(V, F, T, native_symmetry_plane="yz+xz")
        config = native_config(
            freq_min_hz=w["fmin"], freq_max_hz=w["fmax"], freq_count=w["nf"], freq_spacing="log",
            native_symmetry_plane="yz+xz", source_motion="axial", velocity_sources={2: 1.0},
            observation=ObservationConfig(planes=["horizontal", "vertical", "diagonal"],
                                          distance_m=2.0, angle_count=w["angles"]),
        )
        out.append((spec, target, mesh, config))
    return out


def c5_problem():
    """Inputs of the adjoint workload: mesh buffers, wavenumbers, Neumann
    rows, observation points (3 planes x 37) and the 50 seeded off-plane
    vertices the finite-difference check samples."""
    spec, h, mesh, config = workload(name="c5")
    freqs, k, q, pts, buffers = prepare_inputs(mesh, config)
    V = np.asarray(buffers.vertices_3xn_f32, np.float64)
    cand = np.flatnonzero((np.abs(V[0]) > 1e-3) & (np.abs(V[1]) > 1e-3))
    pick = np.sort(np.random.default_rng(5).choice(cand, 50, replace=False))
    return freqs, k, q, pts, buffers, pick


def run_c5(args) -> int:
    """Forward field + directivity/ripple loss + hand-written adjoint
    backward (vertex gradients) per step; the session is rebuilt from the
    vertices every step as in a shape-optimisation loop."""
    import torch

    from hornlab_metal_bem.adjoint import FieldProblem, directivity_ripple_loss, exterior_field

    freqs, k, q, pts, buffers, pick = c5_problem()
    F, n, T = freqs.size, buffers.p1_dof_count, buffers.n_triangles
    solve_mode = args.solve_mode or "fp32_ir"
    extra = {}
    if args.refine_tol is not None:
        extra["refine_tolerance"] = args.refine_tol
    if args.adjoint_refine is not None:
        extra["adjoint_refine_iterations"] = args.adjoint_refine
    prob = FieldProblem.from_buffers(buffers, k.astype(np.float32), pts, symmetry_plane="yz+xz",
                                     solve_mode=solve_mode, source_tags=(2,),
                                     impedance_source_tag=2, **extra)
    dev = torch.device("cuda:0")
    V = torch.tensor(np.asarray(buffers.vertices_3xn_f32, np.float64), device=dev, requires_grad=True)
    qd = torch.tensor(q.astype(np.complex64), device=dev)

    def step():
        V.grad = None
        f = exterior_field(V, qd, prob)
        loss = directivity_ripple_loss(f, 3, pts.shape[1] // 3)
        loss.backward()
        return loss

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    print("c5: warm", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    with torch.no_grad():
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        f = exterior_field(V.detach(), qd, prob)
        torch.cuda.synchronize()
        fwd = time.perf_counter() - t1
    grad = V.grad.cpu().numpy()
    if args.save_gradient:
        np.savez(args.save_gradient, vertices=pick, grad=grad[:, pick], loss=float(loss.detach()),
                 field=f.cpu().numpy())
    out = {
        "metric": "frequencies solved/sec (assemble+LU+field) at 1k/8k dofs; rel-L2 vs FP64",
        "value": round(F * args.steps / dt, 3),
        "unit": "frequencies/s (forward + adjoint backward)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": f"adjoint shape gradient: quarter exponential horn yz+xz, n={n} P1 dofs, "
            f"T={T} triangles, {F} log frequencies 1-12 kHz, 3x37 directivity dB + on-axis ripple loss",
            "n_dofs": n, "n_triangles": T, "frequencies": F, "solve_mode": solve_mode,
        },
        "forward_only_ms": round(1e3 * fwd, 3),
        "loss": float(loss.detach()),
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


def run_c4(args) -> int:
    """Batched uneven sweep: the 16 variants' sessions in one session group
    (size-classed batched LU; the assembly either one launch over all meshes
    with the whole sweep replayed from one hipGraph, or one launch per size
    class with each class's LU on a second hardware queue beside the next
    class's assembly), size-class step and assembly split autotuned;
    per-system latency = time from the sweep's start until the system's mesh
    has its results."""
    import torch

    from hornlab_metal_bem.native import GroupCase, HipBemGroup, HipBemSession

    variants = c4_variants()
    print(f"c4: {len(variants)} variants meshed", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    jobs, sessions, cases = [], [], []
    for spec, target, mesh, config in variants:
        freqs, k, q, pts, buffers = prepare_inputs(mesh, config)
        F, n, n_obs = freqs.size, buffers.p1_dof_count, pts.shape[1]
        job = dict(
            n=n, T=buffers.n_triangles, F=F, n_obs=n_obs,
            k=torch.from_numpy(k.astype(np.float32)).to(dev),
            q=torch.from_numpy(np.ascontiguousarray(q.astype(np.complex64)).view(np.float32)).to(dev),
            obs=torch.from_numpy(np.ascontiguousarray(pts, np.float32)).to(dev),
            field=torch.zeros(F * n_obs * 2, dtype=torch.float32, device=dev),
            tag=torch.zeros(F * 2 * 2, dtype=torch.float64, device=dev),
        )
        jobs.append(job)
        sessions.append(HipBemSession.create_session(buffers, symmetry_plane="yz+xz"))
        cases.append(GroupCase(F, job["k"].data_ptr(), job["k"].data_ptr(), job["q"].data_ptr(),
                               n_obs, job["obs"].data_ptr(), job["field"].data_ptr(),
                               job["tag"].data_ptr()))
    torch.cuda.synchronize()
    group = HipBemGroup(sessions)
    tune = group.autotune(cases, source_tags=[2], impedance_source_tag=2)
    print(f"c4 autotune {tune}", file=sys.stderr, flush=True)

    def step():
        return group.solve_device(cases, source_tags=[2], impedance_source_tag=2)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, member_s, _ = step()
        for j, ls in zip(jobs, member_s):
            lat.extend([ls] * j["F"])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    group.close()
    for s in sessions:
        s.close()
    systems = sum(j["F"] for j in jobs)
    lat = np.asarray(lat) * 1e3
    ns = [j["n"] for j in jobs]
    out = {
        "metric": "frequencies solved/sec (assemble+LU+field) at 1k/8k dofs; rel-L2 vs FP64",
        "value": round(systems * args.steps / dt, 3),
        "unit": "frequencies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": f"batched uneven sweep: {len(jobs)} quarter-horn variants, n {min(ns)}-{max(ns)} "
            f"P1 dofs, {WORKLOADS['c4']['nf']} log frequencies each = {systems} systems, one "
            "session group (size-classed batched LU; "
            + ("assembly per size class, each class's LU beside the next class's assembly)"
               if tune["split_assembly"] else "one assembly launch, replayed hipGraph)"),
            "n_dofs": ns,
            "solve_mode": "fp32",
            "autotuned": {"size_class_step": tune["granularity"],
                          "assembly_split_per_class": tune["split_assembly"],
                          "sweep_seconds_per_candidate": tune["seconds"]},
        },
        "latency_ms_per_system": {"p50": round(float(np.percentile(lat, 50)), 3),
                                  "p99": round(float(np.percentile(lat, 99)), 3)},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


def workload(target_dofs: int | None = None, seed: int = 0, name: str = "c2"):
    from hornlab_metal_bem import ObservationConfig, native_config
    from hornlab_metal_bem.mesh import loaded_mesh_from_arrays
    from hornlab_metal_bem.procedural import HornSpec, edge_length_for_dofs, horn_body

    w = WORKLOADS[name]
    dofs = target_dofs or w["dofs"]
    spec = HornSpec(0.0127, 0.150, 0.150, 0.020, "exponential")
    h = edge_length_for_dofs(spec, dofs, quarter=w["quarter"], seed=seed)
    V, F, T = horn_body(spec, edge_length=h, quarter=w["quarter"], jitter=0.05, seed=seed)
    sym = w["symmetry"]
    mesh = loaded_mesh_from_arrays(V, F, T, native_symmetry_plane=sym)
    config = native_config(
        freq_min_hz=w["fmin"], freq_max_hz=w["fmax"], freq_count=w["nf"], freq_spacing="log",
        native_symmetry_plane=sym, source_motion="axial", velocity_sources={2: 1.0},
        observation=ObservationConfig(planes=["horizontal", "vertical", "diagonal"], distance_m=2.0,
                                      angle_count=w["angles"]),
    )
    return spec, h, mesh, config


def prepare_inputs(mesh, config):
    from hornlab_metal_bem import _resolve_frame
    from hornlab_metal_bem.sweep import (
        _build_frequency_grid,
        _build_neumann_rows,
        _field_points_3xn,
        _k_values_for_native,
        _prepare,
    )

    frame = _resolve_frame(mesh, config)
    freqs = _build_frequency_grid(config)
    obs_points, angles, p1, dp0, buffers, face_scale = _prepare(mesh, frame, config)
    k, kim = _k_values_for_native(freqs, config)
    q = _build_neumann_rows(dp0, mesh.physical_tags, freqs, config, face_scale)
    pts = _field_points_3xn(obs_points)
    return freqs, k, q, pts, buffers


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(buffers, sym, freqs, k, q, pts, sample_idx, threads, near_level=1):
    """The oracle's FP32 port of the reference CPU path on host threads:
    regular + Sauter-Schwab (+ level-1 subdivided near-field) assembly,
    scipy cgesv, field.  Returns (frequencies/s, seconds)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    O.set_threads(threads)
    os.environ.setdefault("OPENBLAS_NUM_THREADS", str(threads))
    g = O.Geometry.from_buffers(buffers, sym)
    t0 = time.perf_counter()
    for i in sample_idx:
        A, rhs, _ = O.assemble(g, float(k[i]), q[i], near_level=near_level, near_threshold=1.5,
                               bits=32)
        p = O.solve(A, rhs, bits=32)
        O.field(g, float(k[i]), p, q[i], pts, bits=32)
    dt = time.perf_counter() - t0
    return len(sample_idx) / dt, dt


def fp64_reference(buffers, sym, k, q, pts, idx, threads):
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    O.set_threads(threads)
    g = O.Geometry.from_buffers(buffers, sym)
    A, rhs, _ = O.assemble(g, float(k[idx]), q[idx], bits=64)
    p = O.solve(A, rhs, bits=64)
    return p, O.field(g, float(k[idx]), p, q[idx], pts, bits=64)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--dofs", type=int, default=None)
    ap.add_argument("--solve-mode", default=None,
                    choices=["fp32", "fp32_ir", "mixed_fp16", "mixed_bf16", "fp64"],
                    help="default: fp32 for c2 (the solve() default), mixed_fp16 (GMRES-IR to "
                         "1e-6) for c3")
    ap.add_argument("--refine-tol", type=float, default=None)
    ap.add_argument("--mixed-tol", type=float, default=1e-6,
                    help="c2 mode comparison: GMRES-IR residual target of the mixed modes")
    ap.add_argument("--no-rcond", action="store_true",
                    help="skip the rcond estimate (solve() runs it by default)")
    ap.add_argument("--no-modes", action="store_true",
                    help="c2: skip the per-frequency plain FP32 / graph / mixed-mode lines")
    ap.add_argument("--no-solve-wall", action="store_true",
                    help="skip the end-to-end solve() wall time")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--lu-graph", type=int, default=0, choices=[0, 1],
                    help="replay the LU launch sequence from a captured hipGraph (solve() does not)")
    ap.add_argument("--save-gradient", default=None, help="c5: write the sampled gradient (.npz)")
    ap.add_argument("--adjoint-refine", type=int, default=None,
                    help="c5: refinement steps of the adjoint solve (FieldProblem default if unset)")
    args = ap.parse_args()
    if args.gpus == 1 and args.workload == "c5":
        return run_c5(args)
    if args.gpus == 1 and args.workload == "c4":
        return run_c4(args)
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: frequencies shard independently across GPUs with no "
            "exchange, and this build runs on single-GPU boxes only (use --gpus 1).",
            file=sys.stderr,
        )
        return 2

    import torch

    from hornlab_metal_bem.native import HipBemSession, discover_native_runtime

    rt = discover_native_runtime()
    if not rt.available:
        print("HIP backend unavailable: " + "; ".join(rt.unavailable_reasons), file=sys.stderr)
        return 1
    wl = WORKLOADS[args.workload]
    sym = wl["symmetry"]
    # mixed precision + GMRES-IR to a 1e-6 residual: the mode the sweep is
    # quoted in (surface pressure within 1e-5 of FP64); FP32 lines below
    solve_mode = args.solve_mode or "mixed_fp16"
    refine_tol = args.refine_tol or (1e-7 if solve_mode in ("fp32", "fp32_ir") else args.mixed_tol)
    rcond = not args.no_rcond
    spec, h, mesh, config = workload(args.dofs, name=args.workload)
    freqs, k, q, pts, buffers = prepare_inputs(mesh, config)
    F, T, n, n_obs = freqs.size, buffers.n_triangles, buffers.p1_dof_count, pts.shape[1]
    dev = torch.device("cuda:0")
    k_d = torch.from_numpy(k.astype(np.float32)).to(dev)
    q_d = torch.from_numpy(np.ascontiguousarray(q.astype(np.complex64)).view(np.float32)).to(dev)
    obs_d = torch.from_numpy(np.ascontiguousarray(pts, np.float32)).to(dev)
    field_d = torch.zeros(F * n_obs * 2, dtype=torch.float32, device=dev)
    p_d = torch.zeros(F * n * 2, dtype=torch.float32, device=dev)
    tag_d = torch.zeros(F * 2 * 2, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    print(f"bench: {args.workload} n={n} T={T} F={F} mode={solve_mode} rcond={rcond}",
          file=sys.stderr, flush=True)

    def timed(step, steps, warmup):
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        diags = []
        t0 = time.perf_counter()
        for _ in range(steps):
            diags.append(step())
        torch.cuda.synchronize()
        return time.perf_counter() - t0, diags

    modes = {}
    with HipBemSession.create_session(buffers, symmetry_plane=sym) as s:
        info = s.info()

        def make_step(mode, tol, batch=None, graph=False):
            def step():
                return s.solve_batch_device(
                    F, k_d.data_ptr(), k_d.data_ptr(), q_d.data_ptr(), n_obs, obs_d.data_ptr(),
                    field_d.data_ptr(), tag_d.data_ptr(), p_ptr=p_d.data_ptr(), source_tags=[2],
                    impedance_source_tag=2, solve_mode=mode, estimate_rcond=rcond,
                    refine_tolerance=1e-7 if tol is None else tol, refine_max_iterations=20,
                    batch_frequencies=batch,
                    lu_graph=graph,
                )
            return step

        dt, diags = timed(make_step(solve_mode, refine_tol, graph=bool(args.lu_graph)), args.steps,
                          args.warmup)
        p_gpu = (p_d.cpu().numpy().view(np.complex64)).reshape(F, n).copy()
        f_gpu = (field_d.cpu().numpy().view(np.complex64)).reshape(F, n_obs).copy()
        print(f"bench: main {1e3 * dt / args.steps:.3f} ms per step", file=sys.stderr, flush=True)
        mode_fields = {}
        if args.workload == "c2" and not args.no_modes:
            # the plain FP32 GPU path (one frequency per launch sequence), the
            # batched and captured-graph FP32 paths and the other mixed mode,
            # each over a few sweeps of the same workload
            variants = [
                ("plain_fp32_per_frequency", "fp32", None, 1, False),
                ("fp32_batched", "fp32", None, None, False),
                ("fp32_lu_graph", "fp32", None, None, True),
                ("mixed_bf16", "mixed_bf16", args.mixed_tol, None, False),
            ]
            for name, mode, tol, batch, graph in variants:
                msteps = max(2, min(10, args.steps // 3))
                mdt, mdiags = timed(make_step(mode, tol, batch, graph), msteps, 1)
                modes[name] = {
                    "value": round(F * msteps / mdt, 3), "ms_per_step": round(1e3 * mdt / msteps, 3),
                    "solve_mode": mode, "batch_frequencies": batch or F, "lu_graph": graph,
                    "refine_tolerance": tol,
                    "dense_solve_ms_per_step": round(1e3 * float(np.mean(
                        [sum(d["dense_solve_seconds"] for d in ds) for ds in mdiags])), 3),
                }
                if mode != "fp32":
                    modes[name]["worst_refine_residual"] = float(max(
                        d["dense_solve_refine_residual_rel"] for d in mdiags[-1]))
                if batch is None:
                    mode_fields[name] = (
                        (p_d.cpu().numpy().view(np.complex64)).reshape(F, n).copy(),
                        (field_d.cpu().numpy().view(np.complex64)).reshape(F, n_obs).copy())
                print(f"bench: {name} {modes[name]['ms_per_step']} ms per step", file=sys.stderr,
                      flush=True)
            plain = modes["plain_fp32_per_frequency"]["value"]
            for name in modes:
                modes[name]["vs_plain_fp32"] = round(modes[name]["value"] / plain, 3)
        # GPU reference sweep for the per-frequency error: FP64-geometry
        # regular assembly and the FP64-accurate solve (GMRES-IR to 1e-13);
        # it is itself checked against the CPU FP64 oracle below
        ref = None
        if not args.no_cpu_baseline:
            s.solve_batch_device(
                F, k_d.data_ptr(), k_d.data_ptr(), q_d.data_ptr(), n_obs, obs_d.data_ptr(),
                field_d.data_ptr(), tag_d.data_ptr(), p_ptr=p_d.data_ptr(), source_tags=[2],
                impedance_source_tag=2, solve_mode="fp64", estimate_rcond=False,
                refine_tolerance=1e-13, refine_max_iterations=20, precise_geometry=True)
            torch.cuda.synchronize()
            ref = ((p_d.cpu().numpy().view(np.complex64)).reshape(F, n).copy(),
                   (field_d.cpu().numpy().view(np.complex64)).reshape(F, n_obs).copy())
            print("bench: reference sweep done", file=sys.stderr, flush=True)
        ref_err = None if ref is None else [
            float(np.linalg.norm(p_gpu[i] - ref[0][i]) / np.linalg.norm(ref[0][i])) for i in range(F)]
    value = F * args.steps / dt
    ms_per_step = 1e3 * dt / args.steps
    vs_plain = round(value / modes["plain_fp32_per_frequency"]["value"], 3) if modes else None

    # stage timings (HIP events on the library stream), per step
    def per_step(key):
        return 1e3 * float(np.mean([sum(d[key] for d in ds) for ds in diags]))

    reg_ms = per_step("regular_assembly_seconds")
    launches_per_step = len({d["batch_index"] for d in diags[0]})
    reg_launch_s = reg_ms * 1e-3 / launches_per_step
    n_img = 4 if sym == "yz+xz" else (1 if sym is None else 2)
    evals_per_step = float(T) ** 2 * 36 * n_img * F
    achieved = evals_per_step / launches_per_step * FLOPS_PER_EVAL / reg_launch_s / 1e12

    # transcendental side of the assembly roofline: rsq + sin + cos per
    # computed pair-point evaluation, each serving both directed blocks
    # (reciprocity), against the device's measured transcendental rate
    from hornlab_metal_bem.native import probe_transcendental_rate

    trans_peak = probe_transcendental_rate()
    trans_launch = 3.0 * evals_per_step / 2.0 / launches_per_step
    trans = {"achieved_ops_per_s": trans_launch / reg_launch_s, "peak_ops_per_s": trans_peak,
             "frac": round(trans_launch / reg_launch_s / trans_peak, 4),
             "ops_per_launch": trans_launch,
             "peak": "hbem_probe_transcendental: v_sin/v_cos/v_rsq in equal parts, 8 chains per "
                     "lane, 8 blocks of 256 per CU"}

    solve_wall = None
    if not args.no_solve_wall:
        # the reference's published number: solve() wall time on the same mesh
        # (host arrays in and out, session setup, result assembly)
        from hornlab_metal_bem import solve

        solve(mesh, config)  # first call pays one-time library and kernel loading
        t0 = time.perf_counter()
        res = solve(mesh, config)
        wall = time.perf_counter() - t0
        solve_wall = {"seconds": round(wall, 4), "frequencies_per_s": round(F / wall, 3),
                      "solve_mode": config.effective_hip_solve_mode,
                      "estimate_rcond": config.hip_estimate_rcond,
                      "frequencies": int(res.frequencies_hz.size)}
        print(f"bench: solve() wall {wall:.3f} s", file=sys.stderr, flush=True)

    threads = min(16, os.cpu_count() or 1)
    cpu = cpu_plain = None
    acc = {}
    acc_ref_check = {}
    if not args.no_cpu_baseline:
        sample = list(range(0, F, 5)) if args.workload == "c2" else [F // 2]
        cps, cdt = cpu_baseline(buffers, sym, freqs, k, q, pts, sample, threads, near_level=1)
        hz = ", ".join(f"{freqs[i]:.0f}" for i in sample)
        cpu = {
            "value": round(cps, 4),
            "unit": "frequencies/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{len(sample)} of the {F} frequencies ({hz} Hz), FP32 regular + "
            f"Sauter-Schwab + level-1 near-field assembly, scipy cgesv, field; {cdt:.1f} s",
            "cpu_model": cpu_model(),
            "host_cpus": os.cpu_count(),
        }
        cps2, cdt2 = cpu_baseline(buffers, sym, freqs, k, q, pts, sample, threads, near_level=0)
        cpu_plain = {"value": round(cps2, 4), "unit": "frequencies/s", "cores": threads,
                     "kind": "port", "sample": f"the same {len(sample)} frequencies without the "
                     f"near-field pass; {cdt2:.1f} s"}
        print("bench: cpu baselines done", file=sys.stderr, flush=True)
        def rel(a, b):
            return float(np.linalg.norm(a - b) / np.linalg.norm(b))

        checked = [0, F - 1] if args.workload == "c2" else [F - 1]
        if ref_err is not None and int(np.argmax(ref_err)) not in checked:
            checked.append(int(np.argmax(ref_err)))  # the worst frequency against the oracle too
        for i in checked:
            p64, f64 = fp64_reference(buffers, sym, k, q, pts, i, threads)
            key = f"{freqs[i]:.0f}Hz"
            acc[key] = {"surface_p_rel_l2": rel(p_gpu[i], p64), "field_rel_l2": rel(f_gpu[i], f64)}
            if ref is not None:
                acc_ref_check[key] = {"surface_p_rel_l2": rel(ref[0][i], p64),
                                      "field_rel_l2": rel(ref[1][i], f64)}
            for name, (pm, fm) in mode_fields.items():
                modes[name].setdefault("accuracy_vs_fp64", {})[key] = {
                    "surface_p_rel_l2": float(np.linalg.norm(pm[i] - p64) / np.linalg.norm(p64)),
                    "field_rel_l2": float(np.linalg.norm(fm[i] - f64) / np.linalg.norm(f64)),
                }
    out = {
        "metric": "frequencies solved/sec (assemble+LU+field) at 1k/8k dofs; rel-L2 vs FP64",
        "value": round(value, 3),
        "unit": "frequencies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / REFERENCE_FREQ_PER_S, 3),
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": f"{'quarter' if sym else 'full'} exponential horn {sym or 'no symmetry'}, "
            f"n={n} P1 dofs, T={T} triangles, {F} log frequencies "
            f"{wl['fmin']:.0f}-{wl['fmax']:.0f} Hz, 3x{wl['angles']} field points at 2 m",
            "n_dofs": n,
            "n_triangles": T,
            "frequencies": F,
            "edge_length_m": round(h, 5),
            "solve_mode": solve_mode,
            "refine_tolerance": refine_tol if solve_mode != "fp32" else None,
            "estimate_rcond": rcond,
            "lu_graph": bool(args.lu_graph),
            "duffy_pairs": info["planned_pairs"],
        },
        "vs_plain_fp32": vs_plain,
        "setup_s": round(info["resident_context_seconds"], 4),
        "stages_ms_per_step": {
            "assembly": round(per_step("assembly_seconds"), 3),
            "regular_assembly": round(reg_ms, 3),
            "singular_assembly": round(per_step("singular_assembly_seconds"), 3),
            "dense_solve": round(per_step("dense_solve_seconds"), 3),
            "rcond": round(per_step("dense_solve_rcond_seconds"), 3),
            "field": round(per_step("field_seconds"), 3),
        },
        "stages_note": "HIP-event times per stage; the rcond estimate runs on a second stream "
        "beside the refinement and the field, so the stages overlap",
        "roofline": {
            "bound": "valu",
            "kernel": "k_assemble_regular",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": REGULAR_TRAFFIC_BYTES.get(args.workload) if args.dofs is None else None,
            "algorithmic": f"T^2*36*(1+I)*F = {evals_per_step:.4g} G/dG evaluations per step x "
            f"{FLOPS_PER_EVAL} FP32 flops (+{TRANS_PER_EVAL} transcendentals) each",
            "launch_ms": round(reg_launch_s * 1e3, 4),
        },
        "roofline_transcendental": trans,
        "modes": modes or None,
        "solve_wall": solve_wall,
        "cpu_baseline": cpu,
        "cpu_baseline_plain": cpu_plain,
        "accuracy_vs_fp64": acc,
        "accuracy_per_frequency": None if ref is None else {
            "reference": "GPU sweep with FP64-geometry regular assembly and the FP64-accurate "
            "solve (GMRES-IR to 1e-13); its own error against the CPU FP64 oracle (FP64 assembly + "
            "zgesv) at the oracle frequencies is reference_vs_fp64",
            "reference_vs_fp64": acc_ref_check,
            "frequencies_hz": [round(float(f), 1) for f in freqs],
            "surface_p_rel_l2": ref_err,
            "rcond": [float(d["dense_solve_rcond"]) for d in diags[-1]],
            "field_rel_l2": [float(np.linalg.norm(f_gpu[i] - ref[1][i]) / np.linalg.norm(ref[1][i]))
                             for i in range(F)],
        },
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
