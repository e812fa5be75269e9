"""Test meshes: the reference's tiny fixtures (tests/test_metal_native.py
of the reference) plus procedural spheres and horns."""
from __future__ import annotations

from types import SimpleNamespace

import numpy as np

from hornlab_metal_bem.geometry import build_geometry_buffers
from hornlab_metal_bem.mesh import make_pure_function_spaces
from hornlab_metal_bem.procedural import HornSpec, horn_body, icosphere, mirror_quarter_to_full


def buffers_from_arrays(vertices_3xn, elements_3xm, l2g, n_dofs, tags):
    grid = SimpleNamespace(
        vertices=np.asarray(vertices_3xn, np.float64),
        elements=np.asarray(elements_3xm, np.int64),
        number_of_elements=int(np.asarray(elements_3xm).shape[1]),
    )
    p1 = SimpleNamespace(local2global=np.asarray(l2g, np.int64), global_dof_count=n_dofs)
    return build_geometry_buffers(grid, np.asarray(tags, np.int32), p1,
                                  SimpleNamespace(global_dof_count=grid.number_of_elements))


def tiny_two_triangles():
    """Two triangles sharing an edge (reference test_metal_native.py:47-69)."""
    return buffers_from_arrays(
        [[0.0, 1.0, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0], [0.0, 0.0, 0.0, 1.0]],
        [[0, 0], [1, 2], [2, 3]],
        [[0, 1, 2], [0, 2, 3]], 4, [1, 2],
    )


def tiny_yz_xz_quarter():
    """Single-triangle yz+xz quarter (reference test_metal_native.py:566-590)."""
    return buffers_from_arrays(
        [[0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.0, 0.0, 0.0]],
        [[0], [1], [2]], [[0, 1, 2]], 3, [2],
    )


def tiny_yz_xz_full():
    """Full four-triangle image of the quarter (reference :593-626)."""
    return buffers_from_arrays(
        [[0.0, 1.0, 0.0, -1.0, 0.0], [0.0, 0.0, 1.0, 0.0, -1.0], [0.0, 0.0, 0.0, 0.0, 0.0]],
        [[0, 0, 0, 0], [1, 2, 4, 3], [2, 3, 1, 4]],
        [[0, 1, 2], [0, 2, 3], [0, 4, 1], [0, 3, 4]], 5, [2, 2, 2, 2],
    )


def mesh_buffers(V, F, tags):
    from hornlab_metal_bem.mesh import make_pure_grid

    grid = make_pure_grid(V, F)
    p1, dp0 = make_pure_function_spaces(grid)
    return build_geometry_buffers(grid, tags, p1, dp0)


def sphere_buffers(radius=0.1, subdivisions=3):
    V, F = icosphere(radius, subdivisions)
    return mesh_buffers(V, F, np.full(F.shape[0], 2, np.int32))


def horn_buffers(edge_length=0.02, quarter=True, seed=0, spec=HornSpec()):
    V, F, T = horn_body(spec, edge_length=edge_length, quarter=quarter, seed=seed)
    return mesh_buffers(V, F, T)


def horn_quarter_and_full(edge_length=0.025, seed=0, spec=HornSpec()):
    Vq, Fq, Tq = horn_body(spec, edge_length=edge_length, quarter=True, seed=seed)
    Vf, Ff, Tf = mirror_quarter_to_full(Vq, Fq, Tq)
    return mesh_buffers(Vq, Fq, Tq), mesh_buffers(Vf, Ff, Tf), (Vq, Vf)


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
