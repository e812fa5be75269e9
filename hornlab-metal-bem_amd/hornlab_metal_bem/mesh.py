"""Triangle surface meshes and P1/DP0 function spaces (NumPy only).

Mirrors the reference loader (hornlab_metal_bem/mesh.py:51-197): seam merge,
degenerate removal, outward-winding check, ``PureGrid`` with Bempp-shaped
``vertices (3, N)`` / ``elements (3, M)`` arrays, and P1 (vertex) / DP0
(triangle) spaces.  ``.msh`` files are read with meshio when it is installed
and otherwise with a small built-in reader of Gmsh ASCII formats 2.2 and 4.1.
"""
from __future__ import annotations

import logging
import warnings
from dataclasses import dataclass
from pathlib import Path

import numpy as np
from numpy.typing import NDArray

from .result import MeshInfo

logger = logging.getLogger(__name__)

_SYMMETRY_SNAP_TOLERANCE = 1.0e-6


@dataclass
class LoadedMesh:
    grid: object
    physical_tags: NDArray[np.int32]
    info: MeshInfo


@dataclass(frozen=True)
class PureGrid:
    """Bempp-shaped mesh view backed by NumPy arrays."""

    vertices: NDArray[np.float64]  # (3, N)
    elements: NDArray[np.int32]  # (3, M)
    volumes: NDArray[np.float64]  # (M,) triangle areas

    @property
    def number_of_elements(self) -> int:
        return int(self.elements.shape[1])


@dataclass(frozen=True)
class PureFunctionSpace:
    local2global: NDArray[np.int32]
    global_dof_count: int


class MeshError(Exception):
    pass


def _triangle_areas(verts: NDArray[np.float64], tris: NDArray[np.int32]) -> NDArray[np.float64]:
    p0, p1, p2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    return np.ascontiguousarray(0.5 * np.linalg.norm(np.cross(p1 - p0, p2 - p0), axis=1))


def make_pure_grid(vertices_nx3: NDArray[np.float64], triangles_nx3: NDArray[np.int32]) -> PureGrid:
    verts = np.asarray(vertices_nx3, dtype=np.float64)
    tris = np.asarray(triangles_nx3, dtype=np.int32)
    return PureGrid(
        vertices=np.ascontiguousarray(verts.T),
        elements=np.ascontiguousarray(tris.T),
        volumes=_triangle_areas(verts, tris),
    )


def make_pure_function_spaces(grid: object) -> tuple[PureFunctionSpace, PureFunctionSpace]:
    """P1 (one dof per vertex) and DP0 (one dof per triangle) spaces."""
    if not hasattr(grid, "elements") or not hasattr(grid, "vertices"):
        raise MeshError("Pure function spaces require grid.vertices and grid.elements")
    elements = np.asarray(grid.elements, dtype=np.int32)
    vertices = np.asarray(grid.vertices)
    if elements.ndim != 2 or elements.shape[0] != 3:
        raise MeshError(f"grid.elements must have shape (3, n_triangles), got {elements.shape}")
    if vertices.ndim != 2 or vertices.shape[0] != 3:
        raise MeshError(f"grid.vertices must have shape (3, n_vertices), got {vertices.shape}")
    p1 = PureFunctionSpace(
        local2global=np.ascontiguousarray(elements.T, dtype=np.int32),
        global_dof_count=int(vertices.shape[1]),
    )
    dp0 = PureFunctionSpace(
        local2global=np.arange(elements.shape[1], dtype=np.int32)[:, None],
        global_dof_count=int(elements.shape[1]),
    )
    return p1, dp0


def open_boundary_edges(triangles_nx3: NDArray[np.int32]) -> NDArray[np.int32]:
    """Sorted vertex pairs of edges used by exactly one triangle."""
    tris = np.asarray(triangles_nx3)
    if tris.size == 0:
        return np.empty((0, 2), dtype=np.int32)
    edges = np.sort(np.concatenate((tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]])), axis=1)
    unique, counts = np.unique(edges, axis=0, return_counts=True)
    return np.ascontiguousarray(unique[counts == 1], dtype=np.int32)


def signed_volume_indicator(verts: NDArray[np.float64], tris: NDArray[np.int32]) -> float:
    """Sum of p0 . (p1 x p2): positive for outward winding of a closed surface."""
    p0, p1, p2 = verts[tris[:, 0]], verts[tris[:, 1]], verts[tris[:, 2]]
    return float(np.sum(p0 * np.cross(p1, p2)))


def merge_duplicate_vertices(
    verts: NDArray[np.float64], tris: NDArray[np.int32], tol: float
) -> tuple[NDArray[np.float64], NDArray[np.int32], int]:
    """Stitch coincident seam vertices (quantised to ``tol``)."""
    if tol <= 0 or len(verts) == 0:
        return verts, tris, 0
    keys = np.round(verts / tol).astype(np.int64)
    _, first, inverse = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    inverse = inverse.reshape(-1)
    if len(first) == len(verts):
        return verts, tris, 0
    return verts[first], inverse[tris].astype(np.int32), len(verts) - len(first)


def loaded_mesh_from_arrays(
    vertices_nx3: NDArray[np.float64],
    triangles_nx3: NDArray[np.int32],
    physical_tags: NDArray[np.int32],
    *,
    physical_groups: dict[int, str] | None = None,
    merge_tol: float = 1e-9,
    validate: bool = True,
    native_symmetry_plane: str | None = None,
) -> LoadedMesh:
    """Build a ``LoadedMesh`` from in-memory arrays with the loader's checks."""
    verts = np.asarray(vertices_nx3, dtype=np.float64)
    tris = np.asarray(triangles_nx3, dtype=np.int32).copy()
    tags = np.asarray(physical_tags, dtype=np.int32).reshape(-1)
    if verts.ndim != 2 or verts.shape[1] != 3:
        raise MeshError("vertices must have shape (N, 3)")
    if tris.ndim != 2 or tris.shape[1] != 3:
        raise MeshError("triangles must have shape (M, 3)")
    if tags.shape[0] != tris.shape[0]:
        raise MeshError("physical_tags must have one entry per triangle")
    verts, tris, merged = merge_duplicate_vertices(verts, tris, merge_tol)
    if merged:
        logger.info("Merged %d duplicate seam vertices", merged)
    valid = ~((tris[:, 0] == tris[:, 1]) | (tris[:, 1] == tris[:, 2]) | (tris[:, 0] == tris[:, 2]))
    if not np.all(valid):
        tris, tags = tris[valid], tags[valid]
    if validate and native_symmetry_plane is None:
        if signed_volume_indicator(verts, tris) < 0:
            raise MeshError(
                "Mesh triangle winding appears inward (signed volume negative). "
                "Canonical meshes must be emitted with outward normals."
            )
    if validate:
        unique = np.unique(tags)
        if not np.any(unique >= 2):
            raise MeshError(f"No velocity source (tag >= 2) found. Tags: {unique.tolist()}")
    grid = make_pure_grid(verts, tris)
    info = MeshInfo(
        n_vertices=int(len(verts)),
        n_triangles=int(len(tris)),
        physical_groups=dict(physical_groups or {}),
        bounding_box_m=(verts.min(axis=0), verts.max(axis=0)),
    )
    return LoadedMesh(grid=grid, physical_tags=tags, info=info)


def load_mesh(
    path: str | Path,
    scale: float = 1.0,
    validate: bool = True,
    merge_tol: float = 1e-9,
    repair_normals: bool = False,
    native_symmetry_plane: str | None = None,
    aperture_tag: int | None = None,
) -> LoadedMesh:
    """Load a Gmsh ``.msh`` triangle surface mesh with physical tags."""
    path = Path(path)
    if not path.exists():
        raise MeshError(f"Mesh file not found: {path}")
    if aperture_tag is not None:
        raise MeshError("aperture_tag (coupled infinite baffle) is not supported on the HIP path")
    verts, tris, tags, names = _read_msh(path)
    verts = verts * float(scale)
    verts, tris, merged = merge_duplicate_vertices(verts, tris, merge_tol)
    if merged:
        logger.info("Merged %d duplicate seam vertices", merged)
    valid = ~((tris[:, 0] == tris[:, 1]) | (tris[:, 1] == tris[:, 2]) | (tris[:, 0] == tris[:, 2]))
    tris, tags = tris[valid], tags[valid]
    if validate:
        if signed_volume_indicator(verts, tris) < 0:
            if repair_normals:
                tris[:, [1, 2]] = tris[:, [2, 1]]
            elif native_symmetry_plane is None:
                raise MeshError(
                    "Mesh triangle winding appears inward (signed volume negative). "
                    "Canonical meshes must be emitted with outward normals by the mesher; "
                    "pass repair_normals=True only for explicit external-mesh compatibility."
                )
        unique = np.unique(tags)
        if not np.any(unique >= 2):
            raise MeshError(f"No velocity source (tag >= 2) found. Tags: {unique.tolist()}")
        if not np.any(unique == 1):
            logger.warning("No rigid wall (tag 1) in mesh")
    if native_symmetry_plane is None:
        suspected = detect_reduced_symmetry_plane(verts, tris)
        if suspected is not None:
            warnings.warn(
                "Mesh may be a reduced native-symmetry mesh "
                f"(suspected plane {suspected!r}) but native_symmetry_plane is None",
                RuntimeWarning,
                stacklevel=2,
            )
    grid = make_pure_grid(verts, tris)
    info = MeshInfo(
        n_vertices=int(len(verts)),
        n_triangles=int(len(tris)),
        physical_groups=names,
        bounding_box_m=(verts.min(axis=0), verts.max(axis=0)),
    )
    return LoadedMesh(grid=grid, physical_tags=tags, info=info)


def detect_reduced_symmetry_plane(
    vertices_nx3: NDArray[np.float64],
    triangles_nx3: NDArray[np.int32],
    *,
    tolerance: float = _SYMMETRY_SNAP_TOLERANCE,
) -> str | None:
    """Conservative guess of a mirror-reduced mesh: every open edge lies on
    candidate coordinate planes the mesh sits on the positive side of
    (reference: hornlab_metal_bem/mesh.py:219-283)."""
    verts = np.asarray(vertices_nx3, dtype=np.float64)
    tris = np.asarray(triangles_nx3, dtype=np.int32)
    if tris.size == 0:
        return None
    edges = open_boundary_edges(tris)
    if edges.size == 0:
        return None
    used = verts[np.unique(tris.reshape(-1))]
    cands = {}
    for plane, c in (("yz", 0), ("xz", 1), ("xy", 2)):
        vals = used[:, c]
        on = np.abs(vals) <= tolerance
        edge_on = np.all(np.abs(verts[edges, c]) <= tolerance, axis=1)
        if vals.min() >= -tolerance and vals.max() > tolerance and on.sum() >= 2 and edge_on.sum() >= 2:
            cands[plane] = edge_on
    for plane, mask in cands.items():
        if np.all(mask):
            return plane
    if "yz" in cands and "xz" in cands and np.all(cands["yz"] | cands["xz"]):
        return "yz+xz"
    return None


def _read_msh(path: Path):
    try:
        import meshio  # type: ignore
    except ImportError:
        meshio = None
    if meshio is not None:
        mesh = meshio.read(path)
        key = "triangle" if "triangle" in mesh.cells_dict else "triangle3"
        if key not in mesh.cells_dict:
            raise MeshError("No triangles found in mesh")
        tris = np.asarray(mesh.cells_dict[key], dtype=np.int32)
        tags = None
        for name, by_type in mesh.cell_data_dict.items():
            if "physical" in name and key in by_type:
                tags = np.asarray(by_type[key], dtype=np.int32)
                break
        if tags is None:
            raise MeshError("Mesh file has no triangle physical-group tags")
        return np.asarray(mesh.points, dtype=np.float64), tris, tags, _physical_names(path)
    return _read_msh_ascii(path)


def _physical_names(path: Path) -> dict[int, str]:
    names: dict[int, str] = {}
    inside = False
    with path.open("r", encoding="utf-8", errors="replace") as fh:
        for raw in fh:
            line = raw.strip()
            if line == "$PhysicalNames":
                inside = True
                continue
            if line == "$EndPhysicalNames":
                break
            if inside:
                parts = line.split(maxsplit=2)
                if len(parts) == 3 and parts[0].isdigit() and int(parts[0]) == 2:
                    names[int(parts[1])] = parts[2].strip().strip('"')
    return names


def _read_msh_ascii(path: Path):
    """Minimal Gmsh ASCII reader (formats 2.2 and 4.1): nodes, 3-node
    triangles and their physical tags."""
    lines = path.read_text(encoding="utf-8", errors="replace").splitlines()
    it = iter(range(len(lines)))
    version = None
    node_ids: list[int] = []
    coords: list[list[float]] = []
    tri_nodes: list[list[int]] = []
    tri_tags: list[int] = []
    entity_phys: dict[tuple[int, int], int] = {}
    i = 0
    while i < len(lines):
        line = lines[i].strip()
        if line == "$MeshFormat":
            version = lines[i + 1].split()[0]
            if lines[i + 1].split()[1] != "0":
                raise MeshError("binary .msh files need meshio")
            i += 2
        elif line == "$Entities" and version and version.startswith("4"):
            counts = [int(v) for v in lines[i + 1].split()]
            j = i + 2 + counts[0]  # skip points
            for dim, cnt in ((1, counts[1]), (2, counts[2]), (3, counts[3])):
                for _ in range(cnt):
                    parts = lines[j].split()
                    tag = int(parts[0])
                    nphys = int(parts[7])
                    if nphys > 0:
                        entity_phys[(dim, tag)] = int(parts[8])
                    j += 1
            i = j
        elif line == "$Nodes":
            if version and version.startswith("4"):
                header = [int(v) for v in lines[i + 1].split()]
                j = i + 2
                for _ in range(header[0]):
                    _, _, _, nblock = (int(v) for v in lines[j].split())
                    ids = [int(lines[j + 1 + k]) for k in range(nblock)]
                    xyz = [[float(v) for v in lines[j + 1 + nblock + k].split()[:3]] for k in range(nblock)]
                    node_ids.extend(ids)
                    coords.extend(xyz)
                    j += 1 + 2 * nblock
                i = j
            else:
                count = int(lines[i + 1])
                for k in range(count):
                    parts = lines[i + 2 + k].split()
                    node_ids.append(int(parts[0]))
                    coords.append([float(v) for v in parts[1:4]])
                i += 2 + count
        elif line == "$Elements":
            if version and version.startswith("4"):
                header = [int(v) for v in lines[i + 1].split()]
                j = i + 2
                for _ in range(header[0]):
                    edim, etag, etype, nblock = (int(v) for v in lines[j].split())
                    for k in range(nblock):
                        if etype == 2:
                            parts = [int(v) for v in lines[j + 1 + k].split()]
                            tri_nodes.append(parts[1:4])
                            tri_tags.append(entity_phys.get((edim, etag), etag))
                    j += 1 + nblock
                i = j
            else:
                count = int(lines[i + 1])
                for k in range(count):
                    parts = [int(v) for v in lines[i + 2 + k].split()]
                    if parts[1] == 2:
                        ntags = parts[2]
                        tri_tags.append(parts[3] if ntags > 0 else 0)
                        tri_nodes.append(parts[3 + ntags : 6 + ntags])
                i += 2 + count
        else:
            i += 1
    del it
    if not tri_nodes:
        raise MeshError("No triangles found in mesh")
    index = {nid: k for k, nid in enumerate(node_ids)}
    tris = np.array([[index[n] for n in tri] for tri in tri_nodes], dtype=np.int32)
    return (
        np.asarray(coords, dtype=np.float64),
        tris,
        np.asarray(tri_tags, dtype=np.int32),
        _physical_names(path),
    )
