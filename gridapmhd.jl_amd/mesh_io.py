"""Unstructured hexahedral meshes: the gmsh 4.1 ASCII reader and VTK output.

Mirrors what the reference reaches through GridapGmsh and ``writevtk``:
``UnstructuredDiscreteModel(GmshDiscreteModel(msh_file))`` with its face
labeling (src/Applications/expansion.jl:278-285, Tube.jl:83) and the
post-processing output of the applications (src/Applications/hunt.jl:263-273,
expansion.jl:246-258).

``HexMesh`` carries the cells (8 vertices each, local vertex v = i + 2j + 4k
of the reference cube, the order of the HIP kernels), the node coordinates
and the geometric classification gmsh writes: every node, boundary quad and
boundary line belongs to one geometric entity (dim, tag), and a physical
group (a tag name) is a set of entities.  A tag of a Q2 node, an RT face or a
thin-wall face is decided on that classification plus the closure of the
tagged boundary quads (a node on the rim of a tagged surface carries the
surface's tags), which is how the reference meshes tag their walls: they
declare the same name on the surfaces, curves and points of a wall.
"""
from __future__ import annotations

import dataclasses
from typing import Dict, Iterable, List, Optional, Sequence, Set, Tuple

import numpy as np

# gmsh hexahedron (type 5) node order -> local vertex v = i + 2j + 4k:
# gmsh 0..3 counter-clockwise at zeta = 0, 4..7 above them
GMSH_HEX_TO_LOCAL = np.array([0, 1, 3, 2, 4, 5, 7, 6])
# the 4 local vertices of local face f (z-, z+, y-, y+, x-, x+)
FACE_VERTS = np.array([[0, 1, 2, 3], [4, 5, 6, 7], [0, 1, 4, 5], [2, 3, 6, 7], [0, 2, 4, 6], [1, 3, 5, 7]])
# the 12 local edges as vertex pairs, x-edges, y-edges, z-edges
EDGE_VERTS = np.array([[0, 1], [2, 3], [4, 5], [6, 7], [0, 2], [1, 3], [4, 6], [5, 7], [0, 4], [1, 5], [2, 6], [3, 7]])
_GMSH_NNODES = {1: 2, 3: 4, 5: 8, 15: 1}


@dataclasses.dataclass
class HexMesh:
    """Unstructured conforming hex mesh with gmsh's geometric classification."""

    node_xyz: np.ndarray                    # (nnodes, 3)
    cells: np.ndarray                       # (ncells, 8) node indices, local vertex order
    node_entity: np.ndarray                 # (nnodes, 2) (dim, tag) of each node
    cell_entity: np.ndarray                 # (ncells, 2)
    quad_entity: Dict[Tuple[int, ...], Tuple[int, int]]   # sorted 4 node ids -> entity
    line_entity: Dict[Tuple[int, int], Tuple[int, int]]   # sorted 2 node ids -> entity
    groups: Dict[str, Set[Tuple[int, int]]]               # tag name -> entities
    isperiodic: Tuple[bool, bool, bool] = (False, False, False)
    nc = None                               # no lattice: not a Cartesian mesh

    @property
    def ncells(self) -> int:
        return int(self.cells.shape[0])

    def cell_xyz(self) -> np.ndarray:
        return self.node_xyz[self.cells]

    def entities(self, tags) -> Set[Tuple[int, int]]:
        """Entities of one tag name or a list of names (unknown names raise)."""
        if tags is None:
            return set()
        names = [tags] if isinstance(tags, str) else list(tags)
        out = set()
        for n in names:
            if n not in self.groups:
                raise KeyError(f"unknown tag {n!r}; the mesh has {sorted(self.groups)}")
            out |= self.groups[n]
        return out

    def add_tag_from_tags(self, name: str, tags: Iterable[str]) -> None:
        """``add_tag_from_tags!`` of Gridap's face labeling."""
        self.groups[name] = self.entities(list(tags))

    def boundary_quads(self):
        """(cell, local face, entity) of every cell face that is a boundary quad of the file."""
        fv = np.sort(self.cells[:, FACE_VERTS], axis=2)          # (ncells, 6, 4)
        out = []
        for c in range(self.ncells):
            for f in range(6):
                e = self.quad_entity.get(tuple(fv[c, f]))
                if e is not None:
                    out.append((c, f, e))
        return out


def read_gmsh(path: str) -> HexMesh:
    """Read a gmsh 4.1 ASCII file of 8-node hexahedra (with the boundary quads,
    lines and points gmsh writes for the physical groups)."""
    with open(path) as fh:
        lines = [l.strip() for l in fh]
    sec = {}
    i = 0
    while i < len(lines):
        if lines[i].startswith("$") and not lines[i].startswith("$End"):
            name = lines[i][1:]
            j = lines.index("$End" + name, i)
            sec[name] = (i + 1, j)
            i = j
        i += 1
    a, _ = sec["MeshFormat"]
    ver = lines[a].split()
    if not ver[0].startswith("4.1") or ver[1] != "0":
        raise ValueError(f"{path}: only gmsh 4.1 ASCII files are read (found format {ver[0]}, file type {ver[1]})")
    # physical names: (dim, physical tag) -> name
    pnames = {}
    if "PhysicalNames" in sec:
        a, b = sec["PhysicalNames"]
        for l in lines[a + 1:b]:
            d, t, nm = l.split(maxsplit=2)
            pnames[(int(d), int(t))] = nm.strip('"')
    # entities: (dim, tag) -> physical tags
    groups: Dict[str, Set[Tuple[int, int]]] = {}
    if "Entities" in sec:
        a, _ = sec["Entities"]
        counts = [int(v) for v in lines[a].split()]
        k = a + 1
        for dim in range(4):
            for _ in range(counts[dim]):
                t = lines[k].split()
                k += 1
                tag = int(t[0])
                p = 4 if dim == 0 else 7
                nph = int(t[p])
                for ph in t[p + 1:p + 1 + nph]:
                    nm = pnames.get((dim, abs(int(ph))), f"physical_{dim}_{abs(int(ph))}")
                    groups.setdefault(nm, set()).add((dim, tag))
    # nodes
    a, _ = sec["Nodes"]
    nblocks, nnodes = int(lines[a].split()[0]), int(lines[a].split()[1])
    tag_of = np.empty(nnodes, dtype=np.int64)
    xyz = np.empty((nnodes, 3))
    ent = np.empty((nnodes, 2), dtype=np.int64)
    k, n = a + 1, 0
    for _ in range(nblocks):
        d, etag, param, m = (int(v) for v in lines[k].split())
        if param:
            raise ValueError(f"{path}: parametric node coordinates are not supported")
        k += 1
        if m == 0:                      # an entity without nodes of its own
            continue
        tag_of[n:n + m] = [int(lines[k + q]) for q in range(m)]
        k += m
        xyz[n:n + m] = [[float(v) for v in lines[k + q].split()[:3]] for q in range(m)]
        k += m
        ent[n:n + m] = (d, etag)
        n += m
    # nodes are numbered by their tags (gmsh numbers them 1..N)
    order = np.argsort(tag_of, kind="stable")
    tag_of, xyz, ent = tag_of[order], xyz[order], ent[order]
    index = {int(t): q for q, t in enumerate(tag_of)}
    # elements
    a, _ = sec["Elements"]
    nblocks = int(lines[a].split()[0])
    k = a + 1
    cells, cell_ent = [], []
    quads, lines_e = {}, {}
    for _ in range(nblocks):
        d, etag, etype, m = (int(v) for v in lines[k].split())
        k += 1
        if etype in (2, 4, 6, 7):
            raise ValueError(f"{path}: triangle / tetrahedron / prism / pyramid elements (gmsh type {etype}): the "
                             "hexahedral Q2 / dP1 / RT1 / dQ1 discretisation needs an all-hexahedral mesh")
        if etype not in _GMSH_NNODES:
            raise ValueError(f"{path}: element type {etype} is not supported (8-node hexahedra meshes only)")
        for q in range(m):
            nodes = [index[int(v)] for v in lines[k + q].split()[1:]]
            if etype == 5:
                cells.append(np.asarray(nodes)[GMSH_HEX_TO_LOCAL])
                cell_ent.append((d, etag))
            elif etype == 3:
                quads[tuple(sorted(nodes))] = (d, etag)
            elif etype == 1:
                lines_e[tuple(sorted(nodes))] = (d, etag)
        k += m
    if not cells:
        raise ValueError(f"{path}: no hexahedra")
    cells = np.asarray(cells, dtype=np.int64)
    mesh = HexMesh(node_xyz=xyz, cells=cells, node_entity=ent, cell_entity=np.asarray(cell_ent, dtype=np.int64),
                   quad_entity=quads, line_entity=lines_e, groups=groups)
    _orient(mesh)
    return mesh


def _box_entity_dim(e: int) -> int:
    return 0 if e <= 8 else (1 if e <= 20 else (2 if e <= 26 else 3))


def hexmesh_from_cartesian(cm) -> HexMesh:
    """The unstructured view of a non-periodic CartesianMesh: nodes, cells,
    the box entities (1-8 corners, 9-20 edges, 21-26 faces, 27 interior) as
    the geometric classification, and the mesh's tags as physical groups."""
    if any(cm.isperiodic):
        raise ValueError("hexmesh_from_cartesian: periodic meshes have no unstructured view here")
    nx, ny, nz = cm.nc
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    I, J, K = (a.transpose(2, 1, 0).ravel() for a in (I, J, K))
    ent = cm.lattice_entity(2 * I, 2 * J, 2 * K)
    node_entity = np.stack([[_box_entity_dim(int(e)) for e in ent], ent], axis=1).astype(np.int64)
    cells = cm.cell_nodes()
    quads, lines_e = {}, {}
    for c in range(cells.shape[0]):
        for f in range(6):
            vs = cells[c, FACE_VERTS[f]]
            common = None
            for x in range(21, 27):
                if all(x == int(ent[v]) or x in _box_closure(int(ent[v])) for v in vs):
                    common = x
            if common is not None:
                quads[tuple(sorted(int(v) for v in vs))] = (2, common)
        for a_, b_ in EDGE_VERTS:
            va, vb = int(cells[c, a_]), int(cells[c, b_])
            for x in range(9, 21):
                if all(x == int(ent[v]) or x in _box_closure(int(ent[v])) for v in (va, vb)):
                    lines_e[tuple(sorted((va, vb)))] = (1, x)
    groups = {}
    for name, es in cm.tags.items():
        groups[name] = {(_box_entity_dim(int(e)), int(e)) for e in es}
    return HexMesh(node_xyz=cm.node_xyz.copy(), cells=cells.copy(), node_entity=node_entity,
                   cell_entity=np.tile(np.array([[3, 27]]), (cells.shape[0], 1)), quad_entity=quads,
                   line_entity=lines_e, groups=groups)


def _box_closure(e: int) -> Set[int]:
    """Box entities whose closure contains entity e (e itself excluded)."""
    from .meshers import box_entity
    out = set()
    dims = {}
    for w in [(a, b, c) for a in (None, 0, 1) for b in (None, 0, 1) for c in (None, 0, 1)]:
        dims[box_entity(list(w))] = w
    we = dims[e]
    for x, wx in dims.items():
        if x != e and all(a is None or a == b for a, b in zip(wx, we)):
            out.add(x)
    return out


def _orient(mesh: HexMesh) -> None:
    """Reorder the vertices of left-handed cells (negative Jacobian at the
    centre) by mirroring them in the local z direction."""
    X = mesh.cell_xyz()
    sx = np.array([-1, 1, -1, 1, -1, 1, -1, 1]) / 4.0
    sy = np.array([-1, -1, 1, 1, -1, -1, 1, 1]) / 4.0
    sz = np.array([-1, -1, -1, -1, 1, 1, 1, 1]) / 4.0
    J = np.stack([np.einsum("v,cvi->ci", s, X) for s in (sx, sy, sz)], axis=2)
    det = np.linalg.det(J)
    flip = det < 0
    mesh.cells[flip] = mesh.cells[flip][:, [4, 5, 6, 7, 0, 1, 2, 3]]


def write_gmsh(path: str, mesh: HexMesh) -> None:
    """Write a HexMesh as gmsh 4.1 ASCII (hexahedra, boundary quads and lines,
    nodes grouped by entity, physical groups)."""
    ents: Dict[int, Set[int]] = {0: set(), 1: set(), 2: set(), 3: set()}
    for d, t in map(tuple, mesh.node_entity):
        ents[d].add(t)
    for d, t in mesh.quad_entity.values():
        ents[d].add(t)
    for d, t in mesh.line_entity.values():
        ents[d].add(t)
    for d, t in map(tuple, mesh.cell_entity):
        ents[d].add(t)
    names = sorted(mesh.groups)
    phys = {}  # (dim, name) -> physical tag
    for q, nm in enumerate(names):
        for d in range(4):
            if any(e[0] == d for e in mesh.groups[nm]):
                phys[(d, nm)] = q + 1
    out = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$PhysicalNames", str(len(phys))]
    out += [f'{d} {t} "{nm}"' for (d, nm), t in sorted(phys.items(), key=lambda kv: (kv[0][0], kv[1]))]
    out += ["$EndPhysicalNames", "$Entities", " ".join(str(len(ents[d])) for d in range(4))]
    for d in range(4):
        for t in sorted(ents[d]):
            ph = [phys[(d, nm)] for nm in names if (d, t) in mesh.groups[nm]]
            box = "0 0 0" if d == 0 else "0 0 0 0 0 0"
            tail = "" if d == 0 else " 0"
            out.append(f"{t} {box} {len(ph)} " + " ".join(map(str, ph)) + tail)
    out.append("$EndEntities")
    # nodes by entity
    order = np.lexsort((mesh.node_entity[:, 1], mesh.node_entity[:, 0]))
    blocks = []
    for (d, t) in sorted({tuple(e) for e in mesh.node_entity}):
        sel = order[(mesh.node_entity[order, 0] == d) & (mesh.node_entity[order, 1] == t)]
        blocks.append((d, t, sel))
    nn = mesh.node_xyz.shape[0]
    out += ["$Nodes", f"{len(blocks)} {nn} 1 {nn}"]
    for d, t, sel in blocks:
        out.append(f"{d} {t} 0 {sel.size}")
        out += [str(q + 1) for q in sel]
        out += ["%.17g %.17g %.17g" % tuple(mesh.node_xyz[q]) for q in sel]
    out.append("$EndNodes")
    eb = []
    for (d, t) in sorted({tuple(e) for e in mesh.cell_entity}):
        sel = np.nonzero((mesh.cell_entity[:, 0] == d) & (mesh.cell_entity[:, 1] == t))[0]
        eb.append((d, t, 5, [mesh.cells[c][GMSH_HEX_TO_LOCAL] for c in sel]))
    for kind, table, etype in (("quad", mesh.quad_entity, 3), ("line", mesh.line_entity, 1)):
        by = {}
        for key, e in table.items():
            by.setdefault(e, []).append(key)
        for e in sorted(by):
            eb.append((e[0], e[1], etype, [np.asarray(k) for k in by[e]]))
    ne = sum(len(b[3]) for b in eb)
    out += ["$Elements", f"{len(eb)} {ne} 1 {ne}"]
    tag = 1
    for d, t, etype, els in eb:
        out.append(f"{d} {t} {etype} {len(els)}")
        for el in els:
            out.append(f"{tag} " + " ".join(str(int(v) + 1) for v in el))
            tag += 1
    out.append("$EndElements")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


# ---------------------------------------------------------------------------
# VTK output

_VTK_HEX_FROM_LOCAL = np.array([0, 1, 3, 2, 4, 5, 7, 6])   # VTK_HEXAHEDRON node order


def write_vtu(path: str, cell_xyz: np.ndarray, point_data: Optional[Dict[str, np.ndarray]] = None,
              cell_data: Optional[Dict[str, np.ndarray]] = None) -> str:
    """Write hexahedral cells as a VTK XML unstructured grid (ASCII).

    ``cell_xyz`` (ncells, 8, 3) in local vertex order; ``point_data`` arrays
    are per cell vertex, shape (ncells, 8) or (ncells, 8, 3) (every cell keeps
    its own 8 points, so discontinuous fields are written as they are);
    ``cell_data`` arrays are (ncells,) or (ncells, 3)."""
    if not path.endswith(".vtu"):
        path += ".vtu"
    C = cell_xyz.shape[0]
    pts = cell_xyz[:, _VTK_HEX_FROM_LOCAL, :].reshape(-1, 3)

    def arr(name, a, ncomp):
        a = np.asarray(a, dtype=np.float64).reshape(-1)
        return (f'<DataArray type="Float64" Name="{name}" NumberOfComponents="{ncomp}" format="ascii">\n'
                + " ".join("%.10g" % v for v in a) + "\n</DataArray>\n")

    out = ['<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n'
           f'<UnstructuredGrid>\n<Piece NumberOfPoints="{8 * C}" NumberOfCells="{C}">\n<Points>\n',
           arr("Points", pts, 3), "</Points>\n<Cells>\n",
           '<DataArray type="Int64" Name="connectivity" format="ascii">\n'
           + " ".join(map(str, range(8 * C))) + "\n</DataArray>\n",
           '<DataArray type="Int64" Name="offsets" format="ascii">\n'
           + " ".join(str(8 * (c + 1)) for c in range(C)) + "\n</DataArray>\n",
           '<DataArray type="UInt8" Name="types" format="ascii">\n' + " ".join(["12"] * C) + "\n</DataArray>\n",
           "</Cells>\n<PointData>\n"]
    for name, a in (point_data or {}).items():
        a = np.asarray(a)
        nc = 1 if a.ndim == 2 else a.shape[2]
        out.append(arr(name, a[:, _VTK_HEX_FROM_LOCAL].reshape(C * 8, nc), nc))
    out.append("</PointData>\n<CellData>\n")
    for name, a in (cell_data or {}).items():
        a = np.asarray(a)
        out.append(arr(name, a, 1 if a.ndim == 1 else a.shape[1]))
    out.append("</CellData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")
    with open(path, "w") as fh:
        fh.write("".join(out))
    return path
