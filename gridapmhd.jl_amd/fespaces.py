"""Cell-wise DOF tables of the four H1-HDiv fields on a Cartesian hex mesh.

Mirrors ``setup_fe_spaces`` / ``fe_space_u/p/j/phi`` (src/fespaces.jl:13-114)
for the hexahedral default pair of the reference (src/parameters.jl:436-441,
517-526): Q2 vector Lagrangian velocity with strong Dirichlet tags,
discontinuous P1 pressure, Raviart-Thomas order 1 current with strong normal
Dirichlet tags, discontinuous Q1 potential.

Numbering follows Gridap's conforming rule: every d-face owns its dofs; own
dofs are numbered by dimension (vertices, then edges, faces, cells), within a
dimension by face id; face ids are assigned in first-appearance order over
cells in cell order and local face order; free dofs count 0, 1, ...; Dirichlet
dofs are numbered separately.  L2 spaces are cell-major.  The multifield
numbering is Consecutive [u; p; j; phi] (src/fespaces.jl:5).

Local cell dof order (129 per cell), shared with the HIP kernels:
  u   [0, 81)    c*27 + a, Q2 node a = i + 3j + 9k at (i/2, j/2, k/2)
  p   [81, 85)   dP1 basis 1-x-y-z, x, y, z (reference coordinates)
  j   [85, 121)  face dofs (faces z-, z+, y-, y+, x-, x+; 4 each, tangential
                 2x2 Gauss points, lower axis fastest), then 12 interior dofs
  phi [121, 129) dQ1 at vertex v = i + 2j + 4k
"""
from __future__ import annotations

import dataclasses
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from .meshers import EDGE_OFFSETS, FACE_OFFSETS, FACE_AXIS, FACE_SIDE, CartesianMesh

NU, NP, NJ, NPHI = 81, 4, 36, 8
NDOF = NU + NP + NJ + NPHI
OFF_U, OFF_P, OFF_J, OFF_PHI = 0, 81, 85, 121
G2 = (0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0))

# Q2 lattice offsets of the 27 local nodes a = i + 3j + 9k
Q2_OFFSETS = np.array([(i, j, k) for k in range(3) for j in range(3) for i in range(3)], dtype=np.int64)


def _first_appearance_ids(keys: np.ndarray) -> (np.ndarray, int):
    """Ids 0..n-1 of ``keys`` (flattened, cell-major) in first-appearance order."""
    flat = keys.ravel()
    uniq, first, inv = np.unique(flat, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    return rank[inv].reshape(keys.shape), uniq.size


def _tangential_axes(d: int):
    return {0: (1, 2), 1: (0, 2), 2: (0, 1)}[d]


@dataclasses.dataclass
class Spaces:
    """DOF tables of the multifield space (u, p, j, phi) on one mesh."""

    mesh: CartesianMesh
    nfree: Dict[str, int]
    ndir: Dict[str, int]
    cell_dofs: np.ndarray      # (ncells, 129) int32: >=0 multifield free id, <0 -(k+1) Dirichlet index
    rt_sign: np.ndarray        # (ncells, 6) int8
    dir_vals: np.ndarray       # (ndir_total,) Dirichlet values (u first, then j)
    # bookkeeping for interpolation / Dirichlet values
    u_node_rank: np.ndarray    # (ncells, 27) global Q2 node index (node order)
    u_node_free: np.ndarray    # (n_nodes,) bool
    u_node_dof: np.ndarray     # (n_nodes,) first field-local id (free: >=0, dirichlet: -(k+1)), 3 consecutive
    face_ids: np.ndarray       # (ncells, 6) global face id
    face_dirichlet: np.ndarray  # (n_faces,) bool
    face_dof: np.ndarray       # (n_faces,) first field-local j id of the face's 4 dofs
    cell_verts: np.ndarray = None  # (ncells, 8) topological vertex ids
    style: str = "consecutive"
    bc_spec: dict = None           # the u / j Dirichlet tags and thin walls these tables were built from
    thin_wall_faces: List[dict] = dataclasses.field(default_factory=list)
    node_tagger: Optional[Callable] = None   # tags -> bool mask over the Q2 nodes
    face_tagger: Optional[Callable] = None   # tags -> bool mask over the faces

    @property
    def ncells(self) -> int:
        return self.mesh.ncells

    @property
    def field_order(self):
        return ("u", "j", "p", "phi") if self.style == "badia2024" else ("u", "p", "j", "phi")

    @property
    def offsets(self) -> Dict[str, int]:
        o, out = 0, {}
        for f in self.field_order:
            out[f] = o
            o += self.nfree[f]
        return out

    @property
    def ndofs(self) -> int:
        return sum(self.nfree.values())

    def field_slices(self) -> Dict[str, slice]:
        off = self.offsets
        return {f: slice(off[f], off[f] + self.nfree[f]) for f in ("u", "p", "j", "phi")}


def setup_fe_spaces(mesh: CartesianMesh, u_tags, j_tags, thin_walls: Sequence[dict] = (),
                    style: str = "consecutive") -> Spaces:
    """Build the dof tables.  ``u_tags``/``j_tags`` are Dirichlet tag names (or
    box entity ids); ``thin_walls`` is a list of dicts with ``domain`` (tags),
    ``cw``, ``tau``, ``jw`` (src/parameters.jl:938-947).  ``style`` is the
    multifield numbering (src/fespaces.jl:4-9): "consecutive" = [u; p; j; phi],
    "badia2024" = [u; j], p, phi.  ``mesh`` is a CartesianMesh or an
    unstructured ``mesh_io.HexMesh`` (gmsh)."""
    if getattr(mesh, "nc", None) is None:
        return _setup_hex_spaces(mesh, u_tags, j_tags, thin_walls, style)
    nx, ny, nz = mesh.nc
    ncells = mesh.ncells
    LX, LY, LZ = mesh.lattice_shape()
    ijk = mesh.cell_ijk()
    base = 2 * ijk  # lattice coords of cell corner

    def wrap(P):
        P = P.copy()
        for a, n in enumerate((LX, LY, LZ)):
            if mesh.isperiodic[a]:
                P[..., a] %= n
        return P

    def lin(P):
        return P[..., 0] + LX * (P[..., 1] + LY * P[..., 2])

    # ---- topology: vertex / edge / face / cell ids -------------------------
    vert_pts = wrap(base[:, None, :] + 2 * np.array([(v & 1, (v >> 1) & 1, (v >> 2) & 1) for v in range(8)])[None])
    nvx = nx if mesh.isperiodic[0] else nx + 1
    nvy = ny if mesh.isperiodic[1] else ny + 1
    nvz = nz if mesh.isperiodic[2] else nz + 1
    vert_id = (vert_pts[..., 0] // 2) + nvx * ((vert_pts[..., 1] // 2) + nvy * (vert_pts[..., 2] // 2))
    n_verts = nvx * nvy * nvz
    edge_pts = wrap(base[:, None, :] + EDGE_OFFSETS[None])
    edge_id, n_edges = _first_appearance_ids(lin(edge_pts))
    face_pts = wrap(base[:, None, :] + FACE_OFFSETS[None])
    face_id, n_faces = _first_appearance_ids(lin(face_pts))

    # ---- Q2 node order: vertices, edges, faces, cells ----------------------
    node_of_lattice = np.full(LX * LY * LZ, -1, dtype=np.int64)
    node_of_lattice[lin(vert_pts).ravel()] = vert_id.ravel()
    node_of_lattice[lin(edge_pts).ravel()] = n_verts + edge_id.ravel()
    node_of_lattice[lin(face_pts).ravel()] = n_verts + n_edges + face_id.ravel()
    cell_pts = base + 1
    node_of_lattice[lin(cell_pts)] = n_verts + n_edges + n_faces + np.arange(ncells)
    n_nodes = n_verts + n_edges + n_faces + ncells
    assert (node_of_lattice >= 0).all()

    # lattice coords of every node (for tags and positions)
    lat = np.empty((n_nodes, 3), dtype=np.int64)
    allI = np.arange(LX * LY * LZ)
    lat[node_of_lattice] = np.stack([allI % LX, (allI // LX) % LY, allI // (LX * LY)], axis=1)

    u_dir_node = mesh.tagged(u_tags, lat[:, 0], lat[:, 1], lat[:, 2]) if u_tags else np.zeros(n_nodes, bool)
    node_free = ~u_dir_node
    u_node_dof = np.empty(n_nodes, dtype=np.int64)
    fr = np.cumsum(node_free) - node_free  # exclusive count
    dr = np.cumsum(~node_free) - (~node_free)
    u_node_dof[node_free] = 3 * fr[node_free]
    u_node_dof[~node_free] = -(3 * dr[~node_free]) - 1
    n_u_free = 3 * int(node_free.sum())
    n_u_dir = 3 * int((~node_free).sum())

    cell_q2_pts = wrap(base[:, None, :] + Q2_OFFSETS[None])
    u_node_rank = node_of_lattice[lin(cell_q2_pts)]

    # ---- RT: faces (4 dofs each) then cells (12 each) ----------------------
    face_lat = np.empty((n_faces, 3), dtype=np.int64)
    face_lat[face_id.ravel()] = face_pts.reshape(-1, 3)
    j_dir_face = mesh.tagged(j_tags, face_lat[:, 0], face_lat[:, 1], face_lat[:, 2]) if j_tags else np.zeros(n_faces, bool)
    face_free = ~j_dir_face
    ffr = np.cumsum(face_free) - face_free
    fdr = np.cumsum(~face_free) - (~face_free)
    face_dof = np.empty(n_faces, dtype=np.int64)
    face_dof[face_free] = 4 * ffr[face_free]
    face_dof[~face_free] = -(4 * fdr[~face_free]) - 1
    n_j_face_free = 4 * int(face_free.sum())
    n_j_free = n_j_face_free + 12 * ncells
    n_j_dir = 4 * int((~face_free).sum())

    # ---- assemble the multifield table --------------------------------------
    n_p, n_phi = 4 * ncells, 8 * ncells
    if style == "badia2024":
        off_u, off_j = 0, n_u_free
        off_p = off_j + n_j_free
        off_phi = off_p + n_p
    elif style == "consecutive":
        off_u, off_p = 0, n_u_free
        off_j = off_p + n_p
        off_phi = off_j + n_j_free
    else:
        raise ValueError(f"unknown multifield style {style!r}")
    cd = np.empty((ncells, NDOF), dtype=np.int64)
    nd = u_node_dof[u_node_rank]  # (ncells, 27)
    for c in range(3):
        free = nd >= 0
        col = np.where(free, off_u + nd + c, -((-nd - 1) + c) - 1)
        cd[:, c * 27:(c + 1) * 27] = col
    cells = np.arange(ncells, dtype=np.int64)
    cd[:, OFF_P:OFF_P + 4] = off_p + 4 * cells[:, None] + np.arange(4)[None]
    fd = face_dof[face_id]  # (ncells, 6)
    for f in range(6):
        for k in range(4):
            v = fd[:, f]
            # Dirichlet j indices follow the u Dirichlet block
            cd[:, OFF_J + 4 * f + k] = np.where(v >= 0, off_j + v + k, -(n_u_dir + (-v - 1) + k) - 1)
    cd[:, OFF_J + 24:OFF_J + 36] = off_j + n_j_face_free + 12 * cells[:, None] + np.arange(12)[None]
    cd[:, OFF_PHI:OFF_PHI + 8] = off_phi + 8 * cells[:, None] + np.arange(8)[None]
    assert cd.max() < 2**31

    sp = Spaces(
        node_tagger=lambda tags: mesh.tagged(tags, lat[:, 0], lat[:, 1], lat[:, 2]),
        face_tagger=lambda tags: mesh.tagged(tags, face_lat[:, 0], face_lat[:, 1], face_lat[:, 2]),
        mesh=mesh,
        nfree={"u": n_u_free, "p": n_p, "j": n_j_free, "phi": n_phi},
        ndir={"u": n_u_dir, "j": n_j_dir},
        cell_dofs=cd.astype(np.int32),
        rt_sign=np.ones((ncells, 6), dtype=np.int8),
        dir_vals=np.zeros(n_u_dir + n_j_dir),
        u_node_rank=u_node_rank,
        u_node_free=node_free,
        u_node_dof=u_node_dof,
        face_ids=face_id,
        face_dirichlet=j_dir_face,
        face_dof=face_dof,
        cell_verts=vert_id.astype(np.int32),
        style=style,
        bc_spec=dict(u_tags=u_tags, j_tags=j_tags, thin_walls=list(thin_walls)),
    )
    # ---- thin-wall boundary faces -------------------------------------------
    for tw in thin_walls:
        cells_l, lf_l = [], []
        for f in range(6):
            a, s = FACE_AXIS[f], FACE_SIDE[f]
            if mesh.isperiodic[a]:
                continue
            n = mesh.nc[a]
            on = ijk[:, a] == (n - 1 if s else 0)
            P = face_pts[:, f, :]
            sel = on & mesh.tagged(tw["domain"], P[:, 0], P[:, 1], P[:, 2])
            cells_l.append(np.nonzero(sel)[0])
            lf_l.append(np.full(int(sel.sum()), f))
        fc = np.concatenate(cells_l) if cells_l else np.zeros(0, np.int64)
        fl = np.concatenate(lf_l) if lf_l else np.zeros(0, np.int64)
        order = np.lexsort((fl, fc))
        sp.thin_wall_faces.append(dict(cell=fc[order].astype(np.int32), lface=fl[order].astype(np.int8),
                                       tau=float(tw["tau"]), cw=float(tw["cw"]), jw=float(tw.get("jw", 0.0))))
    return sp


def _first_appearance_rows(rows: np.ndarray) -> (np.ndarray, int):
    """Ids 0..n-1 of the rows of ``rows`` (k, m) in first-appearance order."""
    _, first, inv = np.unique(rows, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    return rank[inv], first.size


def _face_gauss_points(cell_xyz: np.ndarray, faces: np.ndarray) -> np.ndarray:
    """Physical positions (n, 4, 3) of the 2x2 face Gauss points (the RT face
    dof points, k = g1 + 2 g2) of local faces ``faces`` (n,) of cells ``cell_xyz`` (n, 8, 3)."""
    n = faces.size
    ref = np.zeros((6, 4, 3))
    for f in range(6):
        d, t = FACE_AXIS[f], _tangential_axes(FACE_AXIS[f])
        for k in range(4):
            ref[f, k, d] = float(FACE_SIDE[f])
            ref[f, k, t[0]] = G2[k % 2]
            ref[f, k, t[1]] = G2[k // 2]
    xi = ref[faces]
    N, _ = _q1_shape(xi.reshape(-1, 3))
    return np.einsum("nkv,nvi->nki", N.reshape(n, 4, 8), cell_xyz)


def _setup_hex_spaces(mesh, u_tags, j_tags, thin_walls, style) -> "Spaces":
    """``setup_fe_spaces`` on an unstructured hex mesh (``mesh_io.HexMesh``).

    Topology: vertices are the mesh nodes used by cells (in node order),
    edges and faces get ids in first-appearance order over cells and local
    edges / faces.  Q2 nodes and their tags: a node is Dirichlet for a tag if
    its geometric entity (the gmsh node, line or quad classification) is in
    the tag, or it lies on the closure of a boundary quad of the tag.  RT face
    dofs belong to the first cell that has the face; another cell sharing it
    sees the same four Gauss points in its own face coordinates (matched by
    physical position) and the normal of its local axis either way (sign
    +1 when the face is on opposite local sides of the two cells, -1 when on
    the same side)."""
    from .mesh_io import EDGE_VERTS, FACE_VERTS
    cells = mesh.cells
    ncells = mesh.ncells
    used, vid = np.unique(cells, return_inverse=True)
    vid = vid.reshape(ncells, 8)
    nv = used.size
    ev = np.sort(vid[:, EDGE_VERTS], axis=2)                        # (C, 12, 2)
    edge_id, n_edges = _first_appearance_ids(ev[..., 0].astype(np.int64) * nv + ev[..., 1])
    fv = np.sort(vid[:, FACE_VERTS], axis=2)                        # (C, 6, 4)
    fid_flat, n_faces = _first_appearance_rows(fv.reshape(-1, 4))
    face_id = fid_flat.reshape(ncells, 6)
    # local Q2 node -> (kind, local entity)
    node_of_local = np.empty((ncells, 27), dtype=np.int64)
    for a, (i, j, k) in enumerate(Q2_OFFSETS):
        ones = [ax for ax, v in enumerate((i, j, k)) if v == 1]
        b = [v // 2 for v in (i, j, k)]
        if not ones:
            node_of_local[:, a] = vid[:, b[0] + 2 * b[1] + 4 * b[2]]
        elif len(ones) == 1:
            e = ones[0]
            le = {0: b[1] + 2 * b[2], 1: 4 + b[0] + 2 * b[2], 2: 8 + b[0] + 2 * b[1]}[e]
            node_of_local[:, a] = nv + edge_id[:, le]
        elif len(ones) == 2:
            d = ({0, 1, 2} - set(ones)).pop()
            lf = 2 * (2 - d) + (i, j, k)[d] // 2
            node_of_local[:, a] = nv + n_edges + face_id[:, lf]
        else:
            node_of_local[:, a] = nv + n_edges + n_faces + np.arange(ncells)
    n_nodes = nv + n_edges + n_faces + ncells
    # boundary quads: (cell, local face, entity)
    bq = mesh.boundary_quads()
    face_edges = [[e for e in range(12) if set(EDGE_VERTS[e]) <= set(FACE_VERTS[f])] for f in range(6)]

    def tagged_nodes(tags):
        ents = mesh.entities(tags)
        out = np.zeros(n_nodes, dtype=bool)
        if not ents:
            return out
        ne = np.array([tuple(e) in ents for e in mesh.node_entity[used]])
        out[:nv] = ne
        for key, e in mesh.line_entity.items():
            if e in ents:
                a_, b_ = np.searchsorted(used, key)
                hit = (ev[..., 0] == min(a_, b_)) & (ev[..., 1] == max(a_, b_))
                out[nv + edge_id[hit]] = True
        for c, f, e in bq:
            if e in ents:
                out[nv + n_edges + face_id[c, f]] = True
                out[vid[c, FACE_VERTS[f]]] = True
                out[nv + edge_id[c, face_edges[f]]] = True
        return out

    u_dir_node = tagged_nodes(u_tags) if u_tags else np.zeros(n_nodes, bool)
    node_free = ~u_dir_node
    u_node_dof = np.empty(n_nodes, dtype=np.int64)
    fr = np.cumsum(node_free) - node_free
    dr = np.cumsum(~node_free) - (~node_free)
    u_node_dof[node_free] = 3 * fr[node_free]
    u_node_dof[~node_free] = -(3 * dr[~node_free]) - 1
    n_u_free = 3 * int(node_free.sum())
    n_u_dir = 3 * int((~node_free).sum())
    # RT: face owners, point permutations, normal signs
    flat_first = np.full(n_faces, -1, dtype=np.int64)
    order = np.arange(ncells * 6)
    flat_first[fid_flat[::-1]] = order[::-1]                       # first (cell, face) of each face
    owner_c, owner_f = flat_first // 6, flat_first % 6
    allc = np.repeat(np.arange(ncells), 6)
    allf = np.tile(np.arange(6), ncells)
    xyz = mesh.cell_xyz()
    P = _face_gauss_points(xyz[allc], allf)                         # (C*6, 4, 3)
    Po = P[flat_first[fid_flat]]                                    # owner's points of every cell face
    dist = np.linalg.norm(P[:, :, None, :] - Po[:, None, :, :], axis=3)   # (C*6, 4 own, 4 owner)
    perm = np.argmin(dist, axis=2)
    scale = np.linalg.norm(Po[:, 0] - Po[:, 3], axis=1)[:, None] + 1e-300
    if not (np.take_along_axis(dist, perm[..., None], 2)[..., 0] <= 1e-8 * scale).all() or \
            not (np.sort(perm, axis=1) == np.arange(4)).all():
        raise ValueError("a shared face's Gauss points do not match between its cells (non-conforming mesh?)")
    side = FACE_SIDE[allf]
    is_owner = flat_first[fid_flat] == np.arange(ncells * 6)
    sign = np.where(is_owner | (side != FACE_SIDE[owner_f[fid_flat]]), 1, -1).reshape(ncells, 6)
    perm = perm.reshape(ncells, 6, 4)
    def tagged_faces(tags):
        ents = mesh.entities(tags)
        out = np.zeros(n_faces, dtype=bool)
        for c, f, e in bq:
            if e in ents:
                out[face_id[c, f]] = True
        return out

    # Dirichlet faces of j
    face_dir = tagged_faces(j_tags) if j_tags else np.zeros(n_faces, dtype=bool)
    face_free = ~face_dir
    ffr = np.cumsum(face_free) - face_free
    fdr = np.cumsum(~face_free) - (~face_free)
    face_dof = np.empty(n_faces, dtype=np.int64)
    face_dof[face_free] = 4 * ffr[face_free]
    face_dof[~face_free] = -(4 * fdr[~face_free]) - 1
    n_j_face_free = 4 * int(face_free.sum())
    n_j_free = n_j_face_free + 12 * ncells
    n_j_dir = 4 * int((~face_free).sum())
    n_p, n_phi = 4 * ncells, 8 * ncells
    if style == "badia2024":
        off_u, off_j = 0, n_u_free
        off_p = off_j + n_j_free
        off_phi = off_p + n_p
    elif style == "consecutive":
        off_u, off_p = 0, n_u_free
        off_j = off_p + n_p
        off_phi = off_j + n_j_free
    else:
        raise ValueError(f"unknown multifield style {style!r}")
    cd = np.empty((ncells, NDOF), dtype=np.int64)
    nd = u_node_dof[node_of_local]
    for c in range(3):
        free = nd >= 0
        cd[:, c * 27:(c + 1) * 27] = np.where(free, off_u + nd + c, -((-nd - 1) + c) - 1)
    cidx = np.arange(ncells, dtype=np.int64)
    cd[:, OFF_P:OFF_P + 4] = off_p + 4 * cidx[:, None] + np.arange(4)[None]
    fd = face_dof[face_id]
    for f in range(6):
        v = fd[:, f]
        for k in range(4):
            kk = perm[:, f, k]
            cd[:, OFF_J + 4 * f + k] = np.where(v >= 0, off_j + v + kk, -(n_u_dir + (-v - 1) + kk) - 1)
    cd[:, OFF_J + 24:OFF_J + 36] = off_j + n_j_face_free + 12 * cidx[:, None] + np.arange(12)[None]
    cd[:, OFF_PHI:OFF_PHI + 8] = off_phi + 8 * cidx[:, None] + np.arange(8)[None]
    assert cd.max() < 2**31
    sp = Spaces(mesh=mesh, nfree={"u": n_u_free, "p": n_p, "j": n_j_free, "phi": n_phi},
                ndir={"u": n_u_dir, "j": n_j_dir}, cell_dofs=cd.astype(np.int32), rt_sign=sign.astype(np.int8),
                dir_vals=np.zeros(n_u_dir + n_j_dir), u_node_rank=node_of_local, u_node_free=node_free,
                u_node_dof=u_node_dof, face_ids=face_id, face_dirichlet=face_dir, face_dof=face_dof,
                cell_verts=vid.astype(np.int32), style=style,
                bc_spec=dict(u_tags=u_tags, j_tags=j_tags, thin_walls=list(thin_walls)),
                node_tagger=tagged_nodes, face_tagger=tagged_faces)
    for tw in thin_walls:
        ents = mesh.entities(tw["domain"])
        sel = sorted((c, f) for c, f, e in bq if e in ents)
        fc = np.array([c for c, _ in sel], dtype=np.int32)
        fl = np.array([f for _, f in sel], dtype=np.int8)
        sp.thin_wall_faces.append(dict(cell=fc, lface=fl, tau=float(tw["tau"]), cw=float(tw["cw"]),
                                       jw=float(tw.get("jw", 0.0))))
    return sp


def skeleton_faces(sp: Spaces) -> dict:
    """Interior faces of the mesh, the skeleton of the skeleton stabilisation
    (``Skeleton(Triangulation(model, tags=fluid))``, src/geometry.jl:138-140;
    every cell is fluid here).  A face shared by two cells, periodic images
    included, joins its first appearance (cell order, local face order) with
    the second: ``cell`` / ``lface`` (nf, 2), and ``perm`` (nf, 9) the 3x3 face
    Gauss point (lower tangential axis fastest) of the second cell at each point
    of the first.  The points are matched through the face's corner vertices
    (topological ids, so periodic faces match too)."""
    cached = sp.__dict__.get("_skeleton")
    if cached is not None:
        return cached
    fid = sp.face_ids.ravel()
    order = np.argsort(fid, kind="stable")               # (cell, local face) pairs by face, first appearance first
    f_sorted = fid[order]
    starts = np.nonzero(np.r_[True, f_sorted[1:] != f_sorted[:-1]])[0]
    counts = np.diff(np.r_[starts, f_sorted.size])
    if (counts > 2).any():
        raise ValueError("a face is shared by more than two cells")
    two = starts[counts == 2]
    first, second = order[two], order[two + 1]
    keep = np.argsort(first, kind="stable")
    first, second = first[keep], second[keep]
    cell = np.stack([first // 6, second // 6], 1)
    lface = np.stack([first % 6, second % 6], 1)
    # corners (i0, i1) of each side in face coordinates: local vertex side<<d | i0<<ta | i1<<tb
    corner_v = np.zeros((6, 4), dtype=np.int64)
    for f in range(6):
        d, (ta, tb) = FACE_AXIS[f], _tangential_axes(FACE_AXIS[f])
        for k in range(4):
            corner_v[f, k] = (FACE_SIDE[f] << d) | ((k & 1) << ta) | ((k >> 1) << tb)
    verts = sp.cell_verts
    P = np.take_along_axis(verts[cell[:, 0]], corner_v[lface[:, 0]], 1)      # (nf, 4)
    M = np.take_along_axis(verts[cell[:, 1]], corner_v[lface[:, 1]], 1)
    match = P[:, :, None] == M[:, None, :]                                   # plus corner k -> minus corner m
    if not (match.sum(2) == 1).all():
        raise ValueError("the two cells of an interior face do not share its four corners")
    m = np.argmax(match, 2)
    A, B, C = m[:, 0], m[:, 1], m[:, 2]                 # minus corners of plus (0,0), (1,0), (0,1)
    bit = lambda v, k: (v >> k) & 1
    qa = np.arange(9) % 3
    qb = np.arange(9) // 3
    s0 = 2 * bit(A, 0)[:, None] + (bit(B, 0) - bit(A, 0))[:, None] * qa + (bit(C, 0) - bit(A, 0))[:, None] * qb
    s1 = 2 * bit(A, 1)[:, None] + (bit(B, 1) - bit(A, 1))[:, None] * qa + (bit(C, 1) - bit(A, 1))[:, None] * qb
    perm = s0 + 3 * s1
    assert (np.sort(perm, 1) == np.arange(9)).all()
    out = dict(cell=cell.astype(np.int32), lface=lface.astype(np.int8), perm=perm.astype(np.int8))
    sp.__dict__["_skeleton"] = out
    return out


# ---------------------------------------------------------------------------
# Geometry helpers (Q1 map) and interpolation

def _q1_shape(xi: np.ndarray):
    """Trilinear shape functions and derivatives at reference points (N,3)."""
    N = np.empty((xi.shape[0], 8))
    dN = np.empty((xi.shape[0], 8, 3))
    for v in range(8):
        b = [(v >> a) & 1 for a in range(3)]
        f = [xi[:, a] if b[a] else 1.0 - xi[:, a] for a in range(3)]
        df = [np.ones(xi.shape[0]) if b[a] else -np.ones(xi.shape[0]) for a in range(3)]
        N[:, v] = f[0] * f[1] * f[2]
        dN[:, v, 0] = df[0] * f[1] * f[2]
        dN[:, v, 1] = f[0] * df[1] * f[2]
        dN[:, v, 2] = f[0] * f[1] * df[2]
    return N, dN


def map_points(cell_xyz: np.ndarray, xi: np.ndarray):
    """Physical points and Jacobians J[i][l] = dx_i/dxi_l.

    cell_xyz (C, 8, 3); xi (P, 3) -> x (C, P, 3), J (C, P, 3, 3)."""
    N, dN = _q1_shape(xi)
    x = np.einsum("pv,cvi->cpi", N, cell_xyz)
    J = np.einsum("pvl,cvi->cpil", dN, cell_xyz)
    return x, J


def interpolate_u(sp: Spaces, fun: Callable, free_only=False):
    """Nodal values of a vector field at every Q2 node -> (free_vals, dir_vals)."""
    xyz = sp.mesh.cell_xyz()
    xi = Q2_OFFSETS / 2.0
    x, _ = map_points(xyz, xi)                     # (C, 27, 3)
    vals = np.asarray(fun(x.reshape(-1, 3)), dtype=np.float64).reshape(-1, 27, 3)
    free = np.zeros(sp.nfree["u"])
    dirv = np.zeros(sp.ndir["u"])
    nd = sp.u_node_dof[sp.u_node_rank]            # (C, 27)
    for c in range(3):
        m = nd >= 0
        free[(nd[m] + c)] = vals[..., c][m]
        dirv[(-nd[~m] - 1) + c] = vals[..., c][~m]
    return free, dirv


def _rt_dof_points():
    """Reference point, axis of every local RT dof (36)."""
    pts = np.zeros((36, 3))
    axis = np.zeros(36, dtype=np.int64)
    for r in range(36):
        if r < 24:
            f, k = divmod(r, 4)
            d = 2 - f // 2
            m = 2 if f % 2 else 0
        else:
            d, k = divmod(r - 24, 4)
            m = 1
        t1, t2 = _tangential_axes(d)
        pts[r, d] = m / 2.0
        pts[r, t1] = G2[k % 2]
        pts[r, t2] = G2[k // 2]
        axis[r] = d
    return pts, axis


def interpolate_j(sp: Spaces, fun: Callable):
    """RT dof values of a vector field (inverse Piola at the dof points)."""
    xyz = sp.mesh.cell_xyz()
    pts, axis = _rt_dof_points()
    x, J = map_points(xyz, pts)                    # (C, 36, 3), (C, 36, 3, 3)
    jv = np.asarray(fun(x.reshape(-1, 3)), dtype=np.float64).reshape(-1, 36, 3)
    det = np.linalg.det(J)
    Ji = np.linalg.inv(J)
    jhat = det[..., None] * np.einsum("cpli,cpi->cpl", Ji, jv)  # det J^-1 j
    vals = jhat[:, np.arange(36), axis] * 1.0
    sign = np.ones((sp.ncells, 36))
    sign[:, :24] = np.repeat(sp.rt_sign.astype(np.float64), 4, axis=1)
    vals = vals * sign
    free = np.zeros(sp.nfree["j"])
    dirv = np.zeros(sp.ndir["j"])
    off_j = sp.offsets["j"]
    cd = sp.cell_dofs[:, OFF_J:OFF_J + 36].astype(np.int64)
    m = cd >= 0
    free[cd[m] - off_j] = vals[m]
    dirv[(-cd[~m] - 1) - sp.ndir["u"]] = vals[~m]
    return free, dirv


def interpolate_p(sp: Spaces, fun: Callable):
    """dP1 values: nodal at the reference points (0,0,0), (1,0,0), (0,1,0), (0,0,1)."""
    xyz = sp.mesh.cell_xyz()
    pts = np.array([[0.0, 0, 0], [1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]])
    x, _ = map_points(xyz, pts)
    return np.asarray(fun(x.reshape(-1, 3)), dtype=np.float64).reshape(-1)


def interpolate_phi(sp: Spaces, fun: Callable):
    """dQ1 values at the 8 cell vertices."""
    xyz = sp.mesh.cell_xyz()
    return np.asarray(fun(xyz.reshape(-1, 3)), dtype=np.float64).reshape(-1)


def dirichlet_values_by_tag(sp: Spaces, field: str, tags, fns) -> np.ndarray:
    """Dirichlet values of ``field`` ("u" or "j") with one function per tag
    (``TrialFESpace(V, [f1, f2, ...])`` over ``dirichlet_tags = [t1, t2, ...]``,
    src/fespaces.jl:116-124): a Dirichlet dof takes the function of the first
    tag of the list that holds it."""
    if isinstance(tags, str):
        tags = [tags]
    if len(tags) != len(fns):
        raise ValueError(f"{len(fns)} values for {len(tags)} Dirichlet tags of {field}")
    interp = interpolate_u if field == "u" else interpolate_j
    vals = [interp(sp, f)[1] for f in fns]
    out = np.zeros(sp.ndir[field])
    done = np.zeros(sp.ndir[field], dtype=bool)
    if field == "u":
        owner = sp.u_node_dof
        mask_of = sp.node_tagger
        width = 3
    else:
        owner = sp.face_dof
        mask_of = sp.face_tagger
        width = 4
    dirs = np.nonzero(owner < 0)[0]
    for t, v in zip(tags, vals):
        m = mask_of(t)[dirs]
        first = -owner[dirs[m]] - 1
        idx = (first[:, None] + np.arange(width)[None]).ravel()
        take = ~done[idx]
        out[idx[take]] = v[idx[take]]
        done[idx] = True
    return out


def interpolate(sp: Spaces, u=None, p=None, j=None, phi=None):
    """Multifield interpolation -> (free vector, Dirichlet vector)."""
    x = np.zeros(sp.ndofs)
    dv = np.zeros(sp.ndir["u"] + sp.ndir["j"])
    sl = sp.field_slices()
    if u is not None:
        fu, du = interpolate_u(sp, u)
        x[sl["u"]] = fu
        dv[:sp.ndir["u"]] = du
    if p is not None:
        x[sl["p"]] = interpolate_p(sp, p)
    if j is not None:
        fj, dj = interpolate_j(sp, j)
        x[sl["j"]] = fj
        dv[sp.ndir["u"]:] = dj
    if phi is not None:
        x[sl["phi"]] = interpolate_phi(sp, phi)
    return x, dv
