"""Cartesian hexahedral duct meshes with Gridap's box face labelling.

Mirrors the reference's mesher layer:
  * ``hunt_generate_base_mesh`` / ``CartesianDiscreteModel(domain, nc;
    isperiodic, map)``  (src/Meshers/hunt_mesher.jl:128-138)
  * ``strechMHD`` / ``hunt_stretch_map``  (src/Meshers/hunt_mesher.jl:5-46)
  * ``hunt_add_tags!``  (src/Meshers/hunt_mesher.jl:94-101)

Conventions (Gridap's, as used by the reference's tags):
  * cells and vertices are lexicographic, x fastest;
  * local vertex ``v = i + 2j + 4k`` of a hex;
  * box entities: vertices 1-8, edges 9-20, faces 21-26 in the order
    z-, z+, y-, y+, x-, x+, interior 27 (so ``"insulating" = [25, 26]`` are the
    x = +-1 side walls and ``[23, 24]`` the y = +-1 Hartmann walls);
  * a periodic direction identifies its last vertex plane with the first, the
    geometry keeps the true coordinates (cells of the last layer reach z = L).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, Dict, Iterable, List, Optional, Sequence

import numpy as np

# Gridap HEX local edges (1-based vertex pairs) and faces, as lattice offsets
# of their mid-point in a 3x3x3 lattice of one cell.
EDGE_OFFSETS = np.array([
    (1, 0, 0), (1, 2, 0), (1, 0, 2), (1, 2, 2),   # x-directed: (1,2) (3,4) (5,6) (7,8)
    (0, 1, 0), (2, 1, 0), (0, 1, 2), (2, 1, 2),   # y-directed: (1,3) (2,4) (5,7) (6,8)
    (0, 0, 1), (2, 0, 1), (0, 2, 1), (2, 2, 1),   # z-directed: (1,5) (2,6) (3,7) (4,8)
], dtype=np.int64)
FACE_OFFSETS = np.array([
    (1, 1, 0), (1, 1, 2),   # z-, z+
    (1, 0, 1), (1, 2, 1),   # y-, y+
    (0, 1, 1), (2, 1, 1),   # x-, x+
], dtype=np.int64)
FACE_AXIS = np.array([2, 2, 1, 1, 0, 0])
FACE_SIDE = np.array([0, 1, 0, 1, 0, 1])

# Box entity ids of Gridap's CartesianDiscreteModel labelling.
_BOX_VERTEX = {(i, j, k): 1 + i + 2 * j + 4 * k for i in (0, 1) for j in (0, 1) for k in (0, 1)}
_BOX_EDGE = {}
for _e, (_a, _b) in enumerate([(1, 2), (3, 4), (5, 6), (7, 8), (1, 3), (2, 4), (5, 7), (6, 8),
                               (1, 5), (2, 6), (3, 7), (4, 8)]):
    _BOX_EDGE[frozenset((_a, _b))] = 9 + _e
_BOX_FACE = {(2, 0): 21, (2, 1): 22, (1, 0): 23, (1, 1): 24, (0, 0): 25, (0, 1): 26}


def box_entity(walls: Sequence[Optional[int]]) -> int:
    """Box entity id of a point given, per axis, the wall it sits on (0, 1) or None."""
    on = [a for a in range(3) if walls[a] is not None]
    if len(on) == 0:
        return 27
    if len(on) == 1:
        a = on[0]
        return _BOX_FACE[(a, walls[a])]
    if len(on) == 3:
        return _BOX_VERTEX[tuple(walls)]
    # edge: the two box vertices spanning the free axis
    free = [a for a in range(3) if walls[a] is None][0]
    vs = []
    for t in (0, 1):
        w = list(walls)
        w[free] = t
        vs.append(_BOX_VERTEX[tuple(w)])
    return _BOX_EDGE[frozenset(vs)]


@dataclasses.dataclass
class CartesianMesh:
    """A structured hex mesh of a box, possibly periodic, with mapped nodes."""

    nc: tuple            # (nx, ny, nz)
    domain: tuple        # (x0, x1, y0, y1, z0, z1)
    isperiodic: tuple    # (px, py, pz)
    node_xyz: np.ndarray  # ((nx+1)(ny+1)(nz+1), 3) geometric nodes (after map)
    tags: Dict[str, List[int]] = dataclasses.field(default_factory=dict)

    @property
    def ncells(self) -> int:
        nx, ny, nz = self.nc
        return nx * ny * nz

    def cell_ijk(self) -> np.ndarray:
        nx, ny, nz = self.nc
        c = np.arange(self.ncells, dtype=np.int64)
        return np.stack([c % nx, (c // nx) % ny, c // (nx * ny)], axis=1)

    def cell_nodes(self) -> np.ndarray:
        """(ncells, 8) geometric node ids, local vertex v = i + 2j + 4k."""
        nx, ny, nz = self.nc
        ijk = self.cell_ijk()
        out = np.empty((self.ncells, 8), dtype=np.int64)
        for v in range(8):
            di, dj, dk = v & 1, (v >> 1) & 1, (v >> 2) & 1
            out[:, v] = (ijk[:, 0] + di) + (nx + 1) * ((ijk[:, 1] + dj) + (ny + 1) * (ijk[:, 2] + dk))
        return out

    def cell_xyz(self) -> np.ndarray:
        """(ncells, 8, 3) vertex coordinates of every cell."""
        return self.node_xyz[self.cell_nodes()]

    def lattice_shape(self) -> tuple:
        """Shape of the Q2 node lattice (2n+1 per axis, 2n for a periodic axis)."""
        return tuple(2 * n if p else 2 * n + 1 for n, p in zip(self.nc, self.isperiodic))

    def add_tag_from_tags(self, name: str, entities: Iterable[int]) -> None:
        """Gridap ``add_tag_from_tags!`` restricted to box entity ids."""
        self.tags[name] = sorted(set(int(e) for e in entities))

    def lattice_entity(self, I: np.ndarray, J: np.ndarray, K: np.ndarray) -> np.ndarray:
        """Box entity id (1..27) of Q2-lattice points (I, J, K)."""
        nx, ny, nz = self.nc
        coords = (I, J, K)
        n2 = (2 * nx, 2 * ny, 2 * nz)
        wall = []
        for a in range(3):
            if self.isperiodic[a]:
                wall.append(np.full(I.shape, -1, dtype=np.int64))
            else:
                w = np.full(I.shape, -1, dtype=np.int64)
                w[coords[a] == 0] = 0
                w[coords[a] == n2[a]] = 1
                wall.append(w)
        # encode (wx, wy, wz) in base 3 -> lookup
        table = np.zeros(27, dtype=np.int64)
        for wx in (-1, 0, 1):
            for wy in (-1, 0, 1):
                for wz in (-1, 0, 1):
                    key = (wx + 1) + 3 * (wy + 1) + 9 * (wz + 1)
                    table[key] = box_entity([None if w < 0 else w for w in (wx, wy, wz)])
        key = (wall[0] + 1) + 3 * (wall[1] + 1) + 9 * (wall[2] + 1)
        return table[key]

    def tagged(self, tags, I, J, K) -> np.ndarray:
        """Boolean mask: lattice points whose box entity is in the union of ``tags``."""
        ents = set()
        for t in _as_list(tags):
            if isinstance(t, str):
                ents.update(self.tags[t])
            else:
                ents.add(int(t))
        e = self.lattice_entity(I, J, K)
        return np.isin(e, np.array(sorted(ents), dtype=np.int64))


def _as_list(x):
    if x is None:
        return []
    if isinstance(x, (list, tuple)):
        return list(x)
    return [x]


def cartesian_model(domain, nc, isperiodic=(False, False, False), map: Optional[Callable] = None) -> CartesianMesh:
    """``CartesianDiscreteModel(domain, nc; isperiodic, map)`` on a 3-D box.

    ``map`` acts on an (N, 3) array of uniform node coordinates and returns the
    mapped coordinates (vectorised form of the reference's point map).
    """
    nx, ny, nz = nc
    x0, x1, y0, y1, z0, z1 = domain
    if isperiodic[2] and nz < 3:
        raise ValueError("a periodic direction needs at least 3 cells")
    xs = x0 + (x1 - x0) * np.arange(nx + 1) / nx
    ys = y0 + (y1 - y0) * np.arange(ny + 1) / ny
    zs = z0 + (z1 - z0) * np.arange(nz + 1) / nz
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    if map is not None:
        pts = np.asarray(map(pts), dtype=np.float64)
    return CartesianMesh(tuple(nc), tuple(domain), tuple(bool(p) for p in isperiodic), pts)


# ---------------------------------------------------------------------------
# Node maps

def strech_mhd(coord: np.ndarray, domain, factor, dirs) -> np.ndarray:
    """Smolentsev stretching (``strechMHD``, src/Meshers/hunt_mesher.jl:5-28)."""
    out = coord.copy()
    for i, d in enumerate(dirs):
        xi0, xi1 = domain[2 * i], domain[2 * i + 1]
        l = xi1 - xi0
        c = (factor[i] + 1.0) / (factor[i] - 1.0)
        x = coord[:, d]
        if l > 0:
            m = (xi0 <= x) & (x <= xi1)
        else:
            m = (xi1 <= x) & (x <= xi0)
        if l == 0:
            continue
        xx = (x[m] - xi0) / l
        st = factor[i] * (c ** xx - 1.0) / (1.0 + c ** xx)
        out[m, d] = st * l + xi0
    return out


def hunt_stretch_map(L: float, Ha: float, kmap_x: float = 1, kmap_y: float = 1, BL_adapted: bool = True):
    """``hunt_stretch_map`` (src/Meshers/hunt_mesher.jl:30-46), vectorised."""
    strech_Ha = math.sqrt(Ha / (Ha - 1.0)) if BL_adapted else None
    strech_side = math.sqrt(math.sqrt(Ha) / (math.sqrt(Ha) - 1.0)) if BL_adapted else None

    def map1(x):
        y = strech_mhd(x, (0.0, -L, 0.0, -L), (strech_side, strech_Ha), (0, 1))
        return strech_mhd(y, (0.0, L, 0.0, L), (strech_side, strech_Ha), (0, 1))

    def map2(x):
        out = x.copy()
        out[:, 0] = np.sign(x[:, 0]) * np.abs(L * x[:, 0]) ** (1.0 / kmap_x)
        out[:, 1] = np.sign(x[:, 1]) * np.abs(L * x[:, 1]) ** (1.0 / kmap_y)
        return out

    return map1 if BL_adapted else map2


def tanh_map(beta: float, axes=(1,)):
    """Node clustering towards both walls of [-1, 1]: ``x -> tanh(beta x)/tanh(beta)``."""
    def m(x):
        out = x.copy()
        for a in axes:
            out[:, a] = np.tanh(beta * x[:, a]) / math.tanh(beta)
        return out
    return m


def _splitmix64(n: int, seed: int) -> np.ndarray:
    """n consecutive SplitMix64 outputs, as uint64."""
    mask = (1 << 64) - 1
    out = np.empty(n, dtype=np.uint64)
    s = seed & mask
    for i in range(n):
        s = (s + 0x9E3779B97F4A7C15) & mask
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        out[i] = z ^ (z >> 31)
    return out


def jitter_nodes(mesh: CartesianMesh, amplitude: float, seed: int) -> None:
    """Seeded random displacement U[-a h, a h]^3 of interior nodes (in place).

    Only nodes strictly inside the box along every non-periodic axis move; a
    node of the periodic direction moves with its periodic image, so the
    identified planes stay congruent.
    """
    nx, ny, nz = mesh.nc
    x0, x1, y0, y1, z0, z1 = mesh.domain
    h = np.array([(x1 - x0) / nx, (y1 - y0) / ny, (z1 - z0) / nz])
    nzv = nz if mesh.isperiodic[2] else nz + 1
    nv = (nx + 1) * (ny + 1) * nzv
    r = _splitmix64(3 * nv, seed)
    u = (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    d = (2.0 * u - 1.0).reshape(nv, 3) * amplitude * h[None, :]
    idx = np.arange((nx + 1) * (ny + 1) * (nz + 1))
    i = idx % (nx + 1)
    j = (idx // (nx + 1)) % (ny + 1)
    k = idx // ((nx + 1) * (ny + 1))
    interior = (i > 0) & (i < nx) & (j > 0) & (j < ny)
    if not mesh.isperiodic[2]:
        interior &= (k > 0) & (k < nz)
    kk = k % nzv
    vid = i + (nx + 1) * (j + (ny + 1) * kk)
    mesh.node_xyz[interior] += d[vid[interior]]


# ---------------------------------------------------------------------------
# Reference application meshes

def hunt_add_tags(mesh: CartesianMesh) -> None:
    """``hunt_add_tags!`` without solid walls (src/Meshers/hunt_mesher.jl:94-101)."""
    mesh.add_tag_from_tags("noslip", list(range(1, 21)) + [23, 24, 25, 26])
    mesh.add_tag_from_tags("conducting", list(range(1, 13)) + list(range(17, 21)) + [23, 24])
    mesh.add_tag_from_tags("insulating", [25, 26])
    # wall groups used by the fully developed duct (src/Applications/FullyDevelopedFlow.jl:136-142)
    mesh.add_tag_from_tags("Ha_walls", list(range(1, 21)) + [23, 24])
    mesh.add_tag_from_tags("side_walls", list(range(1, 21)) + [25, 26])
    mesh.add_tag_from_tags("walls", list(range(1, 21)) + [23, 24, 25, 26])
    # duct ends of a non-periodic z-duct (channel-style inlet/outlet, src/Applications/channel.jl:251-271)
    mesh.add_tag_from_tags("inlet", [21])
    mesh.add_tag_from_tags("outlet", [22])
    mesh.add_tag_from_tags("ends", [21, 22])


def hunt_generate_base_mesh(nc, L=1.0, tw=0.0, Ha=10.0, kmap_x=1, kmap_y=1, BL_adapted=True,
                            nz=3, domain=(-1.0, 1.0, -1.0, 1.0, 0.0, 0.1), periodic_z=True,
                            map: Optional[Callable] = None) -> CartesianMesh:
    """Hunt duct mesh (src/Meshers/hunt_mesher.jl:128-138).

    ``nz`` and ``domain`` default to the reference's fixed (nx, ny, 3) on
    [-1,1]^2 x [0,0.1]; the benchmark ducts pass their own.  ``map`` overrides
    the reference stretch map when given.
    """
    if tw != 0.0:
        raise NotImplementedError("solid walls (tw > 0) are outside the H1-HDiv fluid hot path")
    coord_map = map if map is not None else hunt_stretch_map(L + tw, Ha, kmap_x, kmap_y, BL_adapted)
    mesh = cartesian_model(domain, (nc[0], nc[1], nz), (False, False, periodic_z), coord_map)
    hunt_add_tags(mesh)
    return mesh
