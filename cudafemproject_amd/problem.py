"""Problem container, structured mesh generators and the prepared-mesh format.

The prepared-mesh directory format is the one the reference's dataKeeper reads
(DataKeeper/src/datakeeper.cpp:50-213):

* ``nodes.txt``        N, then N lines ``x y [z]``
* ``elements.txt``     M, then M lines of dim+1 0-based node ids
* ``constraints.txt``  C, then C lines ``node code`` (code 1..7, decoded bitwise)
* ``loads.txt``        L, then L lines ``node fx fy [fz]`` (a lone ``0`` when empty)
* ``stress.txt``       B, then B lines ``n_1 .. n_dim n_interior adjElem nx ny [nz] p``
  (the normal is normalised on read).  The legacy golden fixtures use the
  older 2D form ``n1 n2 adjElem nx ny p``; ``read_prepared`` accepts both.
"""
from __future__ import annotations

import dataclasses
import os
from typing import Optional

import numpy as np

LOADS_REFERENCE, LOADS_PHYSICAL = 0, 1


@dataclasses.dataclass
class Problem:
    nodes: np.ndarray                 # (n_nodes, dim) float64
    elems: np.ndarray                 # (n_elems, dim+1) int32
    constr_node: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    constr_code: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    face_nodes: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 0), np.int32))
    face_adj: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    face_normal: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 0)))
    face_p: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0))
    face_interior: Optional[np.ndarray] = None
    load_node: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros(0, np.int32))
    load_vec: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 0)))
    E: float = 2e7
    nu: float = 0.3
    elem_scale: Optional[np.ndarray] = None
    convention: int = LOADS_REFERENCE
    extra_rhs: Optional[np.ndarray] = None

    @property
    def dim(self) -> int:
        return int(self.nodes.shape[1])

    @property
    def n_nodes(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def n_elems(self) -> int:
        return int(self.elems.shape[0])

    @property
    def n_dof(self) -> int:
        return self.n_nodes * self.dim

    def fixed_dofs(self) -> np.ndarray:
        """Boolean (n_nodes, dim): DOFs in the reference's constraintsIds
        (bitwise decode, datakeeper.cpp:148-164)."""
        fx = np.zeros((self.n_nodes, self.dim), bool)
        for n, t in zip(self.constr_node, self.constr_code):
            if t & 1:
                fx[n, 0] = True
            if t & 2:
                fx[n, 1] = True
            if self.dim == 3 and (t & 3):
                fx[n, 2] = True
        return fx


# --------------------------------------------------------------------------
# structured generators

def plate_tri(nx: int, ny: int, lx: float = 1.0, ly: float = 0.5):
    """Rectangle [0,lx]x[0,ly] on an nx x ny grid, each quad split into two
    counter-clockwise triangles along the (i,j)-(i+1,j+1) diagonal.
    Returns nodes, elems; node id = i + (nx+1)*j."""
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    i, j = i.ravel(), j.ravel()
    n00 = i + (nx + 1) * j
    n10 = n00 + 1
    n01 = n00 + (nx + 1)
    n11 = n01 + 1
    t1 = np.stack([n00, n10, n11], 1)
    t2 = np.stack([n00, n11, n01], 1)
    elems = np.stack([t1, t2], 1).reshape(-1, 3).astype(np.int32)
    return nodes, elems


_KUHN_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def box_tet(nx: int, ny: int, nz: int, lx: float = 1.0, ly: float = 1.0, lz: float = 1.0):
    """Box [0,lx]x[0,ly]x[0,lz] on an nx x ny x nz hex grid, each hex split into
    the 6 Kuhn tetrahedra along its main diagonal (conforming, equal volumes,
    positively oriented).  Returns nodes, elems; node id = i + (nx+1)*(j + (ny+1)*k);
    elements hex-major, 6 per hex in the order of _KUHN_PERMS."""
    xs = np.linspace(0.0, lx, nx + 1)
    ys = np.linspace(0.0, ly, ny + 1)
    zs = np.linspace(0.0, lz, nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    sx, sy, sz = 1, nx + 1, (nx + 1) * (ny + 1)
    step = np.array([sx, sy, sz])
    base = i * sx + j * sy + k * sz
    tets = []
    for p in _KUHN_PERMS:
        v0 = base
        v1 = v0 + step[p[0]]
        v2 = v1 + step[p[1]]
        v3 = v2 + step[p[2]]
        # det[e_a; e_b; e_c] = sign of the permutation; odd ones swap v2, v3
        even = p in [(0, 1, 2), (1, 2, 0), (2, 0, 1)]
        tets.append(np.stack([v0, v1, v2, v3] if even else [v0, v1, v3, v2], 1))
    elems = np.stack(tets, 1).reshape(-1, 4).astype(np.int32)
    return nodes, elems


def box_face_tris(nx, ny, nz, axis: int, side: int):
    """Boundary triangles of box_tet on the face axis=const (side 0: min, 1: max).
    Returns face_nodes (n,3), adj_elem (n,), interior node (n,)."""
    dims = [nx, ny, nz]
    sx, sy, sz = 1, nx + 1, (nx + 1) * (ny + 1)
    step = [sx, sy, sz]
    rng = [np.arange(d) for d in dims]
    rng[axis] = np.array([dims[axis] - 1 if side == 1 else 0])
    k, j, i = np.meshgrid(rng[2], rng[1], rng[0], indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    hexid = i + nx * (j + ny * k)
    base = i * sx + j * sy + k * sz
    fn, fa, fi = [], [], []
    for t, p in enumerate(_KUHN_PERMS):
        v0 = base
        v1 = v0 + step[p[0]]
        v2 = v1 + step[p[1]]
        v3 = v2 + step[p[2]]
        if side == 1 and p[0] == axis:
            # v1, v2, v3 lie on the max face; v0 is the interior node
            fn.append(np.stack([v1, v2, v3], 1)); fa.append(hexid * 6 + t); fi.append(v0)
        if side == 0 and p[2] == axis:
            # v0, v1, v2 lie on the min face; v3 is the interior node
            fn.append(np.stack([v0, v1, v2], 1)); fa.append(hexid * 6 + t); fi.append(v3)
    return (np.concatenate(fn).astype(np.int32), np.concatenate(fa).astype(np.int32),
            np.concatenate(fi).astype(np.int32))


def outward_normals(nodes, face_nodes, interior):
    """Unnormalised normals oriented away from the interior node
    (the orientation rule of the prepared stress.txt)."""
    p0 = nodes[face_nodes[:, 0]]
    if face_nodes.shape[1] == 2:
        t = nodes[face_nodes[:, 1]] - p0
        n = np.stack([t[:, 1], -t[:, 0]], 1)
    else:
        n = np.cross(nodes[face_nodes[:, 1]] - p0, nodes[face_nodes[:, 2]] - p0)
    w = nodes[interior] - p0
    flip = (n * w).sum(1) > 0
    n[flip] *= -1
    return n


def nodes_on_plane(nodes, axis, value, tol=1e-12):
    return np.nonzero(np.abs(nodes[:, axis] - value) <= tol)[0].astype(np.int32)


# --------------------------------------------------------------------------
# the benchmark workloads (sizes as stated in the benchmark configuration)

def plate_problem(nx=64, ny=32, lx=1.0, ly=0.5, E=2e7, nu=0.25, traction=1e3,
                  convention=LOADS_REFERENCE) -> Problem:
    """2D plane-stress plate: left edge fixed (code 3), uniform x-traction on
    the right edge.  In the reference's pressure convention F = -1/2 p L n with
    the outward normal +x, an outward traction t is the pressure p = -t."""
    nodes, elems = plate_tri(nx, ny, lx, ly)
    left = nodes_on_plane(nodes, 0, 0.0)
    j = np.arange(ny)
    a = nx + (nx + 1) * j          # (nx, j)
    b = a + (nx + 1)               # (nx, j+1)
    quad = (nx - 1) + nx * j
    adj = (2 * quad).astype(np.int32)   # t1 of quad (nx-1, j) holds (nx,j),(nx,j+1)
    interior = (a - 1).astype(np.int32)  # (nx-1, j)
    fn = np.stack([a, b], 1).astype(np.int32)
    normals = outward_normals(nodes, fn, interior)
    return Problem(nodes=nodes, elems=elems,
                   constr_node=left, constr_code=np.full(len(left), 3, np.int32),
                   face_nodes=fn, face_adj=adj, face_normal=normals,
                   face_p=np.full(len(adj), -traction), face_interior=interior,
                   E=E, nu=nu, convention=convention)


def cube_problem(n=64, E=2e7, nu=0.25, pressure=1e3, convention=LOADS_PHYSICAL) -> Problem:
    """3D unit cube, n^3 hexes x 6 Kuhn tets; z=0 face clamped (code 7),
    uniform pressure on the z=1 face."""
    nodes, elems = box_tet(n, n, n)
    bot = nodes_on_plane(nodes, 2, 0.0)
    fn, fa, fi = box_face_tris(n, n, n, 2, 1)
    normals = outward_normals(nodes, fn, fi)
    return Problem(nodes=nodes, elems=elems,
                   constr_node=bot, constr_code=np.full(len(bot), 7, np.int32),
                   face_nodes=fn, face_adj=fa, face_normal=normals,
                   face_p=np.full(len(fa), pressure), face_interior=fi,
                   E=E, nu=nu, convention=convention)


def column_problem(nx=16, ny=16, nz=256, lx=0.1, ly=0.1, lz=1.6, E=2e7, nu=0.25,
                   pressure=1.0, convention=LOADS_PHYSICAL) -> Problem:
    """Quasi-1D column: bottom clamped, unit pressure on the top face (the
    dynamics driver scales it with the wavelet)."""
    nodes, elems = box_tet(nx, ny, nz, lx, ly, lz)
    bot = nodes_on_plane(nodes, 2, 0.0)
    fn, fa, fi = box_face_tris(nx, ny, nz, 2, 1)
    normals = outward_normals(nodes, fn, fi)
    return Problem(nodes=nodes, elems=elems,
                   constr_node=bot, constr_code=np.full(len(bot), 7, np.int32),
                   face_nodes=fn, face_adj=fa, face_normal=normals,
                   face_p=np.full(len(fa), pressure), face_interior=fi,
                   E=E, nu=nu, convention=convention)


def cantilever_problem(nx=96, ny=32, nz=32, lx=3.0, ly=1.0, lz=1.0, E0=2e7, nu=0.3,
                       seed=5, patch=0.125, pressure=1e3, convention=LOADS_PHYSICAL) -> Problem:
    """Cantilever: x=0 clamped, downward pressure on the faces of the x=lx end
    whose centroid has z > lz*(1-patch) (a strip at the top of the tip);
    per-element density rho_e ~ U[0.05, 1] (seeded), modulus E0 rho_e^3."""
    nodes, elems = box_tet(nx, ny, nz, lx, ly, lz)
    left = nodes_on_plane(nodes, 0, 0.0)
    fn, fa, fi = box_face_tris(nx, ny, nz, 0, 1)
    cz = nodes[fn][:, :, 2].mean(1)
    sel = cz > lz * (1.0 - patch)
    fn, fa, fi = fn[sel], fa[sel], fi[sel]
    # downward traction applied as a pressure on a face whose outward normal
    # is +x: use the normal -z directly (pressure p along -normal => +z), so
    # give normal +z and pressure +p: F = -p A n = -p A e_z.
    normals = np.tile(np.array([[0.0, 0.0, 1.0]]), (len(fa), 1))
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.05, 1.0, size=elems.shape[0])
    pb = Problem(nodes=nodes, elems=elems,
                 constr_node=left, constr_code=np.full(len(left), 7, np.int32),
                 face_nodes=fn, face_adj=fa, face_normal=normals,
                 face_p=np.full(len(fa), pressure), face_interior=fi,
                 E=E0, nu=nu, elem_scale=rho ** 3, convention=convention)
    pb.rho = rho
    return pb


@dataclasses.dataclass
class BatchInstance:
    """One plate of a batch: mesh, material, constraints and load cases
    f (n_nodes, n_rhs, dim), assembled consistent nodal loads."""
    nodes: np.ndarray
    elems: np.ndarray
    E: float
    nu: float
    constr_node: np.ndarray
    constr_code: np.ndarray
    f: np.ndarray

    def problem(self) -> Problem:
        """The instance as a single Problem (for per-instance reference solves)."""
        return Problem(nodes=self.nodes, elems=self.elems, constr_node=self.constr_node,
                       constr_code=self.constr_code, E=self.E, nu=self.nu)


def _edge_load(nodes, on_edge, axis_along, comp, t, f):
    """Consistent nodal loads of a uniform traction t (component `comp`) on the
    straight edge made of the nodes `on_edge`: t * h / 2 to both ends of every
    segment."""
    idx = on_edge[np.argsort(nodes[on_edge, axis_along])]
    seg = np.diff(nodes[idx, axis_along])
    f[idx[:-1], comp] += t * seg / 2.0
    f[idx[1:], comp] += t * seg / 2.0


def plate_batch(n=512, seed=2024, nx_range=(16, 160), ny_range=(16, 80), lx=1.0, ly=0.5,
                traction=1e3, body=1e3):
    """Batch of 2D plates with seeded uneven grids (nx ~ U{nx_range},
    ny ~ U{ny_range}), E ~ LogNormal(ln 2e7, 0.5), nu ~ U[0.2, 0.35]; left edge
    fixed (code 3); four load cases: x-traction and y-traction on the right
    edge, shear (x-traction) on the top edge, body force -body per unit area
    in y (rho g A / 3 to each node).  Loads on fixed DOFs are zero."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        E = float(rng.lognormal(np.log(2e7), 0.5))
        nu = float(rng.uniform(0.2, 0.35))
        nodes, elems = plate_tri(nx, ny, lx, ly)
        nn = len(nodes)
        left = nodes_on_plane(nodes, 0, 0.0)
        right = nodes_on_plane(nodes, 0, lx)
        top = nodes_on_plane(nodes, 1, ly)
        f = np.zeros((nn, 4, 2))
        fx, fy, fs, fb = (np.zeros((nn, 2)) for _ in range(4))
        _edge_load(nodes, right, 1, 0, traction, fx)
        _edge_load(nodes, right, 1, 1, traction, fy)
        _edge_load(nodes, top, 0, 0, traction, fs)
        X = nodes[elems]
        A = 0.5 * np.abs((X[:, 1, 0] - X[:, 0, 0]) * (X[:, 2, 1] - X[:, 0, 1])
                         - (X[:, 1, 1] - X[:, 0, 1]) * (X[:, 2, 0] - X[:, 0, 0]))
        np.add.at(fb[:, 1], elems.ravel(), np.repeat(-body * A / 3.0, 3))
        for k, fk in enumerate((fx, fy, fs, fb)):
            f[:, k] = fk
        f[left] = 0.0
        out.append(BatchInstance(nodes=nodes, elems=elems, E=E, nu=nu, constr_node=left,
                                 constr_code=np.full(len(left), 3, np.int32), f=f))
    return out


def load_batch(ctx, insts):
    """Load a list of BatchInstance into a Context's batch."""
    ctx.batch_set_mesh([(b.nodes, b.elems) for b in insts])
    ctx.batch_set_material([b.E for b in insts], [b.nu for b in insts])
    ctx.batch_set_constraints([(b.constr_node, b.constr_code) for b in insts])


# --------------------------------------------------------------------------
# prepared-mesh files

def _read_counted(path, ncols, dtype=float):
    with open(path) as f:
        toks = f.read().split()
    if not toks:
        return np.zeros((0, ncols), dtype)
    n = int(float(toks[0]))
    vals = toks[1:1 + n * ncols]
    if n == 0:
        return np.zeros((0, ncols), dtype)
    return np.array([float(v) for v in vals], dtype=np.float64).reshape(n, ncols).astype(dtype)


def read_prepared(mesh_dir: str, dim: int, E: float = 2e7, nu: float = 0.3) -> Problem:
    """Read a prepared-mesh directory (both stress.txt layouts)."""
    nodes = _read_counted(os.path.join(mesh_dir, "nodes.txt"), dim)
    elems = _read_counted(os.path.join(mesh_dir, "elements.txt"), dim + 1).astype(np.int32)
    cons = _read_counted(os.path.join(mesh_dir, "constraints.txt"), 2).astype(np.int32)
    loads = _read_counted(os.path.join(mesh_dir, "loads.txt"), dim + 1)
    with open(os.path.join(mesh_dir, "stress.txt")) as f:
        lines = [ln.split() for ln in f.read().strip().splitlines()]
    nf = int(float(lines[0][0])) if lines else 0
    rows = [r for r in lines[1:1 + nf]]
    fn, fa, fnorm, fp, fi = [], [], [], [], []
    for r in rows:
        v = [float(x) for x in r]
        if len(v) == 2 * dim + 3:        # n_1..n_dim n_interior adj normal p
            fn.append(v[:dim]); fi.append(v[dim]); fa.append(v[dim + 1])
            fnorm.append(v[dim + 2:2 * dim + 2]); fp.append(v[2 * dim + 2])
        elif len(v) == 2 * dim + 2:      # legacy: n_1..n_dim adj normal p
            fn.append(v[:dim]); fi.append(-1); fa.append(v[dim])
            fnorm.append(v[dim + 1:2 * dim + 1]); fp.append(v[2 * dim + 1])
        else:
            raise ValueError(f"stress.txt line with {len(v)} fields for dim {dim}")
    pb = Problem(nodes=nodes.astype(np.float64), elems=elems,
                 constr_node=cons[:, 0].copy() if len(cons) else np.zeros(0, np.int32),
                 constr_code=cons[:, 1].copy() if len(cons) else np.zeros(0, np.int32),
                 face_nodes=np.array(fn, np.int32).reshape(-1, dim),
                 face_adj=np.array(fa, np.int32),
                 face_normal=np.array(fnorm, np.float64).reshape(-1, dim),
                 face_p=np.array(fp, np.float64),
                 face_interior=np.array(fi, np.int32),
                 load_node=loads[:, 0].astype(np.int32) if len(loads) else np.zeros(0, np.int32),
                 load_vec=loads[:, 1:].copy() if len(loads) else np.zeros((0, dim)),
                 E=E, nu=nu)
    return pb


def write_prepared(pb: Problem, mesh_dir: str):
    """Write the current (dataKeeper) prepared-mesh format."""
    os.makedirs(mesh_dir, exist_ok=True)
    d = pb.dim
    with open(os.path.join(mesh_dir, "nodes.txt"), "w") as f:
        f.write(f"{pb.n_nodes}\n")
        np.savetxt(f, pb.nodes, fmt="%.17g")
    with open(os.path.join(mesh_dir, "elements.txt"), "w") as f:
        f.write(f"{pb.n_elems}\n")
        np.savetxt(f, pb.elems, fmt="%d")
    with open(os.path.join(mesh_dir, "constraints.txt"), "w") as f:
        f.write(f"{len(pb.constr_node)}\n")
        for n, c in zip(pb.constr_node, pb.constr_code):
            f.write(f"{n} {c}\n")
    with open(os.path.join(mesh_dir, "loads.txt"), "w") as f:
        f.write(f"{len(pb.load_node)}\n")
        for n, v in zip(pb.load_node, pb.load_vec):
            f.write(f"{n} " + " ".join(f"{x:.17g}" for x in v) + "\n")
    interior = pb.face_interior
    if interior is None or (len(interior) and np.any(interior < 0)):
        interior = _interior_nodes(pb)
    with open(os.path.join(mesh_dir, "stress.txt"), "w") as f:
        f.write(f"{len(pb.face_adj)}\n")
        for k in range(len(pb.face_adj)):
            f.write(" ".join(str(int(x)) for x in pb.face_nodes[k]) +
                    f" {int(interior[k])} {int(pb.face_adj[k])} " +
                    " ".join(f"{x:.17g}" for x in pb.face_normal[k]) +
                    f" {pb.face_p[k]:.17g}\n")


def _interior_nodes(pb: Problem) -> np.ndarray:
    out = np.zeros(len(pb.face_adj), np.int32)
    for k, (fn, e) in enumerate(zip(pb.face_nodes, pb.face_adj)):
        rest = [v for v in pb.elems[e] if v not in set(fn.tolist())]
        out[k] = rest[0]
    return out
