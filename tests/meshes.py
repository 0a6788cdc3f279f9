"""Small test meshes built through the host C-ABI (mhdg_mesh_from_cells)."""
import numpy as np

# Outward loops of a right-handed hexahedron with node q at (q&1, (q>>1)&1, (q>>2)&1); the loop table of the
# reference's hexahedron builder (src/uns/complex/ComplexBuilder.cpp:79-85), kept identical so that faces and
# edges are numbered the same way.
HEX_LOOPS = ((0, 4, 6, 2), (1, 3, 7, 5), (0, 1, 5, 4), (2, 6, 7, 3), (0, 2, 3, 1), (4, 5, 7, 6))


def box_cells(nx, ny, nz, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
    """Nodes, hexahedral cell loops and boundary tags (1..6 = xmin, xmax, ymin, ymax, zmin, zmax) of a box."""
    nid = lambda i, j, k: (k * (ny + 1) + j) * (nx + 1) + i
    xyz = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))
    for k in range(nz + 1):
        for j in range(ny + 1):
            for i in range(nx + 1):
                xyz[nid(i, j, k)] = [lo[0] + (hi[0] - lo[0]) * i / nx, lo[1] + (hi[1] - lo[1]) * j / ny,
                                     lo[2] + (hi[2] - lo[2]) * k / nz]
    cells, tags = [], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                v = [nid(i + (q & 1), j + ((q >> 1) & 1), k + ((q >> 2) & 1)) for q in range(8)]
                cells.append([[v[a] for a in lp] for lp in HEX_LOOPS])
                if i == 0: tags.append(([v[0], v[2], v[4], v[6]], 1))
                if i == nx - 1: tags.append(([v[1], v[3], v[5], v[7]], 2))
                if j == 0: tags.append(([v[0], v[1], v[4], v[5]], 3))
                if j == ny - 1: tags.append(([v[2], v[3], v[6], v[7]], 4))
                if k == 0: tags.append(([v[0], v[1], v[2], v[3]], 5))
                if k == nz - 1: tags.append(([v[4], v[5], v[6], v[7]], 6))
    return xyz, cells, tags


def box_mesh(nx, ny, nz, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), default="outflow", tag_types=None):
    from mpcns_uns_amd import Mesh

    xyz, cells, tags = box_cells(nx, ny, nz, lo, hi)
    m = Mesh.from_cells(xyz, cells, tags=tags)
    m.boundary(tag_types or {}, default=default)
    m._cells = (xyz, cells, tags)
    return m


def floored_state(w, n=40, seed=77):
    """The workload's initial state with cells the step must floor (Solver.cpp:171-191): n seeded cells whose Na+
    density is negative (the density branch) and n others whose H+ internal energy is negative (the pressure
    branch).  Returns (U_H, U_Na, Phi) and the two index sets."""
    UH, UNa, Phi = (np.array(a, copy=True) for a in w["state"])
    nc = UH.size // 5
    rng = np.random.default_rng(seed)
    cells = rng.choice(nc, 2 * n, replace=False)
    rho_cells, p_cells = cells[:n], cells[n:]
    Na = UNa.reshape(-1, 5)
    Na[rho_cells, 0] = -1e-3
    H = UH.reshape(-1, 5)
    ke = 0.5 * (H[p_cells, 1] ** 2 + H[p_cells, 2] ** 2 + H[p_cells, 3] ** 2) / H[p_cells, 0]
    H[p_cells, 4] = ke - 0.1
    return (UH, UNa, Phi), rho_cells, p_cells


def rough_gradients(w, scale=50.0, seed=5):
    """Gradients (15 per cell) large enough that the linear face states of many cells have rho or p <= 0, which
    sends fluid_flux_rhs's face_state to its first-order fallback (FluxRHS.cpp:16-27)."""
    rng = np.random.default_rng(seed)
    return scale * rng.normal(size=15 * w["mesh"].n_cells)


def fallback_faces(mesh, W, grad):
    """Number of (face, side) pairs whose linear reconstruction has rho <= 0 or p <= 0 (the fallback branch)."""
    c2f_ptr, c2f = mesh.array("c2f.ptr"), mesh.array("c2f.idx")
    fc, cc = mesh.array("face_center").reshape(-1, 3), mesh.array("cell_center").reshape(-1, 3)
    cell = np.repeat(np.arange(mesh.n_cells), np.diff(c2f_ptr))
    dx = fc[c2f] - cc[cell]
    G = grad.reshape(-1, 5, 3)[cell]
    Wc = W.reshape(-1, 5)[cell]
    rho = Wc[:, 0] + np.einsum("ij,ij->i", G[:, 0], dx)
    p = Wc[:, 4] + np.einsum("ij,ij->i", G[:, 4], dx)
    return int(np.count_nonzero((rho <= 0) | (p <= 0)))


def d2_of(mesh, Phi):
    """(d2 Phi)_c = sum over the cell's faces of the signed flux, by integer incidence (c2f)."""
    ptr, idx, sgn = mesh.array("c2f.ptr"), mesh.array("c2f.idx"), mesh.array("c2f.sgn")
    return np.add.reduceat(sgn * Phi[idx], ptr[:-1])


def hex_workload(hall, nx=6, ny=5, nz=4, jitter=0.1, seed=11, noise=0.01, **cfg_kw):
    """A Hall (or ideal) workload on a jittered hexahedral box away from the dipole (every boundary face outflow):
    cells of six quadrilateral faces, so the device runs its row code for rows other than the tetrahedron's four faces
    and three edges. Returns the workload dict of mpcns_uns_amd.workloads.build plus the cells and tags."""
    from mpcns_uns_amd import Mesh, workloads

    lo, hi = (1.5, -1.0, -1.0), (3.5, 1.0, 1.0)
    xyz, cells, tags = box_cells(nx, ny, nz, lo, hi)
    rng = np.random.default_rng(seed)
    h = np.array([(hi[0] - lo[0]) / nx, (hi[1] - lo[1]) / ny, (hi[2] - lo[2]) / nz])
    inner = np.all((xyz > np.array(lo) + 1e-9) & (xyz < np.array(hi) - 1e-9), axis=1)
    xyz = xyz + inner[:, None] * rng.uniform(-jitter, jitter, xyz.shape) * h
    mesh = Mesh.from_cells(xyz, cells, tags=tags)
    mesh.boundary({}, default="outflow", u_inf=workloads.U_INF)
    cfg = workloads.physics_config(hall, **cfg_kw)
    PhiAdd, Badd = mesh.background(workloads.B_IMF, workloads.DIPOLE_M, (0.0, 0.0, 0.0), 1e-3)
    UH, UNa, Phi = mesh.seeded_state(cfg, noise, seed, Badd)
    return dict(mesh=mesh, cfg=cfg, state=(UH, UNa, Phi), background=(PhiAdd, Badd), cells=cells, tags=tags, xyz=xyz)


def pyramid_split(xyz, cells, every=2):
    """Split every `every`-th hexahedron into six pyramids about its centroid (a new node), each on one of the hex's
    outward face loops with side triangles (b, a, apex) for base edge a -> b, so the mesh stays conforming: cells of
    six and of five faces, faces of four and of three edges."""
    xyz = [np.asarray(p, float) for p in xyz]
    out = []
    for i, loops in enumerate(cells):
        if i % every:
            out.append(loops)
            continue
        corners = sorted({v for lp in loops for v in lp})
        apex = len(xyz)
        xyz.append(np.mean([xyz[v] for v in corners], axis=0))
        for lp in loops:
            sides = [[lp[(k + 1) % 4], lp[k], apex] for k in range(4)]
            out.append([list(lp)] + sides)
    return np.array(xyz), out


def mixed_workload(hall, nx=6, ny=5, nz=4, jitter=0.1, seed=13, noise=0.01, **cfg_kw):
    """hex_workload's box with every other hexahedron split into six pyramids: rows of varying length everywhere."""
    from mpcns_uns_amd import Mesh, workloads

    lo, hi = (1.5, -1.0, -1.0), (3.5, 1.0, 1.0)
    xyz, cells, tags = box_cells(nx, ny, nz, lo, hi)
    rng = np.random.default_rng(seed)
    h = np.array([(hi[0] - lo[0]) / nx, (hi[1] - lo[1]) / ny, (hi[2] - lo[2]) / nz])
    inner = np.all((xyz > np.array(lo) + 1e-9) & (xyz < np.array(hi) - 1e-9), axis=1)
    xyz = xyz + inner[:, None] * rng.uniform(-jitter, jitter, xyz.shape) * h
    xyz, cells = pyramid_split(xyz, cells)
    mesh = Mesh.from_cells(xyz, cells, tags=tags)
    mesh.boundary({}, default="outflow", u_inf=workloads.U_INF)
    cfg = workloads.physics_config(hall, **cfg_kw)
    PhiAdd, Badd = mesh.background(workloads.B_IMF, workloads.DIPOLE_M, (0.0, 0.0, 0.0), 1e-3)
    UH, UNa, Phi = mesh.seeded_state(cfg, noise, seed, Badd)
    return dict(mesh=mesh, cfg=cfg, state=(UH, UNa, Phi), background=(PhiAdd, Badd), cells=cells, tags=tags, xyz=xyz)
