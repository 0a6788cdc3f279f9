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
