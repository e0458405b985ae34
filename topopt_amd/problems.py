"""Boundary-condition and load sets of the benchmark workloads.

Node numbering follows the reference (core/src/fea/assembly.cpp:41-43):
node(a, b, c) = (c*(ny+1) + b)*(nx+1) + a, DOF = 3*node + component.
Every set returns (bcs, loads) as lists of (node, component, value) tuples, the
same triples as the reference's DirichletBC / NodalLoad
(core/include/topopt/fea.hpp:180-192).
"""
import numpy as np


def node_index(nx, ny, a, b, c):
    return (c * (ny + 1) + b) * (nx + 1) + a


def cantilever(nx, ny, nz):
    """x=0 face fully clamped; unit downward (-z) line load along the bottom edge
    at x=nx (nodes a=nx, c=0, b=0..ny), split equally."""
    bcs = []
    for c in range(nz + 1):
        for b in range(ny + 1):
            n = node_index(nx, ny, 0, b, c)
            bcs += [(n, 0, 0.0), (n, 1, 0.0), (n, 2, 0.0)]
    loads = []
    for b in range(ny + 1):
        loads.append((node_index(nx, ny, nx, b, 0), 2, -1.0 / (ny + 1)))
    return bcs, loads


def mbb(nx, ny, nz):
    """MBB half-beam with y vertical: symmetry u_x=0 on the x=0 face; roller
    u_y=0 on the far bottom edge (a=nx, b=0, all c), plus u_z=0 on that edge to
    remove the z rigid mode; unit downward (-y) load along the top edge of the
    symmetry plane (a=0, b=ny, all c), split equally over the nz+1 nodes."""
    bcs = []
    for c in range(nz + 1):
        for b in range(ny + 1):
            bcs.append((node_index(nx, ny, 0, b, c), 0, 0.0))
    for c in range(nz + 1):
        n = node_index(nx, ny, nx, 0, c)
        bcs += [(n, 1, 0.0), (n, 2, 0.0)]
    loads = [(node_index(nx, ny, 0, ny, c), 1, -1.0 / (nz + 1)) for c in range(nz + 1)]
    return bcs, loads


def bridge(nx, ny, nz):
    """Bridge with z vertical: both ends pinned (all 3 DOF) along the bottom
    edges (a=0, c=0, all b) and (a=nx, c=0, all b). Returns (bcs, [loads...])
    with three point-load cases: unit -z at the deck centre line (b=ny//2,
    c=nz) at x = nx/4, nx/2, 3nx/4. The fourth case (self-weight) comes from
    self_weight()."""
    bcs = []
    for a in (0, nx):
        for b in range(ny + 1):
            n = node_index(nx, ny, a, b, 0)
            bcs += [(n, 0, 0.0), (n, 1, 0.0), (n, 2, 0.0)]
    cases = []
    for a in (nx // 4, nx // 2, 3 * nx // 4):
        cases.append([(node_index(nx, ny, a, ny // 2, nz), 2, -1.0)])
    return bcs, cases


def self_weight(nx, ny, nz, density=1.0, gravity=1.0, direction=(0.0, 0.0, -1.0), h=1.0,
                solid=None):
    """Design-independent self-weight: every solid voxel's weight lumped 1/8 to
    each corner node (core/src/fea/assembly.cpp:70-116). Returns (node, comp,
    value) triples for the non-zero entries."""
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    per_node = gravity * density * h ** 3 / 8.0
    f = np.zeros((nz + 1, ny + 1, nx + 1))
    s = np.ones((nz, ny, nx)) if solid is None else np.asarray(solid, dtype=np.float64).reshape(nz, ny, nx)
    for dc in (0, 1):
        for db in (0, 1):
            for da in (0, 1):
                f[dc:dc + nz, db:db + ny, da:da + nx] += s * per_node
    loads = []
    flat = f.ravel()
    for n in np.nonzero(flat)[0]:
        for comp in range(3):
            v = flat[n] * d[comp]
            if v != 0.0:
                loads.append((int(n), comp, float(v)))
    return loads


def to_arrays(triples):
    """(node, comp, value) triples -> (int32 nodes, int32 comps, float64 values)."""
    t = list(triples)
    return (np.array([x[0] for x in t], dtype=np.int32),
            np.array([x[1] for x in t], dtype=np.int32),
            np.array([x[2] for x in t], dtype=np.float64))


def workload(name):
    """Named workloads: (grid, bcs, load cases, SIMP settings)."""
    if name == "cantilever_60x20x20":
        g = (60, 20, 20)
        bcs, loads = cantilever(*g)
        return dict(grid=g, bcs=bcs, loads=[loads], volfrac=0.3, rmin=1.5, penalty=3.0,
                    iterations=50, tol=1e-8, move=0.2)
    if name == "mbb_192x64x64":
        g = (192, 64, 64)
        bcs, loads = mbb(*g)
        return dict(grid=g, bcs=bcs, loads=[loads], volfrac=0.2, rmin=2.0, penalty=3.0,
                    iterations=100, tol=1e-6, move=0.2)
    if name == "bridge_256x64x96":
        g = (256, 64, 96)
        bcs, cases = bridge(*g)
        cases.append(self_weight(*g))
        return dict(grid=g, bcs=bcs, loads=cases, volfrac=0.15, rmin=2.0, penalty=3.0,
                    iterations=80, tol=1e-6, move=0.2)
    if name == "cantilever_512x256x256":
        g = (512, 256, 256)
        bcs, loads = cantilever(*g)
        return dict(grid=g, bcs=bcs, loads=[loads], volfrac=0.1, rmin=3.0, penalty=3.0,
                    iterations=150, tol=1e-6, move=0.2,
                    continuation=(1.0, 3.0, 0.5, 6))
    raise KeyError(name)
