"""numpy restatement of the Tet4 block inputs (mesh, noise, masses, boundary) used by
the scenes of include/symx_scenes.h, so that CPU tests can build the same problem
without a GPU. Mesh split: mesh_generation.cpp:64-84; masses:
dynamic_elasticity_with_contact.cpp:67-79."""
import numpy as np

from oracle_py import SceneData


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    out = np.empty(n)
    s = seed & 0xFFFFFFFFFFFFFFFF
    M = 0xFFFFFFFFFFFFFFFF
    for i in range(n):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z = z ^ (z >> 31)
        out[i] = (z >> 11) * (1.0 / 9007199254740992.0)
    return out


def tet4_mesh(cells, size=(1.0, 1.0, 1.0)):
    nx, ny, nz = cells
    d = [size[0] / nx, size[1] / ny, size[2] / nz]
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    X = np.stack([I.ravel() * d[0], J.ravel() * d[1], K.ravel() * d[2]], 1)

    def idx(i, j, k):
        return i * (ny + 1) * (nz + 1) + j * (nz + 1) + k

    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    v = [idx(i, j, k), idx(i + 1, j, k), idx(i + 1, j + 1, k), idx(i, j + 1, k),
         idx(i, j, k + 1), idx(i + 1, j, k + 1), idx(i + 1, j + 1, k + 1), idx(i, j + 1, k + 1)]
    pat = [(0, 1, 6, 5), (0, 1, 2, 6), (0, 2, 3, 6), (0, 5, 6, 4), (0, 6, 3, 7), (0, 6, 7, 4)]
    tets = np.stack([np.stack([v[a] for a in t], 1) for t in pat], 1).reshape(-1, 4)
    return X, tets.astype(np.int32)


def tet4_block(cells=(4, 4, 4), size=(1.0, 1.0, 1.0), E=1e5, nu=0.4, density=1000.0, gravity=(0, 0, -9.81), dt=0.01,
               noise=0.01, seed=1, k=1e8, fixed_axis=2) -> SceneData:
    X, tets = tet4_mesh(cells, size)
    nv = X.shape[0]
    fixed = np.abs(X[:, fixed_axis]) < 1e-12 if fixed_axis >= 0 else np.zeros(nv, bool)
    u = splitmix64_uniform(seed, 3 * nv).reshape(nv, 3)
    x = X + np.where(fixed[:, None], 0.0, noise * (2.0 * u - 1.0))
    A = X[tets[:, 0]]
    det = np.linalg.det(np.stack([X[tets[:, 1]] - A, X[tets[:, 2]] - A, X[tets[:, 3]] - A], 2))
    mass = np.zeros(nv)
    np.add.at(mass, tets.ravel(), np.repeat(np.abs(det) / 6.0 * density / 4.0, 4))
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    bc = np.nonzero(fixed)[0]
    return SceneData(tets, X, x, x.copy(), np.zeros_like(x), mass, bc, X[bc], lam, mu, dt, k, gravity)


def sphere_mesh(radius=0.5, cells=50, jitter=0.1, seed=7, center=(0.0, 0.0, 0.51)):
    """Jittered-grid ball (symx_scenes.h sphere scene): (X, tets, contact vertices)."""
    r = radius
    h = 2.0 * r / cells
    G, T = tet4_mesh((cells,) * 3, (2 * r,) * 3)
    G = G + (np.asarray(center) - r)
    u = splitmix64_uniform(seed, 3 * G.shape[0]).reshape(-1, 3)
    J = G + jitter * h * (2.0 * u - 1.0)
    c = G[T].mean(axis=1)
    keep = ((c - np.asarray(center)) ** 2).sum(axis=1) <= r * r
    T = T[keep]
    used = np.zeros(G.shape[0], bool)
    used[T.ravel()] = True
    ids = -np.ones(G.shape[0], np.int64)
    ids[used] = np.arange(used.sum())
    X = J[used]
    T = ids[T].astype(np.int32)
    faces = np.sort(np.concatenate([T[:, [0, 1, 2]], T[:, [0, 1, 3]], T[:, [0, 2, 3]], T[:, [1, 2, 3]]]), axis=1)
    uniq, counts = np.unique(faces, axis=0, return_counts=True)
    contact = np.unique(uniq[counts == 1].ravel()).astype(np.int32)
    return X, T, contact


HEX_NODES = [(0, 0, 0), (2, 0, 0), (2, 2, 0), (0, 2, 0), (0, 0, 2), (2, 0, 2), (2, 2, 2), (0, 2, 2),
             (1, 0, 0), (2, 1, 0), (1, 2, 0), (0, 1, 0), (1, 0, 2), (2, 1, 2), (1, 2, 2), (0, 1, 2),
             (0, 0, 1), (2, 0, 1), (2, 2, 1), (0, 2, 1),
             (0, 1, 1), (2, 1, 1), (1, 2, 1), (1, 0, 1), (1, 1, 0), (1, 1, 2),
             (1, 1, 1)]


def hex_mesh(cells, size=(1.0, 1.0, 1.0), nn=27):
    """Hex8 / Hex27 cuboid over [0, size]: (X, elements) in SymX's element node order."""
    sub = 2 if nn == 27 else 1
    N = [sub * c + 1 for c in cells]
    d = [size[i] / (sub * cells[i]) for i in range(3)]
    I, J, K = np.meshgrid(np.arange(N[0]), np.arange(N[1]), np.arange(N[2]), indexing="ij")
    X = np.stack([I.ravel() * d[0], J.ravel() * d[1], K.ravel() * d[2]], 1)
    i, j, k = np.meshgrid(np.arange(cells[0]), np.arange(cells[1]), np.arange(cells[2]), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    step = 2 if nn == 27 else 1
    cols = []
    for (a, b, c) in HEX_NODES[:nn]:
        da, db, dc = (a, b, c) if nn == 27 else (a // 2, b // 2, c // 2)
        cols.append((step * i + da) * N[1] * N[2] + (step * j + db) * N[2] + (step * k + dc))
    return X, np.stack(cols, 1).astype(np.int32)


# Tet10 element nodes in the units of HEX_NODES (0, 1, 2 along each axis): the vertices of
# the unit tetrahedron, then the mid-edges 01 12 02 03 13 23
TET10_NODES = [(0, 0, 0), (2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)]
TET10_SPLIT = [(0, 1, 6, 5), (0, 1, 2, 6), (0, 2, 3, 6), (0, 5, 6, 4), (0, 6, 3, 7), (0, 6, 7, 4)]
TET10_EDGES = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]


def tet10_mesh(cells, size=(1.0, 1.0, 1.0)):
    """Tet10 cuboid over [0, size] (mesh_generation.cpp:93-159 without the RCM renumbering):
    nodes on the (2n+1)^3 lattice, each cell split into the six tetrahedra of the Tet4 block,
    element nodes v0..v3 then the mid-edges 01 12 02 03 13 23."""
    N = [2 * c + 1 for c in cells]
    d = [size[i] / (2 * cells[i]) for i in range(3)]
    I, J, K = np.meshgrid(np.arange(N[0]), np.arange(N[1]), np.arange(N[2]), indexing="ij")
    X = np.stack([I.ravel() * d[0], J.ravel() * d[1], K.ravel() * d[2]], 1)
    i, j, k = np.meshgrid(np.arange(cells[0]), np.arange(cells[1]), np.arange(cells[2]), indexing="ij")
    base = np.stack([2 * i.ravel(), 2 * j.ravel(), 2 * k.ravel()], 1)
    corner = np.array(HEX_NODES[:8])
    lin = lambda p: p[:, 0] * N[1] * N[2] + p[:, 1] * N[2] + p[:, 2]
    T = []
    for t in TET10_SPLIT:
        cols = [lin(base + corner[a]) for a in t]
        cols += [lin(base + (corner[t[a]] + corner[t[b]]) // 2) for (a, b) in TET10_EDGES]
        T.append(np.stack(cols, 1))
    # cells in i-major order, the six tetrahedra of a cell consecutive
    return X, np.stack(T, 1).reshape(-1, 10).astype(np.int32)


def hex_block(cells=(2, 2, 2), size=(1.0, 1.0, 1.0), nn=27, E_min=1e4, E_max=1e6, nu=0.45, density=1000.0, gravity=(0, 0, -9.81),
              dt=0.01, noise=0.1, seed=3, k=1e8, fixed_axis=2):
    """numpy restatement of symx_scene_hex_block_create's inputs (SceneData with mu_e);
    nn = 10 builds the Tet10 block of the same cuboid."""
    X, T = tet10_mesh(cells, size) if nn == 10 else hex_mesh(cells, size, nn)
    nv = X.shape[0]
    sub = 1 if nn == 8 else 2
    h = np.array([size[i] / (sub * cells[i]) for i in range(3)])
    fixed = np.abs(X[:, fixed_axis]) < 1e-12 if fixed_axis >= 0 else np.zeros(nv, bool)
    u = splitmix64_uniform(seed, 3 * nv).reshape(nv, 3)
    x = X + np.where(fixed[:, None], 0.0, noise * h * (2.0 * u - 1.0))
    ne = T.shape[0]
    ue = splitmix64_uniform(seed ^ 0x9E3779B97F4A7C15, ne)
    E = np.exp(np.log(E_min) + ue * (np.log(E_max) - np.log(E_min)))
    vol = np.prod([size[i] / cells[i] for i in range(3)]) / (6.0 if nn == 10 else 1.0)
    mass = np.zeros(nv)
    np.add.at(mass, T.ravel(), density * vol / nn)
    mu_e = E / (2 * (1 + nu))
    lam_over_mu = 2 * nu / (1 - 2 * nu)
    bc = np.nonzero(fixed)[0]
    S = SceneData(T, X, x, x.copy(), np.zeros_like(x), mass, bc, X[bc], 0.0, 0.0, dt, k, gravity, mu_e=mu_e, lambda_over_mu=lam_over_mu)
    S.E = E
    return S
