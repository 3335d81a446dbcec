"""Synthetic inputs for the hot-path workloads (numpy, host side).

Orderings follow the reference where it has one:
  * create_box: points x-major (x*ny*nz + y*nz + z) and springs x-edges, then
    y-edges, then z-edges (src/graph.cpp:39-42, :66-115);
  * tet and hex grids use the same x-major vertex numbering, cubes x-major.

The 5-tet split uses global parity so that neighbouring cubes agree on their
face diagonals: in cube (i,j,k) the four corners with (i+dx+j+dy+k+dz) even
form the central tet and every odd corner forms a corner tet with its three
(even) neighbours.  Tets are oriented so that fem_linear.sim's
W = -(1/6) det([x0-x3, x1-x3, x2-x3]) is positive.  Hex corner a is
dx + 2*dy + 4*dz.
"""
from __future__ import annotations

import numpy as np

# ---- lattices ------------------------------------------------------------------


def create_box(nx: int, ny: int, nz: int):
    """Point count and int32[ns, 2] spring endpoints in createBox order."""
    idx = np.arange(nx * ny * nz, dtype=np.int64).reshape(nx, ny, nz)
    ex = np.stack([idx[:-1, :, :].ravel(), idx[1:, :, :].ravel()], axis=1)
    ey = np.stack([idx[:, :-1, :].ravel(), idx[:, 1:, :].ravel()], axis=1)
    ez = np.stack([idx[:, :, :-1].ravel(), idx[:, :, 1:].ravel()], axis=1)
    return nx * ny * nz, np.concatenate([ex, ey, ez]).astype(np.int32)


def grid_coords(nx: int, ny: int, nz: int, spacing=1.0):
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1).astype(np.float64) * spacing


def springs_lattice(n: int, seed: int = 0, jitter: float = 0.1, k: float = 1e3,
                    l0: float = 1.0, mass: float = 1.0):
    """Jittered n^3 lattice, 6-neighbour springs, z=0 layer fixed."""
    rng = np.random.default_rng(seed)
    npts, ep = create_box(n, n, n)
    x = grid_coords(n, n, n) + rng.uniform(-jitter, jitter, size=(npts, 3))
    ijk = grid_coords(n, n, n)
    return {
        "npoints": npts,
        "x": x, "v": np.zeros((npts, 3)), "m": np.full(npts, mass),
        "fixed": (ijk[:, 2] == 0).astype(np.uint8),
        "endpoints": ep,
        "k": np.full(ep.shape[0], k), "l0": np.full(ep.shape[0], l0),
    }


def springs_batch(nmesh: int = 256, nmin: int = 8, nmax: int = 36, seed: int = 0,
                  klo: float = 1e2, khi: float = 1e4, jitter: float = 0.1):
    """nmesh independent jittered n^3 spring lattices (n uniform in nmin..nmax,
    stiffness log-uniform in [klo, khi] per lattice, z=0 layer fixed) stored
    as one disjoint union; rows[m]..rows[m+1] are lattice m's points."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, size=nmesh)
    ks = 10.0 ** rng.uniform(np.log10(klo), np.log10(khi), size=nmesh)
    parts = [springs_lattice(int(n), seed=seed * 100003 + m, jitter=jitter, k=float(k))
             for m, (n, k) in enumerate(zip(ns, ks))]
    rows = np.zeros(nmesh + 1, dtype=np.int64)
    rows[1:] = np.cumsum([p["npoints"] for p in parts])
    ep = np.concatenate([p["endpoints"] + rows[m] for m, p in enumerate(parts)])
    cat = {k: np.concatenate([p[k] for p in parts]) for k in ("x", "v", "m", "fixed", "k", "l0")}
    return {"npoints": int(rows[-1]), "rows": rows.astype(np.uint32), "n": ns, "kmesh": ks,
            "endpoints": ep.astype(np.int32), **cat}


# ---- tet and hex grids ------------------------------------------------------------
_TET5 = {
    0: [[0, 3, 5, 6], [0, 3, 5, 1], [3, 0, 6, 2], [5, 6, 0, 4], [6, 5, 3, 7]],
    1: [[1, 2, 4, 7], [1, 2, 4, 0], [2, 1, 7, 3], [4, 7, 1, 5], [7, 4, 2, 6]],
}


def _corner_offsets():
    return np.array([[a & 1, (a >> 1) & 1, (a >> 2) & 1] for a in range(8)], dtype=np.int64)


def _oriented_templates():
    off = _corner_offsets().astype(np.float64)
    out = {}
    for par, tets in _TET5.items():
        fixed = []
        for t in tets:
            t = list(t)
            M = np.stack([off[t[i]] - off[t[3]] for i in range(3)], axis=1)
            if np.linalg.det(M) > 0:
                t[0], t[1] = t[1], t[0]
            fixed.append(t)
        out[par] = np.array(fixed, dtype=np.int64)
    return out


def cube_corners(n: int):
    """int64[n^3, 8] vertex ids of every cube's corners (corner a = dx+2dy+4dz)."""
    N1 = n + 1
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    off = _corner_offsets()
    return ((i[:, None] + off[:, 0]) * N1 * N1 + (j[:, None] + off[:, 1]) * N1
            + (k[:, None] + off[:, 2])), (i + j + k) & 1


def tet5_grid(n: int):
    """Unit cube split into n^3 cubes of 5 tets: coords [(n+1)^3, 3], tets int32[5n^3, 4]."""
    corners, par = cube_corners(n)
    tmpl = _oriented_templates()
    tets = np.empty((n ** 3, 5, 4), dtype=np.int64)
    for p in (0, 1):
        sel = par == p
        tets[sel] = corners[sel][:, tmpl[p]]
    x = grid_coords(n + 1, n + 1, n + 1, 1.0 / n)
    return x, tets.reshape(-1, 4).astype(np.int32)


def hex_grid(n: int):
    corners, _ = cube_corners(n)
    x = grid_coords(n + 1, n + 1, n + 1, 1.0 / n)
    return x, corners.astype(np.int32)


# ---- random material fields ------------------------------------------------------------


def gaussian_random_field(n: int, corr_len: float, seed: int):
    """Zero-mean, unit-std periodic field on n^3 cells: seeded white noise
    smoothed in Fourier space by a Gaussian kernel of std corr_len cells."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((n, n, n))
    f = np.fft.fftfreq(n)
    kx, ky, kz = np.meshgrid(f, f, f, indexing="ij")
    k2 = (2 * np.pi) ** 2 * (kx ** 2 + ky ** 2 + kz ** 2)
    g = np.real(np.fft.ifftn(np.fft.fftn(w) * np.exp(-0.5 * k2 * corr_len ** 2)))
    g -= g.mean()
    s = g.std()
    return g / s if s > 1e-12 else np.zeros_like(g)


def lame(E, nu):
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return mu, lam


def tet_problem(n: int, seed: int = 0, field: str = "iid", mean_log10=5.0, std_log10=0.5,
                corr_len=8.0, nu=0.3, load=(0.0, 0.0, -9.81e3)):
    """Clamped-bottom elasticity on the 5-tet grid (z=0 face constrained)."""
    x, tets = tet5_grid(n)
    if field == "iid":
        rng = np.random.default_rng(seed)
        log10E = rng.normal(mean_log10, std_log10, size=tets.shape[0])
    else:
        g = gaussian_random_field(n, corr_len, seed).ravel()
        log10E = np.repeat(mean_log10 + std_log10 * g, 5)
    E = 10.0 ** log10E
    mu, lam = lame(E, nu)
    nv = x.shape[0]
    c = (np.arange(nv) % (n + 1) == 0).astype(np.int32)  # z index 0
    return {"nverts": nv, "x": x, "c": c, "endpoints": tets, "mu": mu, "lam": lam,
            "load": load}


def hex_problem(n: int, seed: int = 0, std_log10=1.5, corr_len=12.0, source=1.0):
    """Scalar Poisson on n^3 trilinear hexes, T=0 on the x=0 face."""
    x, hexes = hex_grid(n)
    g = gaussian_random_field(n, corr_len, seed).ravel() if n > 1 else np.zeros(1)
    kappa = 10.0 ** (std_log10 * g)
    nv = x.shape[0]
    N1 = n + 1
    dirichlet = (np.arange(nv) // (N1 * N1) == 0).astype(np.int32)
    return {"nverts": nv, "x": x, "dirichlet": dirichlet, "endpoints": hexes, "kappa": kappa,
            "h": 1.0 / n, "source": source}


# ---- tetgen meshes -------------------------------------------------------------------------
def _tetgen_rows(path, ncols):
    """The data rows of a tetgen file as the reference reads them
    (src/mesh.cpp:252-360): the first integer of the file is the row count;
    then, line by line, lines shorter than 3 characters and '#' comments are
    skipped, and each data line gives its index (ignored) and ncols values
    (extra columns ignored), until the count is reached."""
    with open(path) as f:
        first = f.readline()
        head = first.split()
        count = int(head[0])
        rows = []
        for line in f:
            if len(rows) >= count:
                break
            line = line.rstrip("\n")
            if len(line) < 3 or line[0] == "#":
                continue
            rows.append(line.split()[1:1 + ncols])
    return head, rows


def load_tetgen(node_path, ele_path, edge_path=None):
    """A tetgen volume mesh as MeshVol::loadTet / loadTetEdge read it
    (src/mesh.h:42-95, src/mesh.cpp:252-360): {"x": float64 [nv, 3],
    "elements": int32 [ne, nodes per element], "edges": int32 [nedge, 2] or
    None}.  Vertex ids are kept as written (the reference does not rebase)."""
    _, xr = _tetgen_rows(node_path, 3)
    x = np.array([[float(v) for v in r] for r in xr], dtype=np.float64).reshape(-1, 3)
    head, er = _tetgen_rows(ele_path, 0)
    nper = int(head[1]) if len(head) > 1 else 4
    _, er = _tetgen_rows(ele_path, nper)
    el = np.array([[int(v) for v in r] for r in er], dtype=np.int32).reshape(-1, nper)
    edges = None
    if edge_path is not None:
        _, dr = _tetgen_rows(edge_path, 2)
        edges = np.array([[int(v) for v in r] for r in dr], dtype=np.int32).reshape(-1, 2)
    return {"x": x, "elements": el, "edges": edges}


def tet_signed_volume(x, tets):
    """W = -det([x_0 - x_3, x_1 - x_3, x_2 - x_3]) / 6 per tet: the volume
    convention of the linear-tet element (positive for a usable tet)."""
    d = x[tets[:, :3]] - x[tets[:, 3:4]]
    return -np.linalg.det(d.transpose(0, 2, 1)) / 6.0


def tet_mesh_problem(x, tets, E=1e5, nu=0.3, clamp_frac=0.05, load=(0.0, 0.0, -9.81e3)):
    """Clamped-bottom linear elasticity on an arbitrary tet mesh: tets whose
    W is negative get two vertices swapped (so every element is positively
    oriented for the element formula), vertices within clamp_frac of the
    height above the lowest one are clamped, E is a scalar or per tet."""
    t = np.array(tets, dtype=np.int32, copy=True)
    neg = tet_signed_volume(x, t) < 0
    t[neg, 0], t[neg, 1] = tets[neg, 1], tets[neg, 0]
    E = np.broadcast_to(np.asarray(E, dtype=np.float64), (t.shape[0],)).copy()
    mu, lam = lame(E, nu)
    z = x[:, 2]
    c = (z <= z.min() + clamp_frac * (z.max() - z.min())).astype(np.int32)
    return {"nverts": x.shape[0], "x": np.ascontiguousarray(x), "c": c, "endpoints": t,
            "mu": mu, "lam": lam, "load": load}


def femtet_problem(node_path, ele_path):
    """The reference's femTet fixture set-up (test/program-fem-test.cpp:14-75):
    E = 5e3, nu = 0.45, u = 0.5 E / nu, l = E nu / ((1 + nu)(1 - 2 nu)) per
    tet; every vertex starts with v = (0.1, 0, 0.1), c = 0, m = 0, except
    those with y < 1e-4, which are constrained (c = 1, v = 0)."""
    mesh = load_tetgen(node_path, ele_path)
    x = np.ascontiguousarray(mesh["x"], dtype=np.float64)
    nv = x.shape[0]
    E, nu = 5e3, 0.45
    uval = 0.5 * E / nu
    lval = E * nu / ((1 + nu) * (1 - 2 * nu))
    v = np.tile(np.array([0.1, 0.0, 0.1]), (nv, 1))
    c = np.zeros(nv, dtype=np.int32)
    low = x[:, 1] < 0.0001
    c[low] = 1
    v[low] = 0.0
    ep = np.ascontiguousarray(mesh["elements"][:, :4], dtype=np.int32)
    nt = ep.shape[0]
    return {"nverts": nv, "x": x, "v": np.ascontiguousarray(v), "c": c,
            "fe": np.zeros((nv, 3)), "m": np.zeros(nv), "endpoints": ep,
            "u": np.full(nt, uval), "l": np.full(nt, lval)}
