"""Seeded synthetic particle packings, order-parameter colouring and initial
fields for the benchmark workloads.

Fields follow the benchmark definition: eta_k = 0.5 (1 - tanh((|x - x_i| -
R_i) / w)) for the particles i coloured k (maximum over them), c = sum_k
eta_k, mu = 0.  The reference's own initial profile (a cosine,
S/initial_values_spherical.h:51-76) is replaced by this tanh profile.

Colouring: particles i, j interact when |x_i - x_j| <= R_i + R_j + buffer
with buffer = w + interface_buffer_ratio * w + radius_buffer_ratio *
max(R_i, R_j) -- the rule of InitialValuesCloud
(S/initial_values_cloud.h:187-198).  The reference colours the resulting
graph with deal.II's SparsityTools::color_sparsity_pattern, which is not
available here; this module uses a greedy colouring in saturation order
(DSatur).  The
assignments it produces for the benchmark packings are committed as
fixtures (tests/golden/packings.json) and checked bit-exactly.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class Packing:
    centers: np.ndarray      # (n, dim)
    radii: np.ndarray        # (n,)
    lower: tuple             # domain
    upper: tuple
    colors: np.ndarray = None

    @property
    def dim(self):
        return self.centers.shape[1]

    @property
    def n_order_params(self):
        return int(self.colors.max()) + 1


def interaction_graph(p: Packing, w=1.0, interface_buffer_ratio=1.0, radius_buffer_ratio=0.0):
    n = len(p.radii)
    adj = [[] for _ in range(n)]
    for i in range(n):
        for j in range(i + 1, n):
            d = np.linalg.norm(p.centers[i] - p.centers[j])
            buf = w + interface_buffer_ratio * w + radius_buffer_ratio * max(p.radii[i], p.radii[j])
            if d <= p.radii[i] + p.radii[j] + buf:
                adj[i].append(j)
                adj[j].append(i)
    return adj


def greedy_coloring(adj):
    """greedy colouring in saturation order (DSatur): repeatedly colour the
    uncoloured particle with the most distinct neighbour colours (ties: more
    neighbours, then lower index) with the smallest free colour"""
    n = len(adj)
    colors = -np.ones(n, dtype=np.int64)
    for _ in range(n):
        best, pick = None, -1
        for i in range(n):
            if colors[i] >= 0:
                continue
            sat = len({colors[j] for j in adj[i] if colors[j] >= 0})
            key = (sat, len(adj[i]), -i)
            if best is None or key > best:
                best, pick = key, i
        used = {colors[j] for j in adj[pick] if colors[j] >= 0}
        c = 0
        while c in used:
            c += 1
        colors[pick] = c
    return colors


def _box(centers, radii, margin):
    lo = tuple(float(v) for v in (centers - radii[:, None]).min(axis=0) - margin)
    hi = tuple(float(v) for v in (centers + radii[:, None]).max(axis=0) + margin)
    return lo, hi


def _cube(lo, hi):
    """enlarge the box to a cube (uniform cells on an N^dim mesh)"""
    side = max(h - l for l, h in zip(lo, hi))
    mid = [(l + h) / 2 for l, h in zip(lo, hi)]
    return tuple(m - side / 2 for m in mid), tuple(m + side / 2 for m in mid)


def two_particles_2d():
    """2D neck growth: domain [-20,20]x[-10,10], R=5 at (+-5, 0)"""
    p = Packing(np.array([[-5.0, 0.0], [5.0, 0.0]]), np.array([5.0, 5.0]),
                (-20.0, -10.0), (20.0, 10.0))
    p.colors = np.array([0, 1])
    return p


def lattice_packing(n_per_dim, radii, overlap=0.5, margin=5.0):
    """spheres on a cubic lattice whose spacing makes neighbours of mean
    radius overlap by `overlap`"""
    n = n_per_dim ** 3
    assert len(radii) == n
    s = 2.0 * float(np.mean(radii)) - overlap
    idx = np.array([(i, j, k) for k in range(n_per_dim) for j in range(n_per_dim)
                    for i in range(n_per_dim)], dtype=np.float64)
    centers = (idx - (n_per_dim - 1) / 2.0) * s
    lo, hi = _cube(*_box(centers, radii, margin))
    return Packing(centers, np.asarray(radii, dtype=np.float64), lo, hi)


def packing_cfg1(seed=0):
    """2x2x2 packing of 8 spheres, R = 10 (1 + U(-0.1, 0.1))"""
    rng = np.random.default_rng(seed)
    radii = 10.0 * (1.0 + rng.uniform(-0.1, 0.1, size=8))
    p = lattice_packing(2, radii)
    # face-diagonal neighbours interact -> 4 order parameters
    p.colors = greedy_coloring(interaction_graph(p, radius_buffer_ratio=0.75))
    return p


def packing_cfg2(seed=0):
    """3x3x3 packing of 27 spheres, lognormal radii (mean 10, sigma 0.15)"""
    rng = np.random.default_rng(seed)
    sigma = 0.15
    radii = 10.0 * np.exp(rng.normal(-0.5 * sigma ** 2, sigma, size=27))
    p = lattice_packing(3, radii)
    # face and body diagonals interact -> 8 order parameters
    p.colors = greedy_coloring(interaction_graph(p, radius_buffer_ratio=1.25))
    return p


def random_packing(n, mean, sigma, box, seed=0, max_tries=200000, overlap=0.5):
    """random sequential packing: each new sphere is dropped at a random
    point and accepted if it overlaps no placed sphere by more than
    `overlap` and touches (overlaps) at least one when any exist"""
    rng = np.random.default_rng(seed)
    radii = mean * np.exp(rng.normal(-0.5 * sigma ** 2, sigma, size=n))
    centers = []
    tries = 0
    for r in radii:
        while True:
            tries += 1
            if tries > max_tries:
                raise RuntimeError("packing did not converge")
            c = rng.uniform(r, box - r, size=3)
            ok, touch = True, not centers
            for cj, rj in zip(centers, radii[:len(centers)]):
                d = np.linalg.norm(c - cj)
                if d < r + rj - overlap - 1e-12:
                    ok = False
                    break
                if d <= r + rj:
                    touch = True
            if ok and touch:
                centers.append(c)
                break
    centers = np.array(centers)
    p = Packing(centers, radii, (0.0, 0.0, 0.0), (float(box),) * 3)
    p.colors = greedy_coloring(interaction_graph(p))
    return p


def packing_cfg3(seed=0):
    """random sequential packing of 64 spheres, lognormal radii (mean 8,
    sigma 0.2) in a 160^3 box; the interaction buffer (2.075 max radius)
    spreads the grains over 10 order parameters for seed 0"""
    p = random_packing(64, 8.0, 0.2, 160.0, seed=seed)
    p.colors = greedy_coloring(interaction_graph(p, radius_buffer_ratio=2.075))
    return p


def packing_cfg4(seed=0):
    """random sequential packing of 100 spheres, lognormal radii (mean 8,
    sigma 0.2) in a 160^3 box; interaction buffer 2 max radius: 10-12 order
    parameters for seeds 0-4 (the operator is built with 12)"""
    p = random_packing(100, 8.0, 0.2, 160.0, seed=seed)
    p.colors = greedy_coloring(interaction_graph(p, radius_buffer_ratio=2.0))
    return p


def initial_fields(p: Packing, n_cells, w=1.0, n_order_params=None, dtype=np.float64):
    """interleaved nodal state (node-major, x fastest): c, mu, eta_0.."""
    dim = p.dim
    G = n_order_params or p.n_order_params
    axes = [np.linspace(l, u, n + 1) for l, u, n in zip(p.lower, p.upper, n_cells)]
    nn = [len(a) for a in axes]
    n_nodes = int(np.prod(nn))
    out = np.zeros((n_nodes, G + 2), dtype=dtype)
    # evaluate plane by plane along the last axis to bound memory
    plane = int(np.prod(nn[:-1]))
    grids = np.meshgrid(*reversed(axes[:-1]), indexing="ij")
    coords = [g.reshape(-1) for g in reversed(grids)]
    for k, z in enumerate(axes[-1]):
        eta = np.zeros((plane, G))
        for (ctr, r, col) in zip(p.centers, p.radii, p.colors):
            d2 = (z - ctr[-1]) ** 2
            for a in range(dim - 1):
                d2 = d2 + (coords[a] - ctr[a]) ** 2
            e = 0.5 * (1.0 - np.tanh((np.sqrt(d2) - r) / w))
            np.maximum(eta[:, col], e, out=eta[:, col])
        sl = slice(k * plane, (k + 1) * plane)
        out[sl, 2:] = eta
        out[sl, 0] = eta.sum(axis=1)
    return out.ravel()


def read_cloud(source, dim=3):
    """read_particles (S/particle.h:38-86): a particle cloud in CSV form,
    one header line, then one particle per line as "x,y,z,r"; the first
    `dim` coordinates are the centre (a 2D run reads x, y), the fourth column
    the radius; particles are numbered in file order.  `source` is a path or
    an open text stream."""
    if isinstance(source, str):
        with open(source) as f:
            return read_cloud(f, dim)
    centers, radii = [], []
    for k, line in enumerate(source):
        if k == 0:
            continue  # the header
        line = line.strip()
        if not line:
            continue
        cols = [c.strip() for c in line.split(",")]
        centers.append([float(c) for c in cols[:dim]])
        radii.append(float(cols[3]))
    return np.array(centers, dtype=np.float64), np.array(radii, dtype=np.float64)


def packing_from_cloud(source, dim=3, interface_width=2.0, boundary_factor=0.5,
                       interface_buffer_ratio=1.0, radius_buffer_ratio=0.0):
    """the packing of a particle cloud as the reference's InitialValuesCloud
    and driver set it up (S/initial_values_cloud.h:30-260, S/driver.h:
    505-525): domain = bounding box of the particles (centre +- radius)
    enlarged by boundary_factor * the largest radius on every side; order
    parameters from the interaction graph of InitialValuesCloud (particles
    interact when their distance is at most r_i + r_j + interface_width (1 +
    interface_buffer_ratio) + radius_buffer_ratio max(r_i, r_j)), coloured by
    DSatur (the reference colours with SparsityTools::color_sparsity_pattern).
    Defaults: GeometryData of S/parameters.h:57-75."""
    centers, radii = read_cloud(source, dim)
    rmax = float(radii.max())
    lo = tuple(float(v) - boundary_factor * rmax for v in (centers - radii[:, None]).min(axis=0))
    hi = tuple(float(v) + boundary_factor * rmax for v in (centers + radii[:, None]).max(axis=0))
    p = Packing(centers, radii, lo, hi)
    p.colors = greedy_coloring(interaction_graph(
        p, w=interface_width, interface_buffer_ratio=interface_buffer_ratio,
        radius_buffer_ratio=radius_buffer_ratio))
    return p


CONFIGS = {
    0: dict(name="2D two-particle neck growth", dim=2, n_cells=(160, 80),
            packing=two_particles_2d),
    1: dict(name="3D 2x2x2 packing of 8 spheres", dim=3, n_cells=(128, 128, 128),
            packing=packing_cfg1),
    2: dict(name="3D 3x3x3 packing of 27 spheres", dim=3, n_cells=(192, 192, 192),
            packing=packing_cfg2),
}
