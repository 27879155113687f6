"""Synthetic targets for the benchmark workloads (seeded, reproducible).

* c60():     analytic truncated icosahedron, 6:6 bonds 1.40 A, 5:6 bonds 1.45 A
* globule(): random sequential addition of atoms of a fixed composition in an
             ellipsoid with a minimum spacing, charges U[-0.5, 0.5] shifted to
             a chosen net charge (the 1,231-atom ubiquitin-size case, and the
             50,000-atom ellipsoid; above 20,000 atoms candidates are added
             in batches)
* shell():   the same inside a hollow sphere (1,000,000-atom case)
* assemblies(): the 64-target batch: atom counts log-uniform in
             [500, 200000], ellipsoids with axis ratios U[1, 2] at the
             1,231-atom globule's number density, net charge round(sqrt(N)/8),
             probes alternating He / N2

All write the reference's .xyz format (src/MoleculeTarget.cpp:148-319):
count line, comment line, then "Symbol x y z [q]".
"""
import itertools
import math

import numpy as np

ELEMENTS = ("C", "N", "O", "H", "S")
GLOBULE_FRACTIONS = (0.31, 0.08, 0.10, 0.50, 0.01)


def c60(b66=1.40, b56=1.45):
    """Truncated icosahedron with the two C60 bond lengths.

    Vertices are the cyclic permutations of (0, +-1, +-3p), (+-1, +-(2+p),
    +-2p), (+-p, +-2, +-(2p+1)) (p = golden ratio) for unit edges; the two
    bond classes are then set by moving every vertex along its pentagon's
    radial direction: pentagon edges (5:6) scale with the pentagon radius and
    hexagon-only edges (6:6) with the separation between pentagons."""
    p = (1 + 5 ** 0.5) / 2
    base = []
    for a, b, c in [(0, 1, 3 * p), (1, 2 + p, 2 * p), (p, 2, 2 * p + 1)]:
        for sa, sb, sc in itertools.product((1, -1), repeat=3):
            v = (sa * a, sb * b, sc * c)
            for k in range(3):
                w = v[k:] + v[:k]
                if w not in base:
                    base.append(w)
    V = np.array(sorted(set(base)), dtype=np.float64)
    assert len(V) == 60
    # the 12 pentagons: group vertices by nearest icosahedron vertex direction
    ico = []
    for s1, s2 in itertools.product((1, -1), repeat=2):
        for w in [(0, s1, s2 * p), (s1, s2 * p, 0), (s2 * p, 0, s1)]:
            ico.append(w)
    ico = np.array(ico) / np.linalg.norm(ico[0])
    owner = np.argmax(V @ ico.T, axis=1)
    out = np.zeros_like(V)
    # unit-edge geometry: pentagon circumradius r5, pentagon-centre distance D
    r5 = 1 / (2 * math.sin(math.pi / 5))
    for k in range(12):
        idx = np.where(owner == k)[0]
        cen = V[idx].mean(axis=0)
        dirn = cen / np.linalg.norm(cen)
        for i in idx:
            rad = V[i] - cen
            rad /= np.linalg.norm(rad)
            out[i] = rad * r5 * b56
        # place the pentagon so that inter-pentagon (6:6) bonds have length b66
        out[idx] += dirn * _pentagon_height(b56, b66)
    return out


def _pentagon_height(b56, b66):
    # distance from the cage centre to a pentagon plane such that the bond
    # between facing vertices of neighbouring pentagons has length b66
    p = (1 + 5 ** 0.5) / 2
    r5 = b56 / (2 * math.sin(math.pi / 5))
    # two neighbouring pentagon normals n1, n2 (icosahedron vertices) make
    # cos(theta) = 1/sqrt(5); facing vertices lie in the plane of n1, n2.
    ct = 1 / 5 ** 0.5
    st = (1 - ct * ct) ** 0.5
    # vertex of pentagon 1 pointing to pentagon 2: h n1 + r5 u1 with u1 in the
    # n1-n2 plane, perpendicular to n1; solve |v1 - v2| = b66 for h
    # v1 = (h, r5) in (n1, u1) coords; v2 = rotation of (h, -r5) by theta
    # |v1 - v2|^2 = b66^2 -> quadratic in h
    def dist(h):
        v1 = np.array([h, r5])
        x, y = h, -r5
        v2 = np.array([x * ct - y * st, x * st + y * ct])
        return np.linalg.norm(v1 - v2)
    lo, hi = 0.1, 20.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if dist(mid) > b66:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


def _rsa(n, inside, sample_box, min_dist, rng, batch=4096):
    """Random sequential addition with a hash grid of cell min_dist."""
    pts = np.zeros((n, 3))
    grid = {}
    count = 0
    lo, hi = sample_box
    inv = 1.0 / min_dist
    d2 = min_dist * min_dist
    tries = 0
    while count < n:
        cand = rng.uniform(lo, hi, size=(batch, 3))
        cand = cand[inside(cand)]
        for c in cand:
            tries += 1
            key = tuple(np.floor(c * inv).astype(np.int64))
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        lst = grid.get((key[0] + dx, key[1] + dy, key[2] + dz))
                        if lst:
                            for j in lst:
                                e = pts[j] - c
                                if e @ e < d2:
                                    ok = False
                                    break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                pts[count] = c
                grid.setdefault(key, []).append(count)
                count += 1
                if count == n:
                    break
        if tries > 200 * n + 100000:
            raise RuntimeError("random sequential addition did not converge")
    return pts


def _rsa_batched(n, inside, sample_box, min_dist, rng):
    """Random addition in batches for large n: a batch of candidates is
    thinned against the accepted points and against earlier candidates of the
    same batch (k-d trees), so every accepted pair is at least min_dist apart."""
    from scipy.spatial import cKDTree
    lo, hi = sample_box
    pts = np.zeros((0, 3))
    tree = None
    while len(pts) < n:
        m = max(20000, len(pts) // 4)
        cand = rng.uniform(lo, hi, size=(m, 3))
        cand = cand[inside(cand)]
        if tree is not None and len(cand):
            d, _ = tree.query(cand, k=1, distance_upper_bound=min_dist)
            cand = cand[~(d < min_dist)]
        if len(cand) > 1:
            pairs = cKDTree(cand).query_pairs(min_dist, output_type="ndarray")
            keep = np.ones(len(cand), dtype=bool)
            keep[pairs.max(axis=1)] = False  # the later candidate of each close pair
            cand = cand[keep]
        cand = cand[: n - len(pts)]
        if len(cand) == 0:
            raise RuntimeError("random addition did not converge")
        pts = np.concatenate([pts, cand])
        tree = cKDTree(pts)
    return pts


def _composition(n, rng, fractions=GLOBULE_FRACTIONS):
    counts = np.floor(np.array(fractions) * n).astype(int)
    counts[np.argmax(fractions)] += n - counts.sum()
    sym = np.repeat(np.array(ELEMENTS), counts)
    rng.shuffle(sym)
    return sym


def _charges(n, net, rng):
    q = rng.uniform(-0.5, 0.5, size=n)
    q += (net - q.sum()) / n
    return q


def globule(n=1231, axes=(18.0, 15.0, 13.0), net_charge=4.0, min_dist=1.0, seed=2104):
    """Atoms packed in an ellipsoid; returns (symbols, xyz, q)."""
    rng = np.random.default_rng(seed)
    a = np.array(axes, dtype=np.float64)
    inside = lambda P: ((P / a) ** 2).sum(axis=1) <= 1.0
    xyz = (_rsa if n <= 20000 else _rsa_batched)(n, inside, (-a, a), min_dist, rng)
    return _composition(n, rng), xyz, _charges(n, net_charge, rng)


def shell(n=1_000_000, r_in=120.0, r_out=140.0, net_charge=60.0, min_dist=1.0, seed=2104):
    """Atoms packed in a hollow sphere; returns (symbols, xyz, q)."""
    rng = np.random.default_rng(seed)
    inside = lambda P: (lambda r: (r >= r_in) & (r <= r_out))(np.linalg.norm(P, axis=1))
    box = (-np.full(3, r_out), np.full(3, r_out))
    xyz = (_rsa if n <= 20000 else _rsa_batched)(n, inside, box, min_dist, rng)
    return _composition(n, rng), xyz, _charges(n, net_charge, rng)


def write_xyz(path, symbols, xyz, q=None, comment="synthetic"):
    with open(path, "w") as fh:
        fh.write("%d\n%s\n" % (len(xyz), comment))
        for i, (s, r) in enumerate(zip(symbols, xyz)):
            if q is None:
                fh.write("%s %.6f %.6f %.6f\n" % (s, r[0], r[1], r[2]))
            else:
                fh.write("%s %.6f %.6f %.6f %.6f\n" % (s, r[0], r[1], r[2], q[i]))


GLOBULE_DENSITY = 1231 / (4.0 / 3.0 * math.pi * 18.0 * 15.0 * 13.0)  # atoms per A^3


def assemblies(count=64, n_min=500, n_max=200_000, seed=2104):
    """Descriptions of the batch targets: list of dicts with n, axes, net
    charge, probe and a per-target seed (atoms come from globule())."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        ra, rb = rng.uniform(1.0, 2.0, size=2)
        s = (n / GLOBULE_DENSITY / (4.0 / 3.0 * math.pi * ra * rb)) ** (1.0 / 3.0)
        out.append({"n": n, "axes": (s * ra, s * rb, s), "net_charge": float(max(1, round(math.sqrt(n) / 8))),
                    "probe": "He" if k % 2 == 0 else "N2", "seed": seed + 1 + k})
    return out
