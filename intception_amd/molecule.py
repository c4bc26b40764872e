"""Molecule-like screened workloads (configs 3 and 4).

Geometry: N atoms by random sequential addition in a sphere with a minimum
spacing of 1.8 bohr.  Basis per atom: shells s, s, p, p, d with K = 4, 1, 3, 1, 1
and even-tempered exponents alpha_k = alpha_max / 3^k, alpha_max ~ U[20, 2000]
per atom, coefficients U[0.1, 1] then normalised.  Choices the workload
definition leaves open, recorded here and in every batch's meta:

* sphere radius: 12.0 bohr for 200 atoms, scaled by (N / 200)^(1/3);
* exponent assignment: the contracted s takes alpha_0..alpha_3, the second s
  alpha_4, the contracted p alpha_2..alpha_4, the second p alpha_5 and the d
  alpha_4 of the atom's sequence.

Screening: Q_P = sqrt(max |(ab|ab)|) per unique shell pair P = (a <= b) from the
device FP64 path; the unique quartets {P, P'} with Q_P Q_P' >= tau are counted
exactly (pairs sorted by Q, prefix counts) and sampled uniformly.
"""
import numpy as np

from . import basis

SHELLS_PER_ATOM = [(0, [0, 1, 2, 3]), (0, [4]), (1, [2, 3, 4]), (1, [5]), (2, [4])]


def pack_atoms(n_atoms, radius, min_dist=1.8, seed=0, max_tries=2_000_000):
    rng = np.random.Generator(np.random.PCG64(seed))
    pts = np.empty((n_atoms, 3))
    n = 0
    tries = 0
    while n < n_atoms:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("random sequential packing did not converge; enlarge the sphere")
        p = rng.uniform(-radius, radius, size=3)
        if p @ p > radius * radius:
            continue
        if n and np.min(np.sum((pts[:n] - p) ** 2, axis=1)) < min_dist * min_dist:
            continue
        pts[n] = p
        n += 1
    return pts


def molecule_shells(atoms, seed=0, shells_per_atom=SHELLS_PER_ATOM, amax_range=(20.0, 2000.0)):
    """Shell table for the molecule: (ls, nprims, off, cen, exps, coefs, atom)."""
    rng = np.random.Generator(np.random.PCG64(seed + 17))
    ls, nps, cen, ex, raw, atom = [], [], [], [], [], []
    for ia, xyz in enumerate(atoms):
        amax = rng.uniform(*amax_range)
        seq = amax / 3.0 ** np.arange(8)
        for l, ks in shells_per_atom:
            ls.append(l)
            nps.append(len(ks))
            cen.append(xyz)
            ex.append(seq[ks])
            raw.append(rng.uniform(0.1, 1.0, size=len(ks)))
            atom.append(ia)
    coefs = [basis.normalize_contraction(e, c, l) for e, c, l in zip(ex, raw, ls)]
    nps = np.array(nps, dtype=np.int32)
    off = np.zeros(len(nps), dtype=np.int32)
    off[1:] = np.cumsum(nps)[:-1]
    return (np.array(ls, dtype=np.int32), nps, off, np.array(cen), np.concatenate(ex), np.concatenate(coefs),
            np.array(atom, dtype=np.int32))


class Molecule:
    def __init__(self, n_atoms=200, seed=202603, radius=None, spherical=True):
        if radius is None:
            radius = 12.0 * (n_atoms / 200.0) ** (1.0 / 3.0)
        self.n_atoms, self.seed, self.radius, self.spherical = n_atoms, seed, radius, spherical
        self.atoms = pack_atoms(n_atoms, radius, 1.8, seed)
        (self.shell_l, self.shell_nprim, self.shell_off, self.shell_cen, self.exps, self.coefs,
         self.shell_atom) = molecule_shells(self.atoms, seed)
        self.n_shell = len(self.shell_l)
        self.pairs = None
        self.Q = None

    def meta(self):
        return {"n_atoms": self.n_atoms, "seed": self.seed, "radius_bohr": self.radius, "min_spacing_bohr": 1.8,
                "shells": "s,s,p,p,d K=4,1,3,1,1; even-tempered ratio 3, alpha_max U[20,2000]; "
                          "s:a0-a3, s:a4, p:a2-a4, p:a5, d:a4"}

    def batch(self, quartets):
        return basis.QuartetBatch(self.spherical, self.shell_l, self.shell_nprim, self.shell_off, self.shell_cen,
                                  self.exps, self.coefs, np.ascontiguousarray(quartets, dtype=np.int32),
                                  meta=self.meta())

    def pair_batch(self, pp):
        """Batch over shared shell pairs: pp (n, 2) indices into self.pairs."""
        pp = np.ascontiguousarray(pp, dtype=np.int32)
        q = quartets_from_pairs(self.pairs, pp)
        return basis.QuartetBatch(self.spherical, self.shell_l, self.shell_nprim, self.shell_off, self.shell_cen,
                                  self.exps, self.coefs, q, meta=self.meta(),
                                  pair_shells=np.ascontiguousarray(self.pairs, dtype=np.int32), quartet_pairs=pp)

    def screened_quartets(self, tau=1e-12):
        """Every unique quartet {P, P'} with Q_P Q_P' >= tau, as pair indices."""
        order, n_p, count = screened_counts(self.Q, tau)
        total = int(count.sum())
        first = np.repeat(np.arange(len(order)), count)
        start = np.repeat(np.arange(len(order)) - np.concatenate([[0], np.cumsum(count)[:-1]]), count)
        second = np.arange(total) + start
        return np.stack([order[first], order[second]], axis=1).astype(np.int32)

    def all_pairs(self):
        i, j = np.triu_indices(self.n_shell)
        return np.stack([i, j], axis=1).astype(np.int32)

    def prescreen_pairs(self, cutoff=1e-16):
        """Pairs whose Gaussian-product prefactor bound exceeds `cutoff`
        (the most diffuse primitives of both shells set the bound)."""
        P = self.all_pairs()
        a, b = P[:, 0], P[:, 1]
        emin = np.array([self.exps[o:o + n].min() for o, n in zip(self.shell_off, self.shell_nprim)])
        r2 = np.sum((self.shell_cen[a] - self.shell_cen[b]) ** 2, axis=1)
        mu = emin[a] * emin[b] / (emin[a] + emin[b])
        keep = np.exp(-mu * r2) > cutoff
        return P[keep]

    def schwarz(self, pairs=None, device="cuda"):
        """Q_P = sqrt(max |(ab|ab)|) for every pair, FP64 on the device."""
        from . import runtime
        if pairs is None:
            pairs = self.prescreen_pairs()
        q = np.stack([pairs[:, 0], pairs[:, 1], pairs[:, 0], pairs[:, 1]], axis=1)
        b = self.batch(q)
        d = runtime.compute(b, device)
        sz = b.sizes()
        mx = np.maximum.reduceat(np.abs(d), b.out_off.astype(np.int64)) if len(d) else np.zeros(0)
        # reduceat over contiguous blocks
        self.pairs = pairs
        self.Q = np.sqrt(mx[:len(sz)])
        return self.Q


def screened_counts(Q, tau):
    """Sort pairs by Q descending; for sorted position p the partners p' >= p
    with Q_p Q_p' >= tau are positions [p, n_p).  Returns (order, n_p, count_p)."""
    order = np.argsort(-Q, kind="stable")
    Qs = Q[order]
    # n_p = number of sorted entries with Q >= tau / Q_p
    thr = tau / np.maximum(Qs, 1e-300)
    n_p = np.searchsorted(-Qs, -thr, side="right")
    count = np.maximum(0, n_p - np.arange(len(Qs)))
    return order, n_p, count


def sample_screened(Q, tau, n_sample, seed=0):
    """Uniform sample (with replacement) of unique quartets {P, P'} passing
    Q_P Q_P' >= tau; returns pair-index pairs (into the pair list) and the
    exact total count."""
    order, n_p, count = screened_counts(Q, tau)
    total = int(count.sum())
    if total == 0:
        return np.zeros((0, 2), dtype=np.int64), 0
    rng = np.random.Generator(np.random.PCG64(seed))
    cum = np.cumsum(count)
    r = rng.integers(0, total, size=n_sample)
    p = np.searchsorted(cum, r, side="right")
    within = r - (cum[p] - count[p])
    pp = p + within
    return np.stack([order[p], order[pp]], axis=1), total


def quartets_from_pairs(pairs, pp):
    """(ab|cd) shell quartets from pair-of-pair indices."""
    a = pairs[pp[:, 0]]
    c = pairs[pp[:, 1]]
    return np.concatenate([a, c], axis=1).astype(np.int32)
