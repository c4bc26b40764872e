"""Molecule-like screened workloads (configs 3 and 4).

Config 4 (`ThreeCenterMolecule`): three-centre (ab|P) integrals as quartets
(ab|P s) with a unit s-shell (exponent 0) in the fourth slot; see its
docstring for the basis.

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

    def sample_screened(self, tau, n, seed=0):
        """Uniform sample of the screened quartets (pair-index pairs) and the
        exact screened count."""
        return sample_screened(self.Q, tau, n, seed=seed)

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


AUX_L = (0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 4)


class ThreeCenterMolecule(Molecule):
    """Config 4: (ab|P s) quartets of a molecule, P an auxiliary shell and s a
    unit s-shell (exponent 0, coefficient 1) at P's centre.

    Geometry as config 3, scaled to the atom count (random sequential packing,
    radius 12 bohr (N/200)^(1/3), spacing >= 1.8 bohr).  Orbital shells per atom
    s, s, p, p, d with contraction depth K ~ U{1..6} each and even-tempered
    exponents (ratio 3, alpha_max ~ U[20, 2000]; shell k takes K consecutive
    exponents of the atom's sequence starting at 0, 4, 2, 5, 4 as in config 3).
    Auxiliary shells: 12 per atom with l = 0,0,0,1,1,1,2,2,2,3,3,4, K = 1,
    exponent 10^U[log10 0.05, log10 200], normalised.  Pairs: the orbital
    pairs a <= b that pass the Gaussian-product prescreen, then one pair
    (P, s_P) per auxiliary shell.  Screening: Q_ab Q_P >= tau with
    Q = sqrt(max |(pair|pair)|) from the FP64 path."""

    def __init__(self, n_atoms=100, seed=202604, radius=None, spherical=True, n_aux=12):
        if radius is None:
            radius = 12.0 * (n_atoms / 200.0) ** (1.0 / 3.0)
        self.n_atoms, self.seed, self.radius, self.spherical = n_atoms, seed, radius, spherical
        self.atoms = pack_atoms(n_atoms, radius, 1.8, seed)
        rng = np.random.Generator(np.random.PCG64(seed + 29))
        ls, nps, cen, ex, co = [], [], [], [], []
        starts = [0, 4, 2, 5, 4]
        kinds = []   # 0 orbital, 1 auxiliary, 2 unit
        for xyz in self.atoms:
            seq = rng.uniform(20.0, 2000.0) / 3.0 ** np.arange(14)
            for (l, _), s0 in zip(SHELLS_PER_ATOM, starts):
                k = int(rng.integers(1, 7))
                e = seq[s0:s0 + k]
                ls.append(l); nps.append(k); cen.append(xyz); ex.append(e)
                co.append(basis.normalize_contraction(e, rng.uniform(0.1, 1.0, size=k), l))
                kinds.append(0)
        for xyz in self.atoms:
            for l in AUX_L[:n_aux]:
                e = 10.0 ** rng.uniform(np.log10(0.05), np.log10(200.0), size=1)
                ls.append(l); nps.append(1); cen.append(xyz); ex.append(e)
                co.append(basis.normalize_contraction(e, [1.0], l))
                kinds.append(1)
                ls.append(0); nps.append(1); cen.append(xyz); ex.append(np.zeros(1)); co.append(np.ones(1))
                kinds.append(2)
        self.shell_l = np.array(ls, dtype=np.int32)
        self.shell_nprim = np.array(nps, dtype=np.int32)
        self.shell_off = np.zeros(len(nps), dtype=np.int32)
        self.shell_off[1:] = np.cumsum(self.shell_nprim)[:-1]
        self.shell_cen = np.array(cen)
        self.exps = np.concatenate(ex)
        self.coefs = np.concatenate(co)
        self.shell_kind = np.array(kinds, dtype=np.int32)
        self.n_shell = len(ls)
        self.pairs = None
        self.Q = None
        self.n_orb_pair = 0

    def meta(self):
        return {"n_atoms": self.n_atoms, "seed": self.seed, "radius_bohr": self.radius, "min_spacing_bohr": 1.8,
                "shells": "orbital s,s,p,p,d K~U{1..6} even-tempered ratio 3, alpha_max U[20,2000]; auxiliary "
                          "12/atom l=0,0,0,1,1,1,2,2,2,3,3,4 K=1 alpha 10^U[log .05, log 200]; unit s in slot 4"}

    def prescreen_pairs(self, cutoff=1e-16):
        orb = np.nonzero(self.shell_kind == 0)[0]
        i, j = np.triu_indices(len(orb))
        P = np.stack([orb[i], orb[j]], axis=1).astype(np.int32)
        a, b = P[:, 0], P[:, 1]
        emin = np.array([self.exps[o:o + n].min() for o, n in zip(self.shell_off, self.shell_nprim)])
        r2 = np.sum((self.shell_cen[a] - self.shell_cen[b]) ** 2, axis=1)
        mu = emin[a] * emin[b] / (emin[a] + emin[b])
        P = P[np.exp(-mu * r2) > cutoff]
        aux = np.nonzero(self.shell_kind == 1)[0]
        A = np.stack([aux, aux + 1], axis=1).astype(np.int32)   # (P, its unit s-shell)
        self.n_orb_pair = len(P)
        return np.concatenate([P, A])

    def _split(self):
        no = self.n_orb_pair
        return self.Q[:no], self.Q[no:]

    def screened_counts(self, tau):
        """Aux pairs sorted by Q descending; orbital pair i keeps the first
        n_i of them (Q_i Q_P >= tau)."""
        Qo, Qa = self._split()
        order = np.argsort(-Qa, kind="stable")
        Qs = Qa[order]
        thr = tau / np.maximum(Qo, 1e-300)
        n_i = np.searchsorted(-Qs, -thr, side="right")
        return order, n_i

    def screened_quartets(self, tau=1e-12):
        """Every (orbital pair | auxiliary pair) with Q_ab Q_P >= tau, as pair indices."""
        order, n_i = self.screened_counts(tau)
        total = int(n_i.sum())
        first = np.repeat(np.arange(len(n_i)), n_i)
        start = np.concatenate([[0], np.cumsum(n_i)[:-1]])
        within = np.arange(total) - np.repeat(start, n_i)
        return np.stack([first, self.n_orb_pair + order[within]], axis=1).astype(np.int32)

    def sample_screened(self, tau, n, seed=0):
        order, n_i = self.screened_counts(tau)
        total = int(n_i.sum())
        if total == 0:
            return np.zeros((0, 2), dtype=np.int64), 0
        rng = np.random.Generator(np.random.PCG64(seed))
        cum = np.cumsum(n_i)
        r = rng.integers(0, total, size=n)
        i = np.searchsorted(cum, r, side="right")
        within = r - (cum[i] - n_i[i])
        return np.stack([i, self.n_orb_pair + order[within]], axis=1), total
