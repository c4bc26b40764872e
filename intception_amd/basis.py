"""Shells, normalisation, Cartesian->spherical matrices and seeded workloads.

The reference takes contraction coefficients and spherical transformation
matrices from the caller (``contract_array_X``, ``sph_trans_array_X``:
src/intception/generator/wrappers.py:75-85 and
src/intception/generator/callables/main_source_algorithms/contracted_general.py:57-65);
it ships neither.  This module is the caller side: it folds primitive and
contracted normalisation into the coefficients (x^l convention: every Cartesian
component carries the normalisation of x^l) and builds real-solid-harmonic
matrices S (ncc(l) x (2l+1), used as S[icart + isph*ncart]) normalised with
respect to those Cartesians.  The same arrays feed the HIP path and the oracle,
so both compute from identical inputs.

Cartesian order within a shell is x^l first ... z^l last
(src/intception/dsl.py:988-1001); spherical order is m = -l..l.
"""
from dataclasses import dataclass, field
from math import comb, factorial, pi, sqrt

import numpy as np


def ncc(l):
    return (l + 1) * (l + 2) // 2


def ncc0(l):
    return 0 if l < 0 else (l + 1) * (l + 2) * (l + 3) // 6


def nsph(l):
    return 2 * l + 1


def cart_components(l):
    return [(i, j, l - i - j) for i in range(l, -1, -1) for j in range(l - i, -1, -1)]


def _dfact(n):
    """(n)!! with (-1)!! = 1."""
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def primitive_norm(alpha, l):
    """Normalisation of x^l exp(-alpha r^2)."""
    alpha = np.asarray(alpha, dtype=np.float64)
    return np.sqrt((2.0 * alpha / pi) ** 1.5 * (4.0 * alpha) ** l / _dfact(2 * l - 1))


def normalize_contraction(exps, coefs, l):
    """Coefficients with primitive and contracted normalisation folded in."""
    exps = np.asarray(exps, dtype=np.float64)
    c = np.asarray(coefs, dtype=np.float64) * primitive_norm(exps, l)
    p = exps[:, None] + exps[None, :]
    ov = _dfact(2 * l - 1) / (2.0 * p) ** l * (pi / p) ** 1.5
    norm2 = float(c @ ov @ c)
    return c / sqrt(norm2)


def _cart_overlap(l):
    """<m_k|m_k'> for the Cartesian monomials of shell l, each scaled by the
    x^l normalisation (so <x^l|x^l> = 1)."""
    comps = cart_components(l)
    n = len(comps)
    ov = np.zeros((n, n))
    for i, a in enumerate(comps):
        for j, b in enumerate(comps):
            s = [a[k] + b[k] for k in range(3)]
            if any(v % 2 for v in s):
                continue
            ov[i, j] = _dfact(s[0] - 1) * _dfact(s[1] - 1) * _dfact(s[2] - 1) / _dfact(2 * l - 1)
    return ov


def sph_matrix(l):
    """S[icart, isph] (ncc(l) x (2l+1)) for real solid harmonics, m = -l..l.

    Coefficients from the closed form of the real solid harmonics
    (Helgaker, Jorgensen, Olsen, eq. 6.4.47-6.4.52; Schlegel & Frisch,
    IJQC 54, 83 (1995)); each column is then normalised under the Cartesian
    overlap of x^l-normalised monomials."""
    comps = cart_components(l)
    index = {c: i for i, c in enumerate(comps)}
    S = np.zeros((len(comps), 2 * l + 1))
    ov = _cart_overlap(l)
    for col, m in enumerate(range(-l, l + 1)):
        am = abs(m)
        two_vm = 0 if m >= 0 else 1          # 2 v_m
        for t in range((l - am) // 2 + 1):
            for u in range(t + 1):
                # v runs over v_m, v_m + 1, ... with 2v <= |m|
                for two_v in range(two_vm, am + 1, 2):
                    sign = (-1) ** (t + (two_v - two_vm) // 2)
                    coef = (sign * 0.25 ** t * comb(l, t) * comb(l - t, am + t) *
                            comb(t, u) * comb(am, two_v))
                    # monomial x^(2t+|m|-2(u+v)) y^(2(u+v)) z^(l-2t-|m|)
                    ex = 2 * t + am - 2 * u - two_v
                    ey = 2 * u + two_v
                    ez = l - 2 * t - am
                    if ex >= 0 and ey >= 0 and ez >= 0:
                        S[index[(ex, ey, ez)], col] += coef
        nrm = float(S[:, col] @ ov @ S[:, col])
        S[:, col] /= sqrt(nrm)
    return S


SPH = {l: sph_matrix(l) for l in range(7)}


@dataclass
class QuartetBatch:
    """A batch of contracted shell quartets over an SoA shell table.

    shell_l, shell_nprim, shell_off: int arrays (n_shell); shell_cen:
    (n_shell, 3); exps / coefs: concatenated primitive data, shell s owning
    [shell_off[s], shell_off[s] + shell_nprim[s]); quartets: (n_quartet, 4)
    shell ids (a, b, c, d); out_off: int64 offset of each quartet's block in the
    output; each block is row-major [a][b][c][d] over Cartesian (ncc) or
    spherical (2l+1) components."""
    spherical: bool
    shell_l: np.ndarray
    shell_nprim: np.ndarray
    shell_off: np.ndarray
    shell_cen: np.ndarray
    exps: np.ndarray
    coefs: np.ndarray
    quartets: np.ndarray
    out_off: np.ndarray = None
    out_size: int = 0
    meta: dict = field(default_factory=dict)
    # optional shared shell pairs: pair_shells (n_pair, 2), quartet_pairs
    # (n_quartet, 2); then quartets[q] = pair_shells[quartet_pairs[q]].ravel()
    pair_shells: np.ndarray = None
    quartet_pairs: np.ndarray = None

    def __post_init__(self):
        if self.out_off is None:
            self.set_dense_output()

    def ncomp(self, l):
        return nsph(l) if self.spherical else ncc(l)

    def sizes(self):
        nc = np.array([self.ncomp(l) for l in range(16)], dtype=np.int64)
        ls = self.shell_l[self.quartets]
        return nc[ls].prod(axis=1)

    def set_dense_output(self):
        sz = self.sizes()
        self.out_off = np.zeros(len(sz), dtype=np.int64)
        if len(sz):
            self.out_off[1:] = np.cumsum(sz)[:-1]
        self.out_size = int(sz.sum())

    @property
    def n_quartet(self):
        return len(self.quartets)

    def classes(self):
        return self.shell_l[self.quartets]

    def shell_dict(self, s):
        o, n = int(self.shell_off[s]), int(self.shell_nprim[s])
        return {"l": int(self.shell_l[s]), "exps": self.exps[o:o + n].copy(),
                "coefs": self.coefs[o:o + n].copy(), "cen": self.shell_cen[s].copy()}

    def quartet_shells(self, q):
        return [self.shell_dict(int(s)) for s in self.quartets[q]]

    def subset(self, idx):
        """Batch with the given quartets (same shell table, dense output)."""
        qp = None if self.quartet_pairs is None else np.ascontiguousarray(self.quartet_pairs[idx])
        return QuartetBatch(self.spherical, self.shell_l, self.shell_nprim, self.shell_off,
                            self.shell_cen, self.exps, self.coefs,
                            np.ascontiguousarray(self.quartets[idx]), meta=dict(self.meta),
                            pair_shells=self.pair_shells, quartet_pairs=qp)


def random_shells(rng, ls, nprims, box=10.0, log_alpha=(-1.0, 2.0)):
    """Independent random shells: centre U[0,box)^3, exponents 10^U[log_alpha),
    coefficients U[0.1,1] then normalised."""
    ls = np.asarray(ls, dtype=np.int32)
    nprims = np.asarray(nprims, dtype=np.int32)
    n = len(ls)
    off = np.zeros(n, dtype=np.int32)
    if n:
        off[1:] = np.cumsum(nprims)[:-1]
    tot = int(nprims.sum())
    cen = rng.uniform(0.0, box, size=(n, 3))
    exps = 10.0 ** rng.uniform(log_alpha[0], log_alpha[1], size=tot)
    raw = rng.uniform(0.1, 1.0, size=tot)
    coefs = np.empty(tot)
    # vectorised normalisation grouped by (l, nprim)
    for l in np.unique(ls):
        for k in np.unique(nprims[ls == l]):
            sel = np.nonzero((ls == l) & (nprims == k))[0]
            idx = off[sel][:, None] + np.arange(k)[None, :]
            e = exps[idx]
            c = raw[idx] * primitive_norm(e, int(l))
            p = e[:, :, None] + e[:, None, :]
            ov = _dfact(2 * int(l) - 1) / (2.0 * p) ** int(l) * (pi / p) ** 1.5
            nrm = np.einsum("si,sij,sj->s", c, ov, c)
            coefs[idx] = c / np.sqrt(nrm)[:, None]
    return ls, nprims, off, cen, exps, coefs


def make_random_quartets(seed, classes, nprim_choices, spherical, box=10.0, log_alpha=(-1.0, 2.0)):
    """One quartet per row of ``classes`` (n x 4 angular momenta), each with
    four independent shells; per-shell K drawn uniformly from nprim_choices."""
    rng = np.random.Generator(np.random.PCG64(seed))
    classes = np.asarray(classes, dtype=np.int32)
    nq = len(classes)
    ls = classes.reshape(-1)
    nprims = rng.choice(np.asarray(nprim_choices, dtype=np.int32), size=len(ls))
    ls, nprims, off, cen, exps, coefs = random_shells(rng, ls, nprims, box, log_alpha)
    quartets = np.arange(4 * nq, dtype=np.int32).reshape(nq, 4)
    return QuartetBatch(bool(spherical), ls, nprims, off, cen, exps, coefs, quartets,
                        meta={"seed": seed})


def batch_from_shells(quartet_shells, spherical):
    """QuartetBatch from a list of quartets, each a list of 4 shell dicts
    (l, exps, coefs, cen); every shell gets its own table entry."""
    ls, nps, cen, ex, co = [], [], [], [], []
    for sh in quartet_shells:
        for s in sh:
            ls.append(int(s["l"]))
            nps.append(len(s["exps"]))
            cen.append(np.asarray(s["cen"], dtype=np.float64))
            ex.append(np.asarray(s["exps"], dtype=np.float64))
            co.append(np.asarray(s["coefs"], dtype=np.float64).reshape(-1))
    nps = np.array(nps, dtype=np.int32)
    off = np.zeros(len(nps), dtype=np.int32)
    off[1:] = np.cumsum(nps)[:-1]
    q = np.arange(len(ls), dtype=np.int32).reshape(-1, 4)
    return QuartetBatch(bool(spherical), np.array(ls, dtype=np.int32), nps, off, np.array(cen),
                        np.concatenate(ex), np.concatenate(co), q)


def unit_s_shell():
    """The unit s-shell of three-centre integrals (ab|P) = (ab|P s): exponent 0,
    coefficient 1, so the ket pair has eta = exponent of P, Q = C, K_CD = 1."""
    return 0, 1, np.array([0.0]), np.array([1.0])


def random_three_center(seed, n, lmax_orb=2, lmax_aux=4, spherical=True, box=10.0):
    """n quartets (a b | P s): orbital shells l <= lmax_orb, K in 1..6,
    exponents 10^U[-1,2); auxiliary shells l <= lmax_aux, K = 1, exponents
    10^U[log .05, log 200); one shared unit s-shell."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo = rng.integers(0, lmax_orb + 1, size=2 * n)
    ko = rng.integers(1, 7, size=2 * n)
    so = random_shells(rng, lo, ko, box, (-1.0, 2.0))
    la = rng.integers(0, lmax_aux + 1, size=n)
    sa = random_shells(rng, la, np.ones(n, dtype=np.int32), box, (np.log10(0.05), np.log10(200.0)))
    ls = np.concatenate([so[0], sa[0], [0]]).astype(np.int32)
    nprims = np.concatenate([so[1], sa[1], [1]]).astype(np.int32)
    cen = np.concatenate([so[3], sa[3], np.zeros((1, 3))])
    exps = np.concatenate([so[4], sa[4], [0.0]])
    coefs = np.concatenate([so[5], sa[5], [1.0]])
    off = np.zeros(len(ls), dtype=np.int32)
    off[1:] = np.cumsum(nprims)[:-1]
    unit = len(ls) - 1
    q = np.stack([np.arange(0, 2 * n, 2), np.arange(1, 2 * n, 2), 2 * n + np.arange(n),
                  np.full(n, unit)], axis=1).astype(np.int32)
    return QuartetBatch(bool(spherical), ls, nprims, off, cen, exps, coefs, q, meta={"seed": seed})


CONFIG0_CLASSES = [(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (1, 1, 1, 1)]
CONFIG2_CLASSES = [(2, 2, 2, 2), (3, 2, 3, 2), (3, 3, 2, 2), (3, 3, 3, 3)]


def config0(n_per_class=10000, seed=202600):
    """Cartesian (ss|ss),(ps|ss),(ps|ps),(pp|ss),(pp|pp); K = 3; 10 bohr cube."""
    cls = np.repeat(np.array(CONFIG0_CLASSES, dtype=np.int32), n_per_class, axis=0)
    b = make_random_quartets(seed, cls, [3], spherical=False)
    b.meta["config"] = 0
    return b


def config1(n_quartet=1_000_000, seed=202601):
    """81 classes over {s,p,d}^4 drawn uniformly, K per shell U{1..6}, spherical."""
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    cls = rng.integers(0, 3, size=(n_quartet, 4), dtype=np.int32)
    b = make_random_quartets(seed, cls, [1, 2, 3, 4, 5, 6], spherical=True)
    b.meta["config"] = 1
    return b


def config2(n_quartet=200_000, seed=202602):
    """(dd|dd),(fd|fd),(ff|dd),(ff|ff), K U{1..3}, exponents 10^U[log .05, log 50)."""
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    pick = rng.integers(0, 4, size=n_quartet)
    cls = np.array(CONFIG2_CLASSES, dtype=np.int32)[pick]
    b = make_random_quartets(seed, cls, [1, 2, 3], spherical=True,
                             log_alpha=(np.log10(0.05), np.log10(50.0)))
    b.meta["config"] = 2
    return b
