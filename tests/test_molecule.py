"""Molecule workloads (configs 3 and 4) on the CPU: packing, shell tables and
the exact Schwarz-screened quartet sets (counts, enumeration and sampling)
against brute force on small molecules with synthetic Schwarz factors."""
import numpy as np

from intception_amd import molecule


def _fake_schwarz(mol, seed):
    mol.pairs = mol.prescreen_pairs()
    mol.Q = 10.0 ** np.random.default_rng(seed).uniform(-9, 0, len(mol.pairs))


def test_packing_spacing_and_radius():
    pts = molecule.pack_atoms(60, 8.0, 1.8, seed=3)
    d = np.linalg.norm(pts[:, None] - pts[None], axis=2) + np.eye(60) * 99
    assert d.min() >= 1.8
    assert np.all(np.linalg.norm(pts, axis=1) <= 8.0)


def test_four_center_screened_set_is_exact():
    mol = molecule.Molecule(12, seed=7)
    _fake_schwarz(mol, 1)
    tau = 1e-6
    q = mol.screened_quartets(tau)
    n = len(mol.pairs)
    i, j = np.triu_indices(n)
    keep = mol.Q[i] * mol.Q[j] >= tau
    want = {(a, b) for a, b in zip(i[keep], j[keep])}
    got = {tuple(sorted(x)) for x in q.tolist()}
    assert len(q) == len(got) == len(want)
    assert got == want
    s, total = mol.sample_screened(tau, 5000, seed=4)
    assert total == len(want)
    assert np.all(mol.Q[s[:, 0]] * mol.Q[s[:, 1]] >= tau)


def test_three_center_shell_table():
    mol = molecule.ThreeCenterMolecule(6, seed=11)
    kinds = mol.shell_kind
    assert np.sum(kinds == 0) == 6 * 5
    assert np.sum(kinds == 1) == 6 * 12 and np.sum(kinds == 2) == 6 * 12
    aux = np.nonzero(kinds == 1)[0]
    # each unit s-shell follows its auxiliary shell, same centre, exponent 0
    assert np.all(kinds[aux + 1] == 2)
    assert np.allclose(mol.shell_cen[aux], mol.shell_cen[aux + 1])
    assert np.all(mol.exps[mol.shell_off[aux + 1]] == 0.0)
    assert sorted(set(mol.shell_l[aux].tolist())) == [0, 1, 2, 3, 4]
    assert np.all(mol.shell_nprim[kinds == 0] >= 1) and np.all(mol.shell_nprim[kinds == 0] <= 6)


def test_three_center_screened_set_is_exact():
    mol = molecule.ThreeCenterMolecule(8, seed=13)
    _fake_schwarz(mol, 2)
    tau = 1e-7
    q = mol.screened_quartets(tau)
    no = mol.n_orb_pair
    assert np.all(q[:, 0] < no) and np.all(q[:, 1] >= no)
    i, j = np.meshgrid(np.arange(no), np.arange(no, len(mol.pairs)), indexing="ij")
    keep = mol.Q[i] * mol.Q[j] >= tau
    want = set(zip(i[keep].tolist(), j[keep].tolist()))
    got = set(map(tuple, q.tolist()))
    assert len(q) == len(got) == len(want) and got == want
    s, total = mol.sample_screened(tau, 4000, seed=5)
    assert total == len(want)
    assert set(map(tuple, s.tolist())) <= want
    # the quartets are (orbital pair | auxiliary pair): the fourth shell is a unit s
    qs = molecule.quartets_from_pairs(mol.pairs, q[:50])
    assert np.all(mol.shell_kind[qs[:, 3]] == 2)
