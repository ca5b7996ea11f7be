"""Pins on the mobility oracle (no GPU): the equal-radius known answer
against the reference's PairwiseMobMatrix, symmetry and positive
definiteness, the Oseen far field, the numpy FP64 fixture, and the
walkthrough against the polydisperse form at a = 1."""
from pathlib import Path

import numpy as np
import pytest

import orc

GOLDEN = Path(__file__).resolve().parent / "golden"


def dense(pos, rad, mu):
    n = len(pos)
    return orc.rpy_apply(pos, rad, mu, np.eye(3 * n))


def test_equal_radius_limit_is_reference_pair_matrix():
    rng = np.random.default_rng(1)
    for _ in range(20):
        a = rng.uniform(0.5, 2.0)
        d = rng.normal(0, 1, 3)
        d *= rng.uniform(2.2, 9.0) * a / np.linalg.norm(d)
        pos = np.array([[0.0, 0.0, 0.0], d])
        M = dense(pos, np.array([a, a]), 1.0)
        block = M[0:3, 3:6] * 6 * np.pi * a          # units 1/(6 pi mu a)
        A, _ = orc.rpy_pair_ref(d / a)                 # reference: a = 1
        np.testing.assert_allclose(block, A, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(M[0:3, 0:3] * 6 * np.pi * a, np.eye(3), rtol=1e-14)


def test_reference_rotation_block_is_levi_civita():
    d = np.array([0.3, -1.1, 2.0])
    _, B = orc.rpy_pair_ref(d)
    e = d / np.linalg.norm(d)
    r = np.linalg.norm(d)
    expect = 3 / (4 * r * r) * np.array([[0, -e[2], e[1]], [e[2], 0, -e[0]], [-e[1], e[0], 0]])
    np.testing.assert_allclose(B, expect, atol=1e-14)   # -Y1NB eps(j,i,k) e_k, Y1NB = -3/(4r^2)


def test_mobility_symmetric_positive_definite():
    data = np.load(GOLDEN / "rpy_n256.npz")
    pos, rad = data["pos"][:80], data["rad"][:80]
    M = dense(pos, rad, 1.0)
    np.testing.assert_allclose(M, M.T, atol=1e-14)
    assert np.linalg.eigvalsh(M).min() > 0


def test_far_field_is_oseen():
    a = np.array([0.02, 0.03])
    for r in [10.0, 100.0]:
        d = np.array([r, 0.0, 0.0]) / np.sqrt(1.0)
        M = dense(np.array([[0.0, 0.0, 0.0], d]), a, 2.0)
        e = d / r
        oseen = (np.eye(3) + np.outer(e, e)) / (8 * np.pi * 2.0 * r)
        assert np.abs(M[0:3, 3:6] - oseen).max() < 1e-3 / (r * r) * np.abs(oseen).max() * 10


def test_oracle_matches_numpy_fixture():
    d = np.load(GOLDEN / "rpy_n256.npz")
    U = orc.rpy_apply(d["pos"], d["rad"], float(d["mu"]), d["F"].astype(np.float64))
    np.testing.assert_allclose(U, d["U"], rtol=1e-11, atol=1e-11 * np.abs(d["U"]).max())


def test_walkthrough_is_polydisperse_form_at_unit_radius():
    rng = np.random.default_rng(5)
    n = 40
    pos = rng.uniform(0, 30, (n, 3))
    f = rng.normal(0, 1, (n, 3))
    ubg = rng.normal(0, 0.1, (n, 3))
    u, w, pn = orc.rpy_walkthrough(pos, f, ubg, 0.01)
    U = orc.rpy_apply(pos, np.ones(n), 1.0, f.reshape(-1, 1)).reshape(n, 3) * 6 * np.pi
    np.testing.assert_allclose(u, U + ubg, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(pn, pos + u * 0.01, rtol=1e-14)
