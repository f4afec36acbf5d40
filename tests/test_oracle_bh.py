"""Pins the oracle's Barnes-Hut restatement (CPU only): closed forms for two
ions and for the reference's centre-of-charge rule, theta -> 0 against the
direct sum, accuracy falling monotonically with theta, and the symmetry
variants at theta = 0 against the direct sums with images."""
import numpy as np
import pytest

import oracle
from solveit_hip import synth


def _aos(x, q, status=None, species=None):
    n = len(q)
    pos4 = np.zeros((n, 4), np.float32)
    pos4[:, :3] = np.asarray(x, np.float32)
    pos4[:, 3] = q
    vel4 = np.zeros((n, 4), np.float32)
    vel4[:, 3] = 0 if species is None else species
    acc4 = np.zeros((n, 4), np.float32)
    acc4[:, 3] = 1 if status is None else status
    return pos4, vel4, acc4


def test_two_ions_closed_form():
    x = [[0.0, 0.0, 0.0], [3e-9, 4e-9, 0.0]]
    pos4, vel4, acc4 = _aos(x, [1.0, 1.0])
    out, ncell = oracle.bh_ref(pos4, vel4, acc4, theta=0.5)
    kc = oracle.k_const() / 111.0
    r = np.array(x[0], np.float32) - np.array(x[1], np.float32)
    d = np.sqrt(float(np.float32(r @ r)) + 6.4e-19)
    expect = kc * r / d ** 3
    np.testing.assert_allclose(out[0], expect, rtol=1e-6)
    np.testing.assert_allclose(out[1], -expect, rtol=1e-6)
    assert ncell == 9  # root split once into 8 children


def test_centre_of_charge_rule():
    """Root cube [0, 8 nm]^3: its child 0 holds a +1/-1 pair near the origin
    (net charge 0) and a +1 ion at b = 3 nm; the target sits at 8 nm.  The
    reference weights child centres by |net charge| over sum |q|, so child 0's
    centre is (1 x b) / 3; at theta = 0.7 the target opens the root and
    accepts child 0 as one monopole (its own leaf is skipped)."""
    b = np.array([3e-9, 3e-9, 3e-9])
    far = np.array([8e-9, 8e-9, 8e-9])
    x = [np.zeros(3), np.full(3, 1e-10), b, far]
    pos4, vel4, acc4 = _aos(x, [1.0, -1.0, 1.0, 1.0], status=[2, 2, 2, 1])
    out, _ = oracle.bh_ref(pos4, vel4, acc4, theta=0.7)
    c = (np.float32(1.0) * b.astype(np.float32)) / np.float32(3.0)
    r = far.astype(np.float32) - c
    d = np.float32(np.sqrt(float(np.float32(r @ r)) + 6.4e-19))
    expect = oracle.k_const() / 111.0 * r / np.float64(d) ** 3
    np.testing.assert_allclose(out[3], expect, rtol=1e-6)
    # and not the |q|-weighted mean of the three ions
    wrong = far - (0 + 1e-10 + 3e-9) / 3.0
    assert abs(out[3][0] - oracle.k_const() / 111.0 * wrong[0] / np.linalg.norm(wrong) ** 3) > 1e-3 * abs(out[3][0])


def test_theta_zero_is_direct():
    d = synth.cone(2500, seed=31)
    d["status"][::17] = 2
    d["status"][::41] = 0
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    out, _ = oracle.bh_ref(pos4, vel4, acc4, theta=0.0)
    ref = oracle.direct_ref(pos4, vel4, acc4)
    m = d["status"] == 1
    assert np.all(out[~m] == 0)
    l2, mx, _ = oracle.rel_l2(out, ref, m)
    assert l2 < 5e-5 and mx < 1e-3  # float accumulation of the same pair terms


def test_accuracy_monotone_in_theta():
    d = synth.cone(3000, seed=32)
    d["q"][:] = 1.0  # one sign: the monopole centre is the |q|-weighted mean
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    ref = oracle.direct_f64(d)
    errs = [oracle.rel_l2(oracle.bh_ref(pos4, vel4, acc4, theta=t)[0], ref)[0]
            for t in (0.2, 0.4, 0.6)]
    assert errs[0] < errs[1] < errs[2] < 2e-2


@pytest.mark.parametrize("sym", ["rotational", "translational"])
def test_symmetry_theta_zero_is_direct(sym):
    if sym == "rotational":
        d = synth.cone(1200, seed=33)
        kw = dict(sym="rotational")
    else:
        d = synth.cell_fill(800, seed=34)
        kw = dict(sym="translational", lattice=synth.full_lattice(1, (1, 1, 1)),
                  domain=(2e-6, 2e-6, 2e-6))
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    out, _ = oracle.bh_ref(pos4, vel4, acc4, theta=0.0, **kw)
    ref = oracle.direct_ref(pos4, vel4, acc4, **kw)
    l2, mx, _ = oracle.rel_l2(out, ref, d["status"] == 1)
    assert l2 < 5e-5


def test_root_contains_sources():
    d = synth.cone(500, seed=35)
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    r = oracle.bh_root(pos4, acc4)
    m = (acc4[:, 3] == 1) | (acc4[:, 3] == 2)
    assert r[3] > 0
    assert np.all(pos4[m, :3] >= r[:3] - 1e-12) and np.all(pos4[m, :3] <= r[:3] + r[3] * (1 + 1e-6))
