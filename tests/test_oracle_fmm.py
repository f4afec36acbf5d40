"""Pins the oracle's FMM (CPU only): the softened-kernel Taylor coefficients,
exactness when everything is near field, and convergence to the direct sum."""
import numpy as np
import pytest

import oracle
from solveit_hip import synth


def _multi(p):
    out = []
    for n in range(p + 1):
        for a in range(n, -1, -1):
            for b in range(n - a, -1, -1):
                out.append((a, b, n - a - b))
    return out


@pytest.mark.parametrize("eps2", [0.0, 0.3])
def test_taylor_coefficients_reproduce_softened_kernel(eps2):
    p = 8
    R = np.array([2.0, -1.5, 3.0])
    T = oracle.fmm_taylor(p, R, eps2)
    mi = _multi(p)
    G = lambda r: 1.0 / np.sqrt(r @ r + eps2)
    for s in (0.2, 0.05):
        w = np.array([0.3, 0.2, -0.4]) * s
        approx = sum(T[k] * w[0] ** a * w[1] ** b * w[2] ** c for k, (a, b, c) in enumerate(mi))
        # truncation error ~ (|w|/|R|)^(p+1)
        bound = 10 * (np.linalg.norm(w) / np.linalg.norm(R)) ** (p + 1) + 1e-14
        assert abs(G(R + w) - approx) < bound


def test_taylor_first_and_second_derivatives_closed_form():
    R = np.array([0.7, -0.2, 1.1])
    e2 = 0.05
    T = oracle.fmm_taylor(2, R, e2)
    s = R @ R + e2
    mi = _multi(2)
    for k, (a, b, c) in enumerate(mi):
        g = (a, b, c)
        if sum(g) == 1:
            i = g.index(1)
            assert abs(T[k] - (-R[i] * s ** -1.5)) < 1e-14
        if sum(g) == 2 and 2 in g:
            i = g.index(2)
            assert abs(T[k] - 0.5 * (3 * R[i] ** 2 * s ** -2.5 - s ** -1.5)) < 1e-14


def test_all_near_field_equals_direct():
    d = synth.cone(600, seed=3)
    ref = oracle.direct_f64(d)
    a, st = oracle.fmm_f64(d, 4, 1, ring=2)   # 2x2x2 leaves, all within ring
    assert st["m2l"] == 0
    assert oracle.rel_l2(a, ref)[0] < 1e-13


def test_fmm_converges_to_direct():
    d = synth.cone(8192, seed=4)
    ref = oracle.direct_f64(d)
    errs = []
    for p in (4, 6, 8):
        a, st = oracle.fmm_f64(d, p, 5, ring=2)
        assert st["m2l"] > 0
        errs.append(oracle.rel_l2(a, ref)[0])
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 2e-5


def test_shell_orders():
    """Per-shell truncation: orders equal to p change nothing; the engine's
    reduced orders (p, p-2, p-3 on shells 3/4/5) stay within 1e-5 of the direct
    sum at p=10 and are less accurate than the uniform order."""
    d = synth.cone(8192, seed=5)
    ref = oracle.direct_f64(d)
    base, _ = oracle.fmm_f64(d, 10, 5, ring=2)
    try:
        oracle.fmm_set_shell_orders([10] * 8)
        same, _ = oracle.fmm_f64(d, 10, 5, ring=2)
        np.testing.assert_array_equal(same, base)
        oracle.fmm_set_shell_orders([0, 0, 0, 10, 8, 7])
        red, _ = oracle.fmm_f64(d, 10, 5, ring=2)
        oracle.fmm_set_shell_orders([0, 0, 0, 10, 2, 2])
        crude, _ = oracle.fmm_f64(d, 10, 5, ring=2)
    finally:
        oracle.fmm_set_shell_orders(None)
    e_base, e_red, e_crude = (oracle.rel_l2(a, ref)[0] for a in (base, red, crude))
    assert e_base <= e_red < 1e-5
    assert e_crude > 3 * e_red


def test_reduced_shell_orders_of_lower_orders():
    """The engine's shell orders below p = 10 lose little against the uniform
    order, where (p, p-2, p-3) would lose much more (8192-ion cone, leaf
    level 5, near field of two cells)."""
    d = synth.cone(8192, seed=5)
    ref = oracle.direct_f64(d)
    try:
        for p, used, cut in ((8, [8, 7, 6], [8, 6, 5]), (6, [6, 5, 4], [6, 4, 3]), (4, [4, 4, 4], [4, 2, 1])):
            oracle.fmm_set_shell_orders(None)
            e_full = oracle.rel_l2(oracle.fmm_f64(d, p, 5, ring=2)[0], ref)[0]
            oracle.fmm_set_shell_orders([0, 0, 0] + used)
            e_used = oracle.rel_l2(oracle.fmm_f64(d, p, 5, ring=2)[0], ref)[0]
            oracle.fmm_set_shell_orders([0, 0, 0] + cut)
            e_cut = oracle.rel_l2(oracle.fmm_f64(d, p, 5, ring=2)[0], ref)[0]
            assert e_used <= 1.6 * e_full, (p, e_used, e_full)
            assert e_cut > 1.5 * e_used, (p, e_cut, e_used)
    finally:
        oracle.fmm_set_shell_orders(None)
