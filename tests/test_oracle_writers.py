"""Pins the oracle of the output writers (CPU only).

The files are read back with restatements of the reference's own MATLAB
loaders (MATLAB_Analysis/load_particle_data.m, load_events_data.m), the
consumers whose reading order fixes the layout; the original layout, which
has no loader there, is checked against a field-by-field transcription of
saveData's fwrite sequence (integration_schemes.cpp:280-298); Efield1 is
checked against exact rational arithmetic rounded to float."""
import struct
from fractions import Fraction

import numpy as np

from oracle import writers as W

INTMAX_U32 = 0xFFFFFFFF
MASS = {0: 111, 1: 309, 2: 507, 4: 198, 5: 87, 6: 285, 7: 483}


def load_particle_data(buf):
    """load_particle_data.m: snapshots of (pos, vel, Efield1, Efield2, ind,
    species, time) from a newOriginal state file."""
    off = 0

    def rd(fmt):
        nonlocal off
        v = struct.unpack_from(fmt, buf, off)
        off += struct.calcsize(fmt)
        return v

    snaps = []
    test_val = rd("<I")[0]
    while True:
        if test_val == INTMAX_U32:
            snaps.append({"n_est": rd("<I")[0], "time": rd("<f")[0], "ind": [], "species": [], "f": []})
            snaps[-1]["ind"].append(rd("<I")[0])
            snaps[-1]["species"].append(rd("<I")[0])
        else:
            snaps[-1]["ind"].append(test_val)
            snaps[-1]["species"].append(rd("<I")[0])
        snaps[-1]["f"].append(rd("<12f"))
        if off >= len(buf):
            break
        test_val = rd("<I")[0]
    for s in snaps:
        s["f"] = np.array(s["f"], np.float32)
    return snaps


def load_events_data(buf):
    """load_events_data.m: (time, pos, vel, ind, species, event) per event."""
    out = []
    for off in range(0, len(buf), 40):
        t, = struct.unpack_from("<f", buf, off)
        pos = struct.unpack_from("<3f", buf, off + 4)
        vel = struct.unpack_from("<3f", buf, off + 16)
        ind, sp, ev = struct.unpack_from("<3I", buf, off + 28)
        out.append((t, pos, vel, ind, sp, ev))
    return out


def _state(n, seed):
    rng = np.random.default_rng(seed)
    pos4 = rng.normal(size=(n, 4)).astype(np.float32)
    pos4[:, 3] = rng.choice([-1.0, 1.0], n)
    vel4 = (rng.normal(size=(n, 4)) * 1e3).astype(np.float32)
    vel4[:, 3] = rng.choice([0, 1, 2, 4, 5, 6, 7], n)
    acc4 = (rng.normal(size=(n, 4)) * 1e12).astype(np.float32)
    acc4[:, 3] = rng.choice([1.0, 2.0, 0.0, -1.0], n, p=[0.6, 0.2, 0.15, 0.05])
    e2 = (rng.normal(size=(n, 3)) * 1e9).astype(np.float32)
    mass = np.array([MASS[int(s)] for s in vel4[:, 3]])
    e1 = W.efield1(acc4[:, :3], pos4[:, 3], mass)
    return pos4, vel4, acc4, e1, e2


def test_new_original_snapshots_read_back_by_the_reference_loader():
    pos4, vel4, acc4, e1, e2 = _state(60, 1)
    b1 = W.save_data(W.SAVE_NEW_ORIGINAL, pos4, vel4, acc4, e1, e2, 5, 50, np.float32(2e-12))
    acc4b = acc4.copy()
    acc4b[7, 3] = 0.0
    b2 = W.save_data(W.SAVE_NEW_ORIGINAL, pos4, vel4, acc4b, e1, e2, 0, 60, np.float32(3e-12))
    snaps = load_particle_data(b1 + b2)
    assert len(snaps) == 2
    for s, (st, first, last, t) in zip(snaps, ((acc4, 5, 50, 2e-12), (acc4b, 0, 60, 3e-12))):
        idx = [i for i in range(first, last) if st[i, 3] in (1.0, 2.0)]
        assert s["time"] == np.float32(t) and s["n_est"] == last
        assert s["ind"] == idx
        assert s["species"] == [int(vel4[i, 3]) for i in idx]
        np.testing.assert_array_equal(s["f"][:, 0:3], pos4[idx, :3])
        np.testing.assert_array_equal(s["f"][:, 3:6], vel4[idx, :3])
        np.testing.assert_array_equal(s["f"][:, 6:9], e1[idx])
        np.testing.assert_array_equal(s["f"][:, 9:12], e2[idx])
    assert 7 in snaps[0]["ind"] or acc4[7, 3] not in (1.0, 2.0)
    assert 7 not in snaps[1]["ind"]


def test_original_layout_matches_the_fwrite_sequence():
    # three ions, the middle one out of bounds (not written)
    pos4 = np.array([[1, 2, 3, 1], [4, 5, 6, -1], [7, 8, 9, -1]], np.float32)
    vel4 = np.array([[10, 11, 12, 5], [13, 14, 15, 0], [16, 17, 18, 2]], np.float32)
    acc4 = np.array([[0, 0, 0, 1], [0, 0, 0, 0], [0, 0, 0, 2]], np.float32)
    e1 = np.arange(9, dtype=np.float32).reshape(3, 3) + 100
    e2 = np.arange(9, dtype=np.float32).reshape(3, 3) + 200
    got = W.save_data(W.SAVE_ORIGINAL, pos4, vel4, acc4, e1, e2, 0, 3, np.float32(0))
    want = b""
    for gg in (0, 2):
        for v in (pos4[gg, 0], pos4[gg, 1], pos4[gg, 2], vel4[gg, 0], vel4[gg, 1], vel4[gg, 2],
                  e1[gg, 0], e1[gg, 1], e1[gg, 2], e2[gg, 0], e2[gg, 1], e2[gg, 2]):
            want += struct.pack("<f", v)
        want += struct.pack("<I", int(vel4[gg, 3])) + struct.pack("<i", gg) + struct.pack("<i", 3)
    assert got == want


def test_full_conditions_write_every_ion():
    pos4, vel4, acc4, e1, e2 = _state(40, 2)
    b = W.save_full(W.SAVE_NEW_ORIGINAL, pos4, vel4, e1, e2, 33, np.float32(1e-9))
    s, = load_particle_data(b)
    assert s["ind"] == list(range(40)) and s["n_est"] == 33
    np.testing.assert_array_equal(s["f"][:, 6:9], e1)
    b0 = W.save_full(W.SAVE_ORIGINAL, pos4, vel4, e1, e2, 33, np.float32(0))
    assert len(b0) == 40 * 60
    assert struct.unpack_from("<i", b0, 60 + 52)[0] == 1 and struct.unpack_from("<i", b0, 56)[0] == 33


def _round_f32(x):
    """The float32 nearest to the rational x (ties to even)."""
    f = np.float32(float(x))
    cands = [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    best = min(cands, key=lambda c: (abs(Fraction(float(c)) - x),
                                     int(np.array(c, np.float32).view(np.uint32)) & 1))
    return best


def test_efield1_is_rounded_once_from_the_exact_product():
    rng = np.random.default_rng(3)
    n = 3000
    acc = (rng.normal(size=(n, 3)) * 10.0 ** rng.uniform(-5, 18, (n, 1))).astype(np.float32)
    acc[:5] = 0.0
    q = rng.choice([-1.0, 1.0], n).astype(np.float32)
    mass = rng.choice(list(MASS.values()), n)
    got = W.efield1(acc, q, mass)
    for i in range(0, n, 7):
        for c in range(3):
            p = float(W.FACTOR) * float(acc[i, c])  # the double product
            x = Fraction(p) * mass[i] / Fraction(float(q[i]))
            want = _round_f32(x)
            if x == 0:
                # signed zero: the sign of a times the sign of q
                assert got[i, c] == 0.0 and np.signbit(got[i, c]) == (np.signbit(acc[i, c]) != (q[i] < 0))
            else:
                assert got[i, c] == want, (i, c, acc[i, c], got[i, c], want)


def test_event_records_read_back_by_the_reference_loader():
    pos4, vel4, acc4, _, _ = _state(20, 4)
    acc4[[3, 9], 3] = 0.0
    acc4[11, 3] = -1.0
    b = W.event_records(np.float32(5e-12), pos4, vel4, acc4, [3, 9, 11])
    ev = load_events_data(b)
    assert [e[3] for e in ev] == [3, 9, 11]
    assert [e[5] for e in ev] == [0, 0, 0xFFFFFFFF]  # outOfBounds, crashed as uint32
    for e in ev:
        assert e[0] == np.float32(5e-12)
        np.testing.assert_array_equal(e[1], pos4[e[3], :3])
        np.testing.assert_array_equal(e[2], vel4[e[3], :3])
        assert e[4] == int(vel4[e[3], 3])
