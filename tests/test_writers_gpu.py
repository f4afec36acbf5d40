"""The resident-state writers through the C ABI against the writer oracle
(oracle/writers.py): snapshots in both SaveDataOptions layouts, the full
conditions, the boundary-condition events and the Efield1 / Efield2 arrays,
byte for byte from the state the engine downloads.  Efield2 is also checked
against the oracle's field at the kick positions (1 float ulp: FP64 fields
that agree to 1e-12, rounded to float)."""
import numpy as np
import pytest

import oracle
from oracle import writers as W
from solveit_hip import synth

pytestmark = pytest.mark.gpu

MASS = {0: 111, 1: 309, 2: 507, 4: 198, 5: 87, 6: 285, 7: 483}


@pytest.fixture(scope="module")
def eng():
    import solveit_hip
    import torch
    assert torch.cuda.is_available()
    e = solveit_hip.Engine(0)
    e.efield_setup(synth.tip_field_mesh())
    yield e
    e.close()


def _plume(n, seed):
    d = synth.cone(n, seed=seed)
    for k in ("x", "y", "z", "vx", "vy", "vz", "q"):
        d[k] = d[k].astype(np.float32).astype(np.float64)
    d["vz"] = d["vz"] * 300.0  # some ions cross the z limit within a few steps
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    acc4[::17, 3] = 2.0      # sources only
    acc4[5::41, 3] = 0.0     # already out of bounds
    return d, pos4, vel4, acc4


def _download(eng, n):
    pos4 = np.zeros((n, 4), np.float32)
    vel4 = np.zeros((n, 4), np.float32)
    acc4 = np.zeros((n, 4), np.float32)
    eng.state_download_host(pos4, vel4, acc4)
    e1, e2 = eng.state_fields_host(n)
    return pos4, vel4, acc4, e1, e2


def _e1(pos4, vel4, acc4):
    mass = np.array([MASS[int(s)] for s in vel4[:, 3]])
    return W.efield1(acc4[:, :3], pos4[:, 3], mass)


def test_snapshots_events_and_fields_match_oracle(eng, tmp_path):
    import solveit_hip as S
    n = 3000
    dt = 1e-12
    d, pos4, vel4, acc4 = _plume(n, 21)
    opts = S.make_step_opts(method=S.SHP_METHOD_DIRECT, precision=S.SHP_PREC_FP64, use_efield=True,
                            bc_space=(0, 0, 8.0e-6, 0))
    print("case writers: 3000-ion cone, FP64 direct, z limit 8 um", flush=True)
    eng.state_upload_host(pos4, vel4, acc4, d["t_inj"])
    f_new = tmp_path / "state_new.bin"
    f_org = tmp_path / "state_org.bin"
    f_full = tmp_path / "state_full.bin"
    f_ev = tmp_path / "events.bin"
    want_new, want_org, want_full = b"", b"", b""
    nev = 0
    with S.CFile(f_new) as fn, S.CFile(f_org) as fo, S.CFile(f_full) as ff, S.CFile(f_ev) as fe:
        for k in range(4):
            eng.state_run(opts, k * dt, dt, 1)
            t = np.float32((k + 1) * dt)
            first, last = (0, n) if k % 2 == 0 else (100, n - 200)
            # three saves between downloads: the two staging slots are reused
            eng.state_save(fn, S.SHP_SAVE_NEW_ORIGINAL, first, last, t)
            eng.state_save(fo, S.SHP_SAVE_ORIGINAL, first, last, t)
            eng.state_save_full(ff, S.SHP_SAVE_NEW_ORIGINAL, last, t)
            nev += eng.state_events(fe)
            p4, v4, a4, e1, e2 = _download(eng, n)
            np.testing.assert_array_equal(e1, _e1(p4, v4, a4))
            want_new += W.save_data(W.SAVE_NEW_ORIGINAL, p4, v4, a4, e1, e2, first, last, t)
            want_org += W.save_data(W.SAVE_ORIGINAL, p4, v4, a4, e1, e2, first, last, t)
            want_full += W.save_full(W.SAVE_NEW_ORIGINAL, p4, v4, e1, e2, last, t)
        eng.state_flush()
    assert f_new.read_bytes() == want_new
    assert f_org.read_bytes() == want_org
    assert f_full.read_bytes() == want_full
    # events: every ion that left (status 1 -> 0) once, with its final state
    # (an ion out of bounds no longer moves), at a drift's end time, in the
    # reference's order (by drift, then index)
    ev = np.frombuffer(f_ev.read_bytes(), W.EVENT)
    left = np.flatnonzero((acc4[:, 3] == 1.0) & (a4[:, 3] == 0.0))
    print(f"  {len(ev)} events, {len(left)} ions left", flush=True)
    assert len(ev) == nev == len(left) > 0
    assert sorted(ev["index"].tolist()) == left.tolist()
    np.testing.assert_array_equal(ev["pos"], p4[ev["index"], :3])
    np.testing.assert_array_equal(ev["vel"], v4[ev["index"], :3])
    assert np.all(ev["event"] == 0) and np.all(ev["species"] == v4[ev["index"], 3].astype(np.uint32))
    ends = {np.float32(j * dt / 2) for j in range(1, 9)}
    assert set(ev["time"].tolist()) <= {float(x) for x in ends}
    order = np.lexsort((ev["index"], ev["time"]))
    assert np.array_equal(order, np.arange(len(ev)))


def test_efield2_is_the_field_at_the_kick(eng):
    import solveit_hip as S
    n = 2000
    dt = 1e-12
    d, pos4, vel4, acc4 = _plume(n, 22)
    acc4[:, 3] = 1.0
    acc4[::9, 3] = 2.0
    opts = S.make_step_opts(method=S.SHP_METHOD_DIRECT, precision=S.SHP_PREC_FP64, use_efield=True)
    print("case Efield2 at the kick", flush=True)
    eng.state_upload_host(pos4, vel4, acc4, d["t_inj"])
    _, _, _, _, e2_0 = _download(eng, n)
    assert not np.any(e2_0)  # no kick yet
    eng.state_run(opts, 0.0, dt, 1)
    _, _, a4, _, e2 = _download(eng, n)
    # the kick sees the first half drift: x + v dt/2 (injected at 0, no folding)
    xk = d["x"] + d["vx"] * (dt / 2)
    yk = d["y"] + d["vy"] * (dt / 2)
    zk = d["z"] + d["vz"] * (dt / 2)
    E, _, _ = oracle.efield_f64(synth.tip_field_mesh(), xk, yk, zk)
    kicked = acc4[:, 3] == 1.0
    np.testing.assert_array_max_ulp(e2[kicked], E[kicked].astype(np.float32), maxulp=1)
    assert not np.any(e2[~kicked])  # sources only are not kicked (particles.cpp:504)
    assert np.count_nonzero(e2[kicked]) > 0.9 * 3 * kicked.sum()
