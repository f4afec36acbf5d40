"""Oracle of the output writers: a byte-level restatement of the reference's
state and event files, written from the reference's float arrays.

Test infrastructure only: the tests use it as the checker of the engine's
writer entry points (shp_state_save, shp_state_save_full, shp_state_events,
shp_state_fields_host); nothing in the product imports it.

* efield1          integration_schemes.cpp:350-355 (Efield1 = factor a m / q)
* save_data        Integrator::saveData, integration_schemes.cpp:263-331
* save_full        Integrator::saveFullCurrentConditions, :197-258
* event_records    ChargedParticles::writeToEventsFile, particles.cpp:666-684
"""
import numpy as np

SAVE_ORIGINAL = 0       # SaveDataOptions::original (structures.h:244-247)
SAVE_NEW_ORIGINAL = 1   # SaveDataOptions::newOriginal

# PhysicalConstant::AMU / PhysicalConstant::ELEM_CHARGE: both float
# constants (constants.h:33-34), so the quotient is computed in float.
FACTOR = np.float64(np.float32(1.66053904e-27) / np.float32(1.6021766e-19))

_REC_ORIGINAL = np.dtype([("f", "<f4", 12), ("species", "<u4"), ("index", "<i4"), ("last", "<i4")])
_REC_NEW = np.dtype([("index", "<i4"), ("species", "<u4"), ("f", "<f4", 12)])
EVENT = np.dtype([("time", "<f4"), ("pos", "<f4", 3), ("vel", "<f4", 3), ("index", "<u4"),
                  ("species", "<u4"), ("event", "<u4")])


def efield1(acc_xyz, q, mass):
    """factor * acceleration * getMass(j) / position.w per component: the
    double product factor * a (a float), then the long double mass
    (getMass returns long double, particles.h:138) and the float charge,
    rounded to float once."""
    a = np.asarray(acc_xyz, np.float32).astype(np.float64)
    p = FACTOR * a
    m = np.asarray(mass, np.longdouble)[:, None]
    qq = np.asarray(q, np.float32).astype(np.longdouble)[:, None]
    return ((p.astype(np.longdouble) * m) / qq).astype(np.float32)


def _species_u32(vel4):
    # uint32_t temp = velocity.w (a float species code)
    return np.asarray(vel4, np.float32)[:, 3].astype(np.uint32)


def _records(fmt, idx, pos4, vel4, e1, e2, last):
    f = np.concatenate([pos4[idx, :3], vel4[idx, :3], e1[idx], e2[idx]], axis=1).astype(np.float32)
    sp = _species_u32(vel4)[idx]
    if fmt == SAVE_ORIGINAL:
        r = np.zeros(len(idx), _REC_ORIGINAL)
        r["f"] = f
        r["species"] = sp
        r["index"] = idx
        r["last"] = last
    else:
        r = np.zeros(len(idx), _REC_NEW)
        r["index"] = idx
        r["species"] = sp
        r["f"] = f
    return r.tobytes()


def _header(fmt, last, time):
    if fmt == SAVE_ORIGINAL:
        return b""
    return (np.array([0xFFFFFFFF], "<u4").tobytes() + np.array([last], "<i4").tobytes()
            + np.array([time], "<f4").tobytes())


def save_data(fmt, pos4, vel4, acc4, e1, e2, first, last, time):
    """The bytes saveData writes: ions of [first, last) whose acceleration.w is
    inbound (1) or affectsAcceleration (2), in index order."""
    st = np.asarray(acc4, np.float32)[:, 3]
    idx = np.arange(first, last)
    idx = idx[(st[idx] == 1.0) | (st[idx] == 2.0)]
    return _header(fmt, last, time) + _records(fmt, idx, np.asarray(pos4, np.float32),
                                               np.asarray(vel4, np.float32), e1, e2, last)


def save_full(fmt, pos4, vel4, e1, e2, last, time):
    """The bytes saveFullCurrentConditions writes: every ion, whatever its status."""
    idx = np.arange(np.asarray(pos4).shape[0])
    return _header(fmt, last, time) + _records(fmt, idx, np.asarray(pos4, np.float32),
                                               np.asarray(vel4, np.float32), e1, e2, last)


def event_records(time, pos4, vel4, acc4, indices):
    """writeToEventsFile for each index: time, position, velocity, index,
    species and the event (acceleration.w as uint32)."""
    idx = np.asarray(indices, np.int64)
    r = np.zeros(len(idx), EVENT)
    r["time"] = np.asarray(time, np.float32)
    r["pos"] = np.asarray(pos4, np.float32)[idx, :3]
    r["vel"] = np.asarray(vel4, np.float32)[idx, :3]
    r["index"] = idx
    r["species"] = _species_u32(vel4)[idx]
    r["event"] = np.asarray(acc4, np.float32)[idx, 3].astype(np.int64).astype(np.uint32)
    return r.tobytes()
