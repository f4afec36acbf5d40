"""The allocation-free direct entry point (shp_direct_dev_noalloc) against
shp_direct_dev, through the C ABI.

Both run the same kernels with the same split and reduction order, so the
bar is bit-for-bit equality.  Two calls in a row, the second on a smaller
input, check that the shared per-device scratch holds nothing from the
previous call that the next one reads; sizes above the stated maximum are
rejected before any device work."""
import ctypes

import numpy as np
import pytest

import solveit_hip
from solveit_hip import synth

pytestmark = pytest.mark.gpu

MAX_N = 262144
ERR_INVALID = -1  # SHP_ERR_INVALID


@pytest.fixture(scope="module")
def eng():
    import torch
    assert torch.cuda.is_available()
    e = solveit_hip.Engine(0)
    yield e
    e.close()


def _forces(eng, fn, st, precision, sym):
    import torch
    n = st.n
    out = tuple(torch.full((n,), np.nan, dtype=torch.float64, device="cuda:0") for _ in range(3))
    p = st.as_struct()
    torch.cuda.synchronize()
    rc = fn(eng.h, ctypes.byref(p), precision, ctypes.byref(sym), out[0].data_ptr(),
            out[1].data_ptr(), out[2].data_ptr())
    assert rc == 0, solveit_hip.lib.shp_last_error().decode()
    assert solveit_hip.lib.shp_sync(eng.h) == 0
    return torch.stack(out, 1).cpu().numpy()


@pytest.mark.parametrize("precision,sym", [
    (solveit_hip.SHP_PREC_FP32C, "none"),
    (solveit_hip.SHP_PREC_FP64, "none"),
    (solveit_hip.SHP_PREC_FP64, "rotational"),
])
def test_noalloc_equals_direct_dev_twice(eng, precision, sym):
    import torch
    lib = solveit_hip.lib
    s = solveit_hip.make_symmetry(sym, nsec=6)
    for n, seed in ((40000 if sym == "none" else 6000, 71), (3000, 72)):
        d = synth.cone(n, seed=seed)
        d["status"][::11] = 2
        d["status"][::29] = 0
        st = solveit_hip.DeviceState.from_numpy(d)
        torch.cuda.synchronize()
        print(f"case noalloc prec={precision} sym={sym} n={n}", flush=True)
        got = _forces(eng, lib.shp_direct_dev_noalloc, st, precision, s)
        ref = _forces(eng, lib.shp_direct_dev, st, precision, s)
        np.testing.assert_array_equal(got, ref)
        assert np.all(got[d["status"] != 1] == 0)


def test_noalloc_rejects_sizes_beyond_the_maximum(eng):
    import torch
    lib = solveit_hip.lib
    st = solveit_hip.DeviceState.from_numpy(synth.cone(1000, seed=73))
    out = [torch.zeros(1000, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    torch.cuda.synchronize()
    ptrs = [o.data_ptr() for o in out]
    # n above the maximum: rejected on the host before any device work
    p = st.as_struct()
    p.n = MAX_N + 1
    s = solveit_hip.make_symmetry("none")
    rc = lib.shp_direct_dev_noalloc(eng.h, ctypes.byref(p), solveit_hip.SHP_PREC_FP64, ctypes.byref(s),
                                    *ptrs)
    assert rc == ERR_INVALID
    assert "at most" in lib.shp_last_error().decode()
    # more images than the scratch holds: rejected on the host before any kernel
    q = st.as_struct()
    s = solveit_hip.make_symmetry("rotational", nsec=2048)
    rc = lib.shp_direct_dev_noalloc(eng.h, ctypes.byref(q), solveit_hip.SHP_PREC_FP64, ctypes.byref(s),
                                    *ptrs)
    assert rc == ERR_INVALID
    assert "images" in lib.shp_last_error().decode()
    assert np.all(torch.stack(out).cpu().numpy() == 0)
