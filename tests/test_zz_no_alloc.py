"""rh_divergence_max_noalloc (the entry point that allocates nothing) against
rh_divergence_max through the C ABI: equal results on two calls in a row, the
second on a smaller grid (the scratch array keeps the first call's contents),
for f32 and f64 faces, plus the validation failure it shares."""
import ctypes as C

import numpy as np
import pytest

import rheon_amd as R
from rheon_amd import _native as N
from rheon_amd.scenes import GridSpec, taylor_green_noise


def run(fn, spec, vel):
    lib = N.lib()
    g = R.make_grid(spec)
    prec = R.F32 if vel[0].dtype == np.float32 else R.F64
    bufs = []
    for v in vel:
        p = C.c_void_p()
        R._check(lib.rh_device_alloc(v.nbytes, C.byref(p)), "rh_device_alloc")
        R._check(lib.rh_copy_to_device(p, v.ctypes.data, v.nbytes), "rh_copy_to_device")
        bufs.append(p)
    out = C.c_double(-1.0)
    st = getattr(lib, fn)(C.byref(g), prec, bufs[0], bufs[1], bufs[2], C.byref(out))
    for p in bufs:
        lib.rh_device_free(p)
    return st, out.value


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_matches_divergence_max(dtype):
    big = GridSpec((96, 80, 72), (1 / 96,) * 3)
    small = GridSpec((24, 17, 9), (0.07, 0.11, 0.19), (-2.0, 4.0, 0.3))
    for spec, seed in ((big, 1), (small, 2)):
        vel = taylor_green_noise(spec, seed, 0.1, dtype)
        want = run("rh_divergence_max", spec, vel)
        got = run("rh_divergence_max_noalloc", spec, vel)
        assert want[0] == R.OK and got[0] == R.OK
        assert got[1] == want[1] and got[1] > 0.0
    # a nonzero wall face: the same rejection as rh_divergence_max
    vel = taylor_green_noise(small, 3, 0.1, dtype)
    vel[0][0] = 1.0
    assert run("rh_divergence_max_noalloc", small, vel)[0] == run("rh_divergence_max", small, vel)[0] \
        == R.NONZERO_WALL_VELOCITY
