"""pbfmm_pressure_force_and_moment_noalloc, the allocation-free variant of
pbfmm_pressure_force_and_moment whose partial sums live in the library's
fixed device array, against the existing entry point through the C ABI:
two handles in a row, the second with a smaller mesh (the array keeps the
first call's partial sums), with and without a cell mask.  The two use the
same kernels and summation order, so the bar is 1e-14 of the largest value.
No device memory is allocated by the calls (free device memory unchanged).
"""
import ctypes as C

import numpy as np
import pytest

import pibem_amd as P
from pibem_amd import hip
from pibem_amd import mesh as M

pytestmark = pytest.mark.gpu


def _free_bytes():
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hip().hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


class DeviceBytes:
    def __init__(self, a):
        a = np.ascontiguousarray(a, dtype=np.uint8)
        self.ptr = C.c_void_p()
        assert hip.hip().hipMalloc(C.byref(self.ptr), max(a.size, 1)) == 0
        assert hip.hip().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), a.size, 1) == 0  # host to device

    @property
    def value(self):
        return self.ptr.value

    def __del__(self):
        if self.ptr:
            hip.hip().hipFree(self.ptr)


def _handle(n):
    pr = M.config_sphere(n, 4)
    f = P.BEMFMA(trunc_order=4)
    f.init_fma(pr.mesh, pr.dirichlet)
    bp = P.BEMProblem(f)
    pres = np.random.default_rng(n).uniform(-1, 1, pr.mesh.n_dofs) + pr.mesh.support_xyz[:, 2]
    return pr, f, bp, pres


def test_noalloc_matches_entry_point():
    cases = [_handle(16), _handle(6)]  # 1,536 then 216 cells
    body_all = P.Body(range(6), center_of_gravity=(0.1, -0.2, 0.3))
    body_part = P.Body([0, 2, 5], center_of_gravity=(-0.3, 0.2, 0.1))
    ref, dev = {}, {}
    for k, (pr, f, bp, pres) in enumerate(cases):
        for name, body in (("all", body_all), ("part", body_part)):
            ref[k, name] = np.concatenate(bp.pressure_force_and_moment(body, pres))  # sets up post-solve data
        dev[k] = (hip.DeviceArray.from_numpy(pres), DeviceBytes(body_part.cell_mask(pr.mesh)), hip.DeviceArray(6))
    hip.synchronize()
    free0 = _free_bytes()
    out = {}
    for k, (pr, f, bp, pres) in enumerate(cases):  # the larger mesh first, then the smaller one
        dp, dm, do = dev[k]
        for name, body, mptr in (("all", body_all, None), ("part", body_part, dm.value)):
            bp.pressure_force_and_moment_noalloc(mptr, dp.value, body.center_of_gravity, do.value)
            hip.synchronize()
            out[k, name] = do.numpy()
    assert _free_bytes() == free0
    for key in ref:
        a, b = out[key], ref[key]
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        print(f"\n  {key}: noalloc {a} entry point {b} max rel diff {err:.1e}", end="")
        assert err <= 1e-14
    # the larger mesh again after the smaller one
    pr, f, bp, pres = cases[0]
    dp, dm, do = dev[0]
    bp.pressure_force_and_moment_noalloc(None, dp.value, body_all.center_of_gravity, do.value)
    hip.synchronize()
    assert np.max(np.abs(do.numpy() - ref[0, "all"])) <= 1e-14 * np.max(np.abs(ref[0, "all"]))
    for pr, f, bp, pres in cases:
        f.close()


def test_noalloc_errors():
    pr, f, bp, pres = _handle(6)
    dp, do = hip.DeviceArray.from_numpy(pres), hip.DeviceArray(6)
    with pytest.raises(P.PbfmmError) as ei:  # post-solve data not set up yet
        bp.pressure_force_and_moment_noalloc(None, dp.value, (0, 0, 0), do.value)
    assert ei.value.code == -2
    bp.volume_integral(P.Body(range(6)))  # sets it up
    with pytest.raises(P.PbfmmError) as ei:
        bp.pressure_force_and_moment_noalloc(None, None, (0, 0, 0), do.value)
    assert ei.value.code == -1
    bp.pressure_force_and_moment_noalloc(None, dp.value, (0, 0, 0), do.value)
    hip.synchronize()
    f.close()
