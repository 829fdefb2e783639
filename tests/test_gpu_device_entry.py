"""The device-pointer entry points through the C ABI -- the interface of a solver that
keeps its vectors on the GPU: cmg_residual_tangent_device and cmg_residual_device give
the host entry points' assembled residual and tangent values bit for bit (the same
kernels, stream ordered), in FP64 and mixed precision, with and without the surface
pressure; cmg_element_kernel_device leaves the element outputs that
cmg_element_residuals reports."""
import ctypes

import numpy as np
import pytest

from cardiomechanics_amd import materials as MT
from cardiomechanics_amd import mesh as M
from cardiomechanics_amd.context import MechanicsContext
from cardiomechanics_amd.hip import DeviceArray

pytestmark = pytest.mark.gpu

SUCCESS = 2


def deformed(mesh, amp=0.02, seed=1234):
    L = np.ptp(mesh.coords, axis=0).max()
    return mesh.coords + M.smooth_displacement(mesh.coords, L, amp, seed)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_device_entry_points_equal_the_host_entry_points(precision):
    mesh, mats = M.lv_mesh(dims=(2, 4, 8)), [MT.lv_holzapfel()]
    x = np.ascontiguousarray(deformed(mesh).ravel())
    t = 0.12
    with MechanicsContext(mesh, mats, precision=precision) as ctx:
        L = ctx.L
        ctx.set_pressures([-4000.0])
        vt = ctx.value_dtype
        xd = DeviceArray(x)
        fd = DeviceArray(np.full(ctx.n_dofs, np.nan))
        vd = DeviceArray(np.full(ctx.nnzb * 9, np.nan, dtype=vt))
        try:
            assert not np.array_equal(ctx.residual(x, t, include_pressure=True),
                                      ctx.residual(x, t, include_pressure=False))
            for pressure in (1, 0):
                r_host = ctx.residual(x, t, include_pressure=bool(pressure))
                v_host = ctx.tangent(x, t, include_pressure=bool(pressure))
                rc = L.cmg_residual_tangent_device(ctx.h, xd.ptr, ctypes.c_double(t), pressure, fd.ptr, vd.ptr)
                assert rc == SUCCESS and L.cmg_synchronize(ctx.h) == SUCCESS, L.cmg_last_error()
                assert fd.download().tobytes() == r_host.tobytes()
                assert vd.download().tobytes() == v_host.tobytes()
                fd.upload(np.full(ctx.n_dofs, np.nan))
                rc = L.cmg_residual_device(ctx.h, xd.ptr, ctypes.c_double(t), pressure, fd.ptr)
                assert rc == SUCCESS and L.cmg_synchronize(ctx.h) == SUCCESS, L.cmg_last_error()
                assert fd.download().tobytes() == r_host.tobytes()
                print(f"precision {precision} pressure {pressure}: |r| {np.abs(r_host).max():.3e}, "
                      f"{ctx.nnzb} blocks, device = host bit for bit", flush=True)
            fe_host = ctx.element_residuals(x.reshape(-1, 3), t)
            for with_tangent in (0, 1):
                rc = L.cmg_element_kernel_device(ctx.h, xd.ptr, ctypes.c_double(t), with_tangent)
                assert rc == SUCCESS and L.cmg_synchronize(ctx.h) == SUCCESS, L.cmg_last_error()
            # the element kernel writes the context's element buffers, which the element
            # residual entry point copies out after its own evaluation at the same state
            assert ctx.element_residuals(x.reshape(-1, 3), t).tobytes() == fe_host.tobytes()
        finally:
            for a in (xd, fd, vd):
                a.free()
