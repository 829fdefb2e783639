"""cmg_residual_device_noalloc (the residual entry point whose scratch is a fixed
device array instead of the context's buffers) against cmg_residual_device, through
the C ABI: bitwise the same assembled residual on two calls in a row, the second on a
smaller mesh (the shared scratch keeps the first call's contents), the corrupt-element
report, and the size limit."""
import ctypes

import numpy as np
import pytest

from cardiomechanics_amd import materials as MT
from cardiomechanics_amd import mesh as M
from cardiomechanics_amd.context import MechanicsContext
from cardiomechanics_amd.hip import DeviceArray

pytestmark = pytest.mark.gpu

ERR_INVALID, CORRUPT_ELEMENT, SUCCESS = -100, -6, 2


def deformed(mesh, amp=0.02, seed=1234):
    L = np.ptp(mesh.coords, axis=0).max()
    return mesh.coords + M.smooth_displacement(mesh.coords, L, amp, seed)


def both(ctx, x, t, pressure=1):
    xd = DeviceArray(np.ascontiguousarray(x.ravel()))
    f_ref = DeviceArray(np.zeros(ctx.n_dofs))
    f_new = DeviceArray(np.full(ctx.n_dofs, np.nan))
    L = ctx.L
    rc_ref = L.cmg_residual_device(ctx.h, xd.ptr, ctypes.c_double(t), pressure, f_ref.ptr)
    assert L.cmg_synchronize(ctx.h) == SUCCESS
    rc_new = L.cmg_residual_device_noalloc(ctx.h, xd.ptr, ctypes.c_double(t), pressure, f_new.ptr)
    assert L.cmg_synchronize(ctx.h) == SUCCESS
    out = (rc_ref, rc_new, f_ref.download(), f_new.download())
    for a in (xd, f_ref, f_new):
        a.free()
    return out


def test_noalloc_residual_equals_the_allocating_entry_point():
    cases = [
        (M.lv_mesh(20_000), [MT.lv_holzapfel()], -8000.0, 0.12),   # larger first
        (M.cube_mesh(2), [MT.cube_guccione()], 2000.0, 0.0),       # then smaller
    ]
    for mesh, mats, p, t in cases:
        print(f"[noalloc] {mesh.name}: {mesh.n_elements} elements", flush=True)
        with MechanicsContext(mesh, mats) as ctx:
            ctx.set_pressures([p])
            rc_ref, rc_new, f_ref, f_new = both(ctx, deformed(mesh), t)
            assert rc_ref == SUCCESS and rc_new == SUCCESS, (rc_ref, rc_new, ctx.L.cmg_last_error())
            assert np.abs(f_ref).max() > 0
            assert f_new.tobytes() == f_ref.tobytes()
            rc_ref, rc_new, f_ref, f_new = both(ctx, deformed(mesh), t, pressure=0)
            assert rc_ref == SUCCESS and rc_new == SUCCESS
            assert f_new.tobytes() == f_ref.tobytes()


def test_noalloc_reports_a_corrupt_element():
    mesh, mats = M.cube_mesh(2), [MT.cube_guccione()]
    x = mesh.coords.copy()
    el = mesh.elements[5]
    x[el[0]] = x[el[1:4]].mean(axis=0) * 2 - x[el[0]]    # corner 0 through the opposite face
    with MechanicsContext(mesh, mats) as ctx:
        rc_ref, rc_new, _, _ = both(ctx, x, 0.0)
        assert rc_ref == CORRUPT_ELEMENT and rc_new == CORRUPT_ELEMENT
        assert "corrupt" in ctx.L.cmg_last_error().decode()
        # the flag is cleared by the next call
        rc_ref, rc_new, f_ref, f_new = both(ctx, mesh.coords, 0.0)
        assert rc_new == SUCCESS and f_new.tobytes() == f_ref.tobytes()


def test_noalloc_rejects_meshes_beyond_its_scratch():
    mesh = M.lv_mesh(300_000)            # ~9e6 doubles of element forces > 8 Mi
    assert mesh.n_elements * 30 > 8 * 1024 * 1024
    with MechanicsContext(mesh, [MT.lv_holzapfel()]) as ctx:
        xd = DeviceArray(mesh.coords.ravel())
        fd = DeviceArray(np.zeros(ctx.n_dofs))
        rc = ctx.L.cmg_residual_device_noalloc(ctx.h, xd.ptr, ctypes.c_double(0.0), 1, fd.ptr)
        assert rc == ERR_INVALID
        assert "scratch" in ctx.L.cmg_last_error().decode()
