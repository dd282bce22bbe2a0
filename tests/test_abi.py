"""The C-ABI library: loads without a GPU, exports every function the header
declares, and rejects bad arguments before touching a device."""
import ctypes

import numpy as np
import pytest

from water_flow_simulation_amd import _native as nat


def test_library_exports_every_header_function():
    L = nat.lib()
    names = nat.header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(nat._SIGS), "ctypes signatures out of sync with include/wfs.h"


def test_abi_version_and_device_count():
    L = nat.lib()
    assert L.wfs_abi_version() == 1
    assert nat.device_count() >= 0


@pytest.mark.parametrize("desc,needle", [
    (dict(nx=2, ny=4, nz=4), "nx >= 3"),
    (dict(nx=2048, ny=2048, nz=1024), "2^31"),
    (dict(nx=16, ny=4, nz=4, tau0=0.5), "tau0"),
    (dict(nx=16, ny=4, nz=4, block_x=96), "block_x"),
])
def test_create_rejects_bad_descriptors(desc, needle):
    d = dict(nx=16, ny=4, nz=4, tau0=0.5006, cs=0.17, storage=0, block_x=0, device=0)
    d.update(desc)
    h = ctypes.c_void_p()
    rc = nat.lib().wfs_create(ctypes.byref(nat.wfs_desc(**d)), ctypes.byref(h))
    assert rc == nat.WFS_EINVAL
    assert needle in nat.last_error()
    with pytest.raises(ValueError):
        nat.check(rc)


def test_null_handle_is_rejected():
    L = nat.lib()
    assert L.wfs_step(None, 1) == nat.WFS_EINVAL
    assert L.wfs_nstep(None) == -1
    assert L.wfs_destroy(None) == nat.WFS_OK


def test_no_device_is_an_error_not_a_fallback():
    if nat.device_count() > 0:
        pytest.skip("a GPU is visible")
    from water_flow_simulation_amd.lbm3d import LBM3D
    solid = np.zeros((4, 4, 16), bool)
    z = np.zeros(solid.shape, np.float32)
    with pytest.raises(nat.WfsError):
        LBM3D(None, solid, z, z, z, np.zeros((4, 4), np.float32), 0.5006)


def test_torch_backend_rejected_like_reference():
    from water_flow_simulation_amd.lbm3d import LBM3D

    class Be:
        name = "torch"

    solid = np.zeros((4, 4, 16), bool)
    z = np.zeros(solid.shape, np.float32)
    with pytest.raises(RuntimeError):
        LBM3D(Be(), solid, z, z, z, np.zeros((4, 4), np.float32), 0.5006)
