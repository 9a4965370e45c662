"""CPU tests of the engine's host-side guards: attachments that expire when
their tensor changes, the dense-kernel cache keyed on tensor identity, and
argument checks that run before any pointer reaches the library."""
import gc

import numpy as np
import pytest
import torch

from opticalimaging_amd import engine


def test_attachment_expires_on_in_place_edit():
    t = torch.arange(6, dtype=torch.float32)
    engine.attach(t, "_oi_x", "derived")
    assert engine.attached(t, "_oi_x") == "derived"
    t.mul_(2)
    assert engine.attached(t, "_oi_x") is None
    engine.attach(t, "_oi_x", "again")
    t[0] = 5
    assert engine.attached(t, "_oi_x") is None
    assert engine.attached(t.clone(), "_oi_x") is None          # copies carry nothing
    assert engine.attached(np.zeros(3), "_oi_x") is None
    u = torch.zeros(4)
    engine.attach(u, "_oi_x", 1)
    u.set_(torch.ones(4).untyped_storage())                     # re-pointed storage
    assert engine.attached(u, "_oi_x") is None


def test_dense_cache_is_keyed_on_the_tensor_object():
    cache = engine._DenseCache(capacity=4)
    t = torch.zeros(8, 8, 2, dtype=torch.complex64)
    cache.put(t, "ks-t")
    assert cache.get(t) == "ks-t"
    t.add_(1)
    assert cache.get(t) is None                                   # version changed
    cache.put(t, "ks-t2")
    old_id = id(t)
    del t
    gc.collect()
    reused = False
    keep = []
    for _ in range(200):                   # find a new tensor at the freed object's id
        u = torch.zeros(8, 8, 2, dtype=torch.complex64)
        if id(u) == old_id:
            reused = True
            assert cache.get(u) is None
            break
        keep.append(u)
    for u in keep:
        assert cache.get(u) is None
    print("object id reused:", reused)


def test_dense_cache_capacity():
    cache = engine._DenseCache(capacity=2)
    ts = [torch.zeros(2) + i for i in range(3)]
    for i, t in enumerate(ts):
        cache.put(t, i)
    assert cache.get(ts[0]) is None and cache.get(ts[1]) == 1 and cache.get(ts[2]) == 2


def _kernel_set(ny=16, nx=16):
    return engine.KernelSet(torch.zeros((2, 3, 3), dtype=torch.complex64), -1, -1, ny, nx)


@pytest.mark.parametrize("masks", [
    torch.zeros((2, 16, 16)),                                   # host tensor
    torch.zeros((2, 16, 16), dtype=torch.float64),
    torch.zeros((2, 16, 32))[:, :, ::2],                        # strided view
    np.zeros((2, 16, 16), dtype=np.float32),                    # not a tensor
])
def test_socs_image_rejects_bad_masks_before_launch(masks):
    with pytest.raises(ValueError):
        engine.socs_image(masks, _kernel_set())


def test_socs_image_rejects_grid_mismatch():
    with pytest.raises(ValueError, match="does not match kernel grid"):
        engine.socs_image(torch.zeros((1, 8, 8)), _kernel_set())


def test_residual_and_adjoint_reject_bad_arguments():
    ks = _kernel_set()
    m = torch.zeros((2, 16, 16))
    with pytest.raises(ValueError):
        engine.socs_residual(m, ks, m, m)
    with pytest.raises(ValueError):
        engine.socs_adjoint(m, m.double(), ks)
    with pytest.raises(ValueError):
        engine.socs_adjoint(m, torch.zeros((1, 16, 16)), ks)
    with pytest.raises(ValueError):
        engine.socs_adjoint(m, m, ks)                            # host tensors


def test_device_tensor_checks():
    t = torch.zeros((2, 3))
    with pytest.raises(ValueError, match="HIP device"):
        engine._device_tensor("x", t, torch.float32, (2, 3))
    with pytest.raises(ValueError, match="torch tensor"):
        engine._device_tensor("x", [1, 2], torch.float32, (2,))
