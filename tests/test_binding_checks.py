"""Argument checks of the Python binding: every device buffer is validated
(device, contiguity, element width, length) before anything reaches the C
ABI, so a wrong tensor raises InvalidArgument instead of letting a kernel
read or write past its allocation.  The CPU tests drive check_buffer with a
stand-in for a GPU tensor; the GPU tests pass bad tensors to real handles
(nothing is launched: the checks fire first)."""

import numpy as np
import pytest


class FakeDeviceTensor:
    """The attributes check_buffer reads from a torch tensor on a GPU."""

    def __init__(self, n, width, contiguous=True, is_cuda=True):
        self.n, self.width = n, width
        self.contiguous, self.is_cuda = contiguous, is_cuda

    def numel(self):
        return self.n

    def element_size(self):
        return self.width

    def is_contiguous(self):
        return self.contiguous

    def data_ptr(self):
        return 0x1000


def test_valid_buffer_passes(gd):
    assert gd.check_buffer(FakeDeviceTensor(8, 4), "k", 4, 8) == 0x1000
    assert gd.check_buffer(FakeDeviceTensor(9, 8), "k", 8, 8) == 0x1000


@pytest.mark.parametrize("t,width,need,msg", [
    (FakeDeviceTensor(8, 4, is_cuda=False), 4, 8, "GPU tensor"),
    (FakeDeviceTensor(8, 4, contiguous=False), 4, 8, "contiguous"),
    (FakeDeviceTensor(8, 4), 8, 8, "8-byte"),      # int32 for 64-bit keys
    (FakeDeviceTensor(8, 8), 4, 8, "4-byte"),
    (FakeDeviceTensor(7, 4), 4, 8, "required"),    # too short
])
def test_bad_buffer_is_rejected(gd, t, width, need, msg):
    with pytest.raises(gd.InvalidArgument, match=msg):
        gd.check_buffer(t, "buf", width, need)


def test_none_only_when_optional_or_unused(gd):
    assert gd.check_buffer(None, "found", 1, 10, optional=True) is None
    assert gd.check_buffer(None, "out", 4, 0) is None
    with pytest.raises(gd.InvalidArgument, match="required"):
        gd.check_buffer(None, "keys", 4, 10)


def test_count_checks(gd):
    t = FakeDeviceTensor(10, 4)
    assert gd._count(t, None) == 10
    assert gd._count(t, 10) == 10
    assert gd._count(0x1000, 5) == 5  # raw address: explicit count
    with pytest.raises(gd.InvalidArgument, match="exceeds"):
        gd._count(t, 11)
    with pytest.raises(gd.InvalidArgument, match="negative"):
        gd._count(t, -1)
    with pytest.raises(gd.InvalidArgument, match="explicit count"):
        gd._count(0x1000, None)


@pytest.mark.gpu
def test_handles_reject_bad_tensors_before_launch(gd, io):
    torch = io.torch
    d = gd.GPULSMOpt(max_elements=1 << 12, batch_capacity=1 << 10)
    k32 = io.up(np.arange(100, dtype=np.uint32))
    v32 = io.up(np.arange(100, dtype=np.uint32))
    with pytest.raises(gd.InvalidArgument):
        d.insert(k32, v32, count=101)                     # count > numel
    with pytest.raises(gd.InvalidArgument):
        d.insert(k32.to(torch.int64), v32)                # 8-byte keys
    with pytest.raises(gd.InvalidArgument):
        d.insert(k32.cpu(), v32)                          # host tensor
    with pytest.raises(gd.InvalidArgument):
        d.insert(k32[::2], v32[::2])                      # strided view
    out = io.empty(50)
    with pytest.raises(gd.InvalidArgument):
        d.lookup(k32, out)                                # output too short
    with pytest.raises(gd.InvalidArgument):
        d.lookup(k32, io.empty(100), io.empty(100, 4))    # found not 1 byte
    f = gd.FlixIndex(key_bytes=8, value_bytes=8)
    with pytest.raises(gd.InvalidArgument):
        f.build_bucket_layer_only(k32)                    # int32 for key64
    srv = gd.Server(max_elements=1 << 12, batch_capacity=1 << 10)
    with pytest.raises(gd.InvalidArgument):
        srv.submit(gd.Server.LOOKUP, k32, None, io.empty(10), count=100)
    srv.close()
    # the handles still work after the rejected calls
    d.insert(k32, v32)
    res = io.empty(100)
    d.lookup(k32, res)
    assert np.array_equal(io.down(res, 100), np.arange(100))
