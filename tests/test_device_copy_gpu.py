"""The engine's device-to-device copy (block-chunked 16-byte copy kernel, byte tail, hipMemcpyAsync for
unaligned pointers), through the C ABI: bytes copied exactly, nothing written past the end."""
import numpy as np
import pytest

from vectisdb_amd import capi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("src_off,dst_off", [(0, 0), (16, 32), (3, 0), (0, 5)])
@pytest.mark.parametrize("nbytes", [0, 1, 15, 16, 17, 16383, 16384, 16385, 3_000_007])
def test_device_copy_exact_and_bounded(dev, nbytes, src_off, dst_off):
    import torch
    rng = np.random.default_rng(nbytes + 7 * src_off + 13 * dst_off)
    src_host = rng.integers(0, 256, nbytes + src_off + 64, dtype=np.uint8)
    src = torch.from_numpy(src_host).to(dev)
    dst = torch.full((nbytes + dst_off + 64,), 0xA5, dtype=torch.uint8, device=dev)
    capi.device_copy(dst[dst_off:], src[src_off:], nbytes)
    torch.cuda.synchronize()
    got = dst.cpu().numpy()
    np.testing.assert_array_equal(got[dst_off:dst_off + nbytes], src_host[src_off:src_off + nbytes])
    assert (got[:dst_off] == 0xA5).all()
    assert (got[dst_off + nbytes:] == 0xA5).all()


def test_device_copy_rejects_overrun(dev):
    import torch
    a = torch.zeros(16, dtype=torch.uint8, device=dev)
    b = torch.zeros(8, dtype=torch.uint8, device=dev)
    with pytest.raises(ValueError):
        capi.device_copy(b, a)
