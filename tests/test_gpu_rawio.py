"""Raw data file -> device tensor through the pinned staging buffer."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_raw_file_to_device(tmp_path):
    from yoricktrtt_amd import rawio

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    print("[gpu] rawio to_device", flush=True)
    vol = np.random.default_rng(3).uniform(size=(6, 5, 4)).astype(np.float32)
    rawio.raw_write(tmp_path / "v.raw", vol, offset=512)
    host = rawio.read_volume(tmp_path / "v.raw", 4, 5, 6)
    dev = rawio.to_device(host, "cuda:0", torch.float32)
    assert dev.is_cuda and dev.dtype == torch.float32
    assert np.array_equal(dev.cpu().numpy(), vol)
