"""The achievable-bandwidth probes (stream.hip) on the GPU: each returns a
rate within what the hardware can do (above 0.1 TB/s, below the 8 TB/s HBM
peak for HBM-sized buffers), and a gather from a buffer the L2s hold runs
faster than the same gather from HBM."""
import pytest

pytestmark = pytest.mark.gpu

HBM_PEAK_GBPS = 8000.0


def test_stream_probes_are_plausible(gpu, gbx):
    copy = gbx.stream_copy_bw(256 << 20, 3)
    r3 = gbx.stream_bw(3 * (128 << 20), 3, 3)
    r0 = gbx.stream_bw(256 << 20, 0, 3)
    print(f"copy {copy:.0f}, 3:1 {r3:.0f}, read-only {r0:.0f} GB/s")
    for v in (copy, r3, r0):
        assert 100.0 < v < HBM_PEAK_GBPS


def test_gather_probe_l2_versus_hbm(gpu, gbx):
    hbm = gbx.gather_bw(1 << 30, 3)
    l2 = gbx.gather_bw(2 << 20, 3)
    print(f"gathers of 256 B runs: HBM-sized buffer {hbm:.0f} GB/s, L2-sized buffer {l2:.0f} GB/s")
    assert 100.0 < hbm < HBM_PEAK_GBPS
    assert l2 > hbm


def test_gather_probe_argument_errors(gpu, gbx):
    for args in ((3 << 20, 1), (1 << 19, 1), (1 << 20, 0)):
        with pytest.raises(gbx.GbxError) as e:
            gbx.gather_bw(*args)
        assert e.value.status == 1
