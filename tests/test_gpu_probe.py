"""bs_hbm_probe: the device's achievable streaming bandwidths are physical
(positive, at most the published 8 TB/s peak of MI355X HBM3E) and ordered as
HBM behaves (a pure read stream is not slower than a copy)."""
import pytest

pytestmark = pytest.mark.gpu


def test_hbm_probe_rates_are_physical(bs, gpu_available):
    r = bs.hbm_probe(1 << 30, 3)
    print(r, flush=True)
    for k in ("read_GBps", "write_GBps", "copy_GBps"):
        assert 500.0 < r[k] <= 8000.0 * 1.02, (k, r[k])
    assert r["read_GBps"] >= r["copy_GBps"] * 0.95
