"""The HBM streaming probe (femx_hbm_probe) through the C ABI: every kind
runs and reports a plausible rate (a 64 MiB buffer may be served from the
256 MB last-level cache, so the upper bound is loose), and bad arguments
are rejected before any launch."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["copy", "read", "write", "mix_7r5w"])
def test_probe_kinds(gpu, kind):
    tbs = gpu.hbm_probe(kind, 64 << 20, 512, 3)
    assert 0.1 < tbs < 50.0, tbs


def test_probe_rejects_bad_arguments(gpu):
    with pytest.raises(gpu.FemxError):
        gpu.hbm_probe("copy", 1 << 10, 512, 3)  # below 1 MiB
    with pytest.raises(gpu.FemxError):
        gpu.hbm_probe("copy", 64 << 20, 0, 3)
