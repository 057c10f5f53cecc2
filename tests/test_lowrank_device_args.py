"""pgs_gemslr_setup_device without a usable device: a negative index is
refused before any work, and on a machine without a device the setup
returns PGS_ERR_NO_DEVICE (no host fallback for the low-rank build)."""
import pytest

import pargemslr_amd as pg


def _problem():
    A = pg.laplacian(8, 8, 8)
    return A, pg.config_params(3, 4, 1e-3, 20, 8)


def test_negative_device_is_refused():
    A, p = _problem()
    with pytest.raises(pg.PgsError) as e:
        pg.Gemslr(A, p, lowrank_device=-1)
    assert e.value.code == 101


@pytest.mark.skipif(pg.device_count() > 0, reason="a device is present")
def test_no_device_is_an_error_not_a_fallback():
    A, p = _problem()
    with pytest.raises(pg.PgsError) as e:
        pg.Gemslr(A, p, lowrank_device=0)
    assert e.value.code == 202
