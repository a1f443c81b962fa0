"""CPU checks of the launch schedule (pgalf_set_schedule / pgalf_get_schedule
and galaxyfinder_amd.tune's document format): host-only calls, no kernels."""
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from galaxyfinder_amd import density as D, tune  # noqa: E402
from galaxyfinder_amd._lib import PgalfError  # noqa: E402

DEFAULT = {"nmax": -1, "knn_packed": 1, "knn_xcd_chunk": 64, "tsc_unit_entries": 4096}


@pytest.fixture(autouse=True)
def _restore():
    yield
    D.set_schedule()


def _bin(nmax, packed=1, chunk=64, unit=4096):
    return dict(nmax=nmax, knn_packed=packed, knn_xcd_chunk=chunk, tsc_unit_entries=unit)


def test_defaults():
    for n in (0, 1000, 1 << 40):
        assert D.get_schedule(n) == (DEFAULT, 56)


def test_bins_pick_first_bin_at_or_above_n():
    D.set_schedule([_bin(1000, 0, 0, 1024), _bin(100000, 1, 16, 16384)], 32)
    assert D.get_schedule(1000)[0] == _bin(1000, 0, 0, 1024)
    assert D.get_schedule(1001)[0] == _bin(100000, 1, 16, 16384)
    assert D.get_schedule(100000)[0] == _bin(100000, 1, 16, 16384)
    assert D.get_schedule(100001) == (DEFAULT, 32)
    D.set_schedule()
    assert D.get_schedule(10) == (DEFAULT, 56)


@pytest.mark.parametrize("bins,dm", [
    ([_bin(1000), _bin(1000)], 56),      # not ascending
    ([_bin(1000, packed=2)], 56),
    ([_bin(1000, chunk=-1)], 56),
    ([_bin(1000, chunk=2048)], 56),
    ([_bin(1000, unit=128)], 56),
    ([_bin(1000, unit=1 << 17)], 56),
    ([], 65),
    ([], -1),
])
def test_bad_schedules_are_rejected_and_leave_the_table(bins, dm):
    D.set_schedule([_bin(500, 0, 0, 1024)], 40)
    with pytest.raises(PgalfError):
        D.set_schedule(bins, dm)
    assert D.get_schedule(100) == (_bin(500, 0, 0, 1024), 40)


def test_tune_document_round_trip(tmp_path):
    doc = {"device": "x", "bins": [_bin(16384, 1, 16, 4096), _bin(1 << 62, 1, 64, 16384)],
           "smooth_direct_max": 48, "timings": {}}
    p = tmp_path / "s.json"
    p.write_text(json.dumps(doc))
    tune.load(str(p))
    assert D.get_schedule(10000) == (_bin(16384, 1, 16, 4096), 48)
    assert D.get_schedule(1 << 30) == (_bin(1 << 62, 1, 64, 16384), 48)
    with pytest.raises(ValueError):
        tune.validate({"bins": [_bin(10), _bin(5)]})
    with pytest.raises(ValueError):
        tune.validate({"bins": [{"nmax": 10, "knn_packed": 1}]})
