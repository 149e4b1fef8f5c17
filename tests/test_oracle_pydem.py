"""The reference's pydem tests restated against the CPU oracle, in the
reference's own precision (f32) and in FP64."""
import pytest

import pydem_cases as pc
from oracle import OracleSimulation, lib


@pytest.fixture(params=[32, 64], ids=["f32", "f64"])
def create(request):
    lib().rpo_set_threads(1)  # tiny scenes: threading only adds overhead

    def make(*args):
        return OracleSimulation.create(*args, bits=request.param)
    yield make
    lib().rpo_set_threads(0)


def test_free_fall(create):
    pc.free_fall(create)


def test_hertzian_compression(create):
    pc.hertzian_compression(create)


def test_rolling_incline(create, tmp_path):
    pc.rolling_incline(create, tmp_path)


def test_jkr_pulloff(create):
    pc.jkr_pulloff(create)


def test_sjkr_pulloff(create):
    pc.sjkr_pulloff(create)


def test_contact_models(create):
    pc.contact_models(create)


def test_periodic_column(create):
    pc.periodic_column(create)


def test_overlapping_column(create):
    lib().rpo_set_threads(4)
    pc.overlapping_column(create)
