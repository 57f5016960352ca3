"""CPU: the oracle against independent truths (float64 brute force, closed
forms, ray-plane intersections, hand-computed conversions)."""
import pytest

import oracle_pins as P


@pytest.mark.parametrize("pin", P.ALL_PINS, ids=lambda p: p.__name__)
def test_pin(pin):
    ok, msg = pin()
    assert ok, msg
