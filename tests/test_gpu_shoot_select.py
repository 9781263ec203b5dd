"""The select form of the shooting state machine (shoot_next_sel) against the branch form on the GPU, compiled for
gfx950 as the strong-shock kernel uses it (tests/shoot_check.hip in device mode): 4M single transitions and 256K whole
shootings, bit for bit."""
import pytest

from tests.test_shoot_select import run_check

pytestmark = pytest.mark.gpu


def test_select_form_equals_branch_form_on_the_gpu():
    out = run_check("device")
    assert out["single"][0] == 1 << 22 and out["single"][2] == 0, out
    assert out["chain"][0] == 1 << 18 and out["chain"][1] > 1 << 20 and out["chain"][2] == 0, out
