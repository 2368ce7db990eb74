"""Host-side option translation of the Python mirror (no device calls):
SobolevStepOptions -> rc_step_opts (optimizer.ts:95-175 defaults: mode
'analytical', h = 1e-6, projectionMode 'reassemble'), and the argument checks
that run before any device work."""
import pytest

from rcurves import _lib as L
from rcurves import core


def test_sobolev_defaults_mirror_the_reference_function():
    o = core.SobolevStepOptions()
    assert o.projection_mode == "reassemble"      # optimizer.ts:144-160
    c = o.to_c()
    assert c.projection_frozen == 0
    assert c.gradient_mode == L.MODE_ANALYTICAL and c.fd_h == 1e-6
    assert (c.params.alpha, c.params.beta, c.params.epsilon) == (3.0, 6.0, 1e-10)


def test_frozen_and_finite_difference_translate():
    c = core.SobolevStepOptions(projection_mode="frozen", mode="finiteDiff", h=2e-6).to_c()
    assert c.projection_frozen == 1
    assert c.gradient_mode == L.MODE_FINITE_DIFF and c.fd_h == 2e-6


@pytest.mark.parametrize("kw", [{"projection_mode": "lazy"}, {"mode": "central"}])
def test_unknown_modes_are_rejected_before_device_work(kw):
    with pytest.raises(L.RcError) as err:
        core.SobolevStepOptions(**kw).to_c()
    assert err.value.status == L.RC_BAD_ARG


def test_raw_step_mode_checked_on_the_host():
    with pytest.raises(L.RcError, match="unknown gradient mode"):
        core.step([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], [[0, 1], [2, 3]], None,
                  core.StepOptions(mode="centralDiff", step_size=1e-3))
