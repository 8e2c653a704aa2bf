"""Learning-rate schedule and argument checks of the warm-start training tool (host code; runs without a GPU)."""
import numpy as np
import pytest

from gpu_testbed_poisson_gravity_amd import train_warm_start as T


@pytest.mark.parametrize("decay_frac", [0.25, 0.5, 1.0])
def test_schedule_is_constant_then_linear_to_a_tenth(decay_frac):
    steps, lr = 1000, 2e-3
    rates = np.array([T.learning_rate(s, steps, lr, decay_frac) for s in range(steps)])
    assert rates[-1] == pytest.approx(lr / 10)
    start = int(np.ceil((1.0 - decay_frac) * (steps - 1)))     # first step of the decay
    assert np.all(rates[:start] == lr)
    step_drop = 0.9 / (decay_frac * (steps - 1))               # relative decrement per step
    assert lr * (1 - step_drop) <= rates[start] <= lr          # the decay starts where the constant part ends
    # linear: equal decrements over the decay
    d = np.diff(rates[start:])
    assert np.all(d < 0)
    assert np.allclose(d, d[0], rtol=1e-6, atol=1e-12)
    assert d[0] == pytest.approx(-step_drop * lr, rel=1e-6)


@pytest.mark.parametrize("bad", ["0", "-0.1", "1.5"])
def test_decay_fraction_outside_unit_interval_is_refused(bad):
    # refused by the argument check, before any patch is generated or the GPU is touched
    with pytest.raises(SystemExit):
        T.main(["--decay-frac", bad])
