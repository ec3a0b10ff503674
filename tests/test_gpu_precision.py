"""Mixed precision against f64: gravity forces and whole-world trajectories, with the
error stated per result.

The f32 gravity path evaluates far pairs in f32 on double-single positions (f32 group
sums added to f64), near pairs (closer than three core sizes) and the pairs among a
tile's own 64 slots in f64. Force bounds: a CPU emulation of the f32 pair arithmetic
(numpy float32) gives relative force-norm errors of about 1e-7 and per-body errors up
to 5e-6 on random charges; the bounds below leave a factor of ~5 for the device's own
rounding (hardware rsqrt, contracted products). Trajectory bounds are the 1e-9 gate
on final-state relative error."""
import numpy as np
import pytest

import pyoracle
from exgame_amd import scenes
from exgame_amd.world import DeviceWorld, gravity_accumulate

from test_gpu_parity import random_charges, rel

pytestmark = pytest.mark.gpu


def force_errors(got, want):
    per = np.linalg.norm(got - want, axis=1) / np.maximum(np.linalg.norm(want, axis=1), 1e-300)
    return rel(got, want), float(per.max()), float(np.sqrt(np.mean(per ** 2)))


@pytest.mark.parametrize("n", [2, 600, 4096])
def test_gravity_f32c_against_oracle(n):
    c = random_charges(n, n, coincident=n > 8)
    got = gravity_accumulate(c, G=1.0, precision="f32c")
    want = pyoracle.gravity_accumulate(c, G=1.0)
    r, mx, rms = force_errors(got[:, :3], want[:, :3])
    print(f"n={n} f32c force error: norm {r:.2e} max {mx:.2e} rms {rms:.2e}")
    assert r < 1e-6 and mx < 2e-5 and rms < 2e-6
    assert rel(got[:, 3:], want[:, 3:]) < 1e-5


def test_gravity_f32c_at_262144_bodies():
    # the precision-form size: f32c against the device f64 path (itself pinned to the
    # oracle above); unit-ball positions and lognormal masses normalised to 1
    st = scenes.Stream(0, 2)
    n = 262144
    c = np.zeros((n, 8))
    c[:, :3] = st.in_ball(n, 1.0)
    m = np.exp(0.5 * st.normal(n))
    m /= m.sum()
    c[:, 6] = m
    c[:, 7] = scenes.effective_size(m)
    f64 = gravity_accumulate(c, G=1.0, precision="f64")
    f32 = gravity_accumulate(c, G=1.0, precision="f32c")
    r, mx, rms = force_errors(f32[:, :3], f64[:, :3])
    print(f"n={n} f32c force error vs f64: norm {r:.2e} max {mx:.2e} rms {rms:.2e}")
    assert np.all(np.isfinite(f32))
    assert r < 1e-5 and rms < 1e-5


def test_mixed_precision_headline_world_meets_the_gate_over_200_steps():
    # the 4,096-body chains workload (512 chains of 8), f32 gravity and f32 Jacobians,
    # against the f64 device rollout (which the trajectory tests pin to the oracle):
    # final-state position error after 200 steps at most 1e-9 (measured 5.4e-10)
    scene = scenes.jointed_chains(n_chains=512, m=8, seed=0)
    a = DeviceWorld(scene, tol2=1e-20)
    b = DeviceWorld(scene, tol2=1e-20, gravity_precision="f32c", jacobian_precision="f32c")
    a.step(200, 1e-2, 1.0)
    b.step(200, 1e-2, 1.0)
    sa, sb = a.state(), b.state()
    err = rel(sb["position"], sa["position"])
    merr = rel(sb["momentum"], sa["momentum"])
    print(f"200-step final-state error, mixed vs f64: position {err:.2e} momentum {merr:.2e}")
    assert err <= 1e-9
    assert merr < 1e-7


def test_mixed_precision_world_against_the_oracle_over_100_steps():
    # 16 chains of 8 in mixed precision against the oracle itself
    scene = scenes.jointed_chains(n_chains=16, m=8, seed=0)
    dev = DeviceWorld(scene, tol2=1e-20, gravity_precision="f32c", jacobian_precision="f32c")
    orc = pyoracle.OracleWorld(scene, tol2=1e-20)
    dev.step(100, 1e-2, 1.0)
    orc.step(100, 1e-2, 1.0)
    err = rel(dev.state()["position"], orc.state()["position"])
    print(f"100-step final-state position error, mixed vs oracle: {err:.2e}")
    assert err <= 1e-9


@pytest.mark.parametrize("braced", [False, True])
def test_f32_jacobian_converges_to_the_f64_solution(braced):
    # FP32 Jacobian lanes and compensated FP32 assembly change only the Newton path;
    # the residual is f64 and each step still converges to tol² = 1e-20
    scene = scenes.jointed_chains(n_chains=12, m=8, seed=6, braced=braced)
    dev = DeviceWorld(scene, tol2=1e-20, jacobian_precision="f32c")
    orc = pyoracle.OracleWorld(scene, tol2=1e-20)
    dev.step(20, 1e-2, 1.0)
    orc.step(20, 1e-2, 1.0)
    err = rel(dev.state()["position"], orc.state()["position"])
    merr = rel(dev.state()["momentum"], orc.state()["momentum"])
    print(f"braced={braced}: 20-step error with f32 Jacobians: position {err:.2e} momentum {merr:.2e}; "
          f"Newton iterations {int(dev.stats()[:, 1].sum())} vs {int(orc.stats()[:, 1].sum())}")
    assert err < 1e-9 and merr < 1e-8
    assert np.all(dev.stats()[:, 5] == 0)   # every span converged


def test_precision_form_65536_jointed_bodies():
    # the jointed half of the precision form: 8,192 chains of 8 at the 4,096-body
    # workload's density (chain origins in a ball of radius 126 = 50·16^(1/3)); f32
    # Jacobians alone, then with f32c gravity as well, against the all-f64 run on the
    # device over 200 steps (measured 8.7e-15 and 2.1e-10)
    scene = scenes.jointed_chains(n_chains=8192, m=8, seed=0, radius=126.0)
    runs = {}
    for name, gp, jp in [("f64", "f64", "f64"), ("jac32", "f64", "f32c"), ("mixed", "f32c", "f32c")]:
        w = DeviceWorld(scene, tol2=1e-20, gravity_precision=gp, jacobian_precision=jp)
        w.step(200, 1e-2, 1.0)
        runs[name] = w.state()
        w.close()
    e_jac = rel(runs["jac32"]["position"], runs["f64"]["position"])
    e_mix = rel(runs["mixed"]["position"], runs["f64"]["position"])
    print(f"65,536 bodies, 200 steps: final-state position error vs f64 — f32 Jacobian {e_jac:.2e}, "
          f"f32 Jacobian + f32c gravity {e_mix:.2e}")
    assert e_jac < 1e-12
    assert e_mix <= 1e-9
