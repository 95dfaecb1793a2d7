"""Host-side checks of the study workflows (no GPU).

* The optimizer rules against the reference's committed optimisation study
  (JAX + Optax 0.2.4): JAX's gradient at theta0 follows from the first SGD
  step, g0 = (theta0 - theta1) / 0.1, and fed to each optimizer must give
  the first step of that optimizer's committed trajectory.
* The rules against closed forms for a constant gradient g: bias-corrected
  Adam moves by -lr g / (|g| + eps) at every step; RMSprop by
  -lr g / sqrt((1 - 0.9^k) g^2 + eps) at step k.
* GradientBenchmark / ParameterOptimizationStudy on a cubic toy objective:
  central differences miss f'''(theta) eps^2 / 6 exactly; the batched and
  point-by-point paths give the same numbers; files carry the reference's
  keys.
"""
import json
import os

import numpy as np
import pytest

from ns_solver_jax_differential_distributed_amd import studies as ST

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_results.json")))
OPT = GOLD["optimization_study"]


def _g0():
    p = np.array(OPT["SGD (lr=0.1)"]["parameters"], np.float64)
    return (p[0] - p[1]) / 0.1


def _first_step(opt, g):
    th0 = np.array([1.0, 0.5], np.float32)
    u, _ = opt.update(np.asarray(g, np.float32), opt.init(th0))
    return ST.apply_updates(th0, u)


def test_sgd_first_step_reproduces_the_committed_trajectory():
    th1 = _first_step(ST.sgd(0.1), _g0())
    ref = np.array(OPT["SGD (lr=0.1)"]["parameters"][1], np.float32)
    assert np.all(np.abs(th1 - ref) <= np.spacing(ref))


@pytest.mark.parametrize("name,lr", [("Adam (lr=0.01)", 0.01), ("Adam (lr=0.001)", 0.001)])
def test_adam_first_step_reproduces_the_committed_trajectory(name, lr):
    # |g_x| = 2.6e-5: the step is lr |g| / (|g| + 1e-8) = 0.99962 lr only with
    # eps outside the square root (inside it would be 0.25 lr)
    th1 = _first_step(ST.adam(lr), _g0())
    ref = np.array(OPT[name]["parameters"][1], np.float32)
    assert np.max(np.abs(th1 - ref)) < 2e-7


def test_rmsprop_first_step_reproduces_the_committed_trajectory():
    # x: 0.1 g_x^2 << 1e-8, so the step is lr g_x / sqrt(1e-8) ~ 100 lr g_x
    # (eps inside the root); g_x from the SGD step is known to 2.3 %
    th1 = _first_step(ST.rmsprop(0.01), _g0())
    ref = np.array(OPT["RMSprop (lr=0.01)"]["parameters"][1], np.float64)
    step, ref_step = th1.astype(np.float64) - [1.0, 0.5], ref - [1.0, 0.5]
    assert abs(step[0] / ref_step[0] - 1) < 0.03
    assert abs(step[1] / ref_step[1] - 1) < 1e-5      # 0.01 / sqrt(0.1)


@pytest.mark.parametrize("g", [0.3, -2.5e-5, 4.0])
def test_adam_constant_gradient_closed_form(g):
    lr = 0.01
    opt = ST.adam(lr)
    st = opt.init(np.zeros(1, np.float32))
    for _ in range(6):
        u, st = opt.update(np.array([g], np.float32), st)
        # 2e-5: (1 - b2) is rounded to float32 from a double, 1 - b2^k is
        # computed in float32 (as Optax does); without bias correction the
        # first steps would be off by a factor ~3
        assert abs(float(u[0]) / (-lr * g / (abs(g) + 1e-8)) - 1) < 2e-5


@pytest.mark.parametrize("g", [0.3, -2.5e-5])
def test_rmsprop_constant_gradient_closed_form(g):
    lr = 0.01
    opt = ST.rmsprop(lr)
    st = opt.init(np.zeros(1, np.float32))
    for k in range(1, 8):
        u, st = opt.update(np.array([g], np.float32), st)
        want = -lr * g / np.sqrt((1 - 0.9 ** k) * g * g + 1e-8)
        assert abs(float(u[0]) / want - 1) < 2e-5


class _Cubic:
    """J(theta) = sum a_i theta_i^3 + b theta_0 theta_1, with a batch path."""

    a = np.array([0.7, -1.3])
    b = 0.4

    def __call__(self, th):
        th = np.asarray(th, np.float64)
        return float(np.sum(self.a * th ** 3) + self.b * th[0] * th[1])

    def batch(self, rows):
        return np.array([self(r) for r in rows])


class _CubicGrad:
    def __call__(self, th):
        th = np.asarray(th, np.float64)
        return 3 * _Cubic.a * th ** 2 + _Cubic.b * th[::-1]

    def value_and_grad_batch(self, rows):
        return (_TorchLike(np.array([_Cubic()(r) for r in rows])),
                _TorchLike(np.stack([self(r) for r in rows])))


class _TorchLike(np.ndarray):
    def __new__(cls, a):
        return np.asarray(a).view(cls)

    def detach(self):
        return self

    def cpu(self):
        return self

    def numpy(self):
        return np.asarray(self)


def test_gradient_benchmark_truncation_error_and_batched_path(tmp_path):
    # powers of two: theta +- eps are exact in float32, so FD - exact is the
    # truncation term alone
    eps = np.array([2.0 ** -10, 2.0 ** -7, 2.0 ** -4], np.float32)
    th = [1.0, 0.5]

    class NoBatch(_Cubic):
        batch = None

    gb = ST.GradientBenchmark(_Cubic(), _CubicGrad(), param_vectors=[th], epsilon_range=eps, jax_timing_runs=2)
    gb1 = ST.GradientBenchmark(NoBatch(), _CubicGrad(), param_vectors=[th], epsilon_range=eps, jax_timing_runs=2)
    r, r1 = gb.run_all()[0], gb1.run_all()[0]
    for e in eps:
        e = float(e)
        # FD - exact = a_i * e^2 per component (f''' = 6 a_i; the
        # theta_0 theta_1 term differences exactly)
        fd = r["epsilon_sweep"][e]["grad_fd"]
        assert np.allclose(fd - _CubicGrad()(th), _Cubic.a * e * e, rtol=1e-6, atol=1e-12)
        assert np.array_equal(fd, r1["epsilon_sweep"][e]["grad_fd"])
    gb.save_results(str(tmp_path / "gb.npz"))
    keys = set(np.load(tmp_path / "gb.npz").files)
    ref_keys = {k for k in GOLD["gradient_benchmark"] if k.startswith("param_0_")}
    assert keys == ref_keys


def test_optimization_study_lockstep_equals_sequential(tmp_path):
    cfg = lambda: [{"name": "Adam (lr=0.01)", "optimizer": ST.adam(0.01)},
                   {"name": "SGD (lr=0.1)", "optimizer": ST.sgd(0.1)},
                   {"name": "RMSprop (lr=0.01)", "optimizer": ST.rmsprop(0.01)}]

    class Plain:
        def __call__(self, th):
            return _CubicGrad()(th)

    a = ST.ParameterOptimizationStudy(_Cubic(), _CubicGrad(), [1.0, 0.5], cfg(), optim_steps=12).run_all()
    b = ST.ParameterOptimizationStudy(lambda t: _Cubic()(t), Plain(), [1.0, 0.5], cfg(), optim_steps=12).run_all()
    for n in a:
        assert np.array_equal(a[n]["history"]["parameters"], b[n]["history"]["parameters"])
        assert np.array_equal(a[n]["history"]["objectives"], b[n]["history"]["objectives"])
        assert a[n]["history"]["parameters"].shape == (13, 2)
    st = ST.ParameterOptimizationStudy(_Cubic(), _CubicGrad(), [1.0, 0.5], cfg(), optim_steps=3)
    st.run_all()
    st.save_results(str(tmp_path / "opt.npz"))
    f = np.load(tmp_path / "opt.npz")
    assert {k for k in f.files if k.startswith("SGD")} == {
        "SGD (lr=0.1)_" + s for s in ("obj_initial", "obj_final", "improvement", "steps", "objectives", "parameters")}
