"""CLI and profiler contracts on the CPU (no GPU work).

The runner's argument parser and output naming follow the reference's
runner tests (unit_tests/test_distributed_solver.py:629-683, runner.py:31-56,
162-164); ProfileSummary / benchmark_callable follow its profiler API tests
(:569-622); the shipped sample configs give the reference's dt and step
counts (config_N_128: dt = f32(0.003125), 160 steps).
"""
import json
import os

import pytest

from ns_solver_jax_differential_distributed_amd import NS_solver as S
from ns_solver_jax_differential_distributed_amd import runner
from ns_solver_jax_differential_distributed_amd.solver_profiler import ProfileSummary, SolverKernelProfiler

PKG = os.path.dirname(os.path.abspath(S.__file__))
CFG = os.path.join(PKG, "sample_config")


def test_empty_args_give_defaults():
    a = runner.build_parser().parse_args([])
    assert a.config is None and a.profile is False
    assert a.profile_warmup == 1 and a.profile_runs == 5


def test_combined_flags():
    a = runner.build_parser().parse_args(["--config", "my.json", "--profile", "--profile-warmup", "2",
                                          "--profile-runs", "8"])
    assert (a.config, a.profile, a.profile_warmup, a.profile_runs) == ("my.json", True, 2, 8)


@pytest.mark.parametrize("argv", [["--profile-warmup", "notanint"], ["--profile-runs", "3.5"]])
def test_non_integer_counts_are_rejected(argv):
    with pytest.raises(SystemExit):
        runner.build_parser().parse_args(argv)


def test_output_naming():
    npy, png = runner.output_paths("result", 512)
    assert npy == os.path.join("result", "velocity_magnitude_single_N512.npy")
    assert png == os.path.join("result", "velocity_magnitude_single_N512.png")


@pytest.mark.parametrize("N,steps", [(128, 160), (256, 320), (512, 640), (1024, 1280), (2048, 2560),
                                     (4096, 5120)])
def test_sample_configs(N, steps):
    # cfl 0.4, v_max 1, L 1: dt = 0.4 / N (the diffusive limit 1/(4 nu N^2) is
    # larger at nu = 1e-4 for every N here), steps = ceil(0.5 / dt)
    cfg = S.load_config(os.path.join(CFG, f"config_N_{N}.json"))
    assert cfg.N == N and cfg.nu == 1e-4 and cfg.advection_scheme == "central"
    assert cfg.steps == steps
    assert cfg.dt == float(__import__("numpy").float32(0.4 / N))
    with open(os.path.join(CFG, f"config_N_{N}.json")) as f:
        assert set(json.load(f)) == {"N", "nu", "t", "advection_scheme"}


def test_benchmark_callable_contract():
    calls = []

    def counted():
        calls.append(1)
        return 0

    r = SolverKernelProfiler.benchmark_callable("count", counted, warmup=2, runs=5)
    assert isinstance(r, ProfileSummary) and r.label == "count"
    assert len(calls) == 7                        # 2 warmup + 5 measured
    assert 0.0 < r.min_sec <= r.mean_sec <= r.max_sec
    r0 = SolverKernelProfiler.benchmark_callable("no-warmup", lambda: 0, warmup=0, runs=2)
    assert isinstance(r0, ProfileSummary)


def test_profile_summary_fields():
    s = ProfileSummary(label="s", mean_sec=0.5, min_sec=0.4, max_sec=0.6, std_sec=0.05)
    assert (s.label, s.mean_sec, s.min_sec, s.max_sec, s.std_sec) == ("s", 0.5, 0.4, 0.6, 0.05)
