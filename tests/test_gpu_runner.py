"""The CLI and the stage profiler on the GPU.

* runner.main on the shipped config_N_128.json writes the speed field the
  reference's runner wrote for the same config
  (result/velocity_magnitude_single_N128.npy, a committed fixture) under the
  reference's file name, plus its .png.
* --profile: every stage row of the reference's single-core table, and the
  two fused passes, with positive, ordered times.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _cfg(N):
    from ns_solver_jax_differential_distributed_amd import NS_solver as S
    return os.path.join(os.path.dirname(os.path.abspath(S.__file__)), "sample_config", f"config_N_{N}.json")


def test_runner_writes_the_reference_speed_field(tmp_path):
    from ns_solver_jax_differential_distributed_amd import runner
    out = runner.main(["--config", _cfg(128), "--out", str(tmp_path)])
    npy = tmp_path / "velocity_magnitude_single_N128.npy"
    assert out["npy"] == str(npy) and out["steps"] == 160
    speed = np.load(npy)
    ref = np.load(os.path.join(HERE, "golden", "speed_single_N128.npy"))
    assert speed.shape == ref.shape == (128, 128) and speed.dtype == np.float32
    assert np.max(np.abs(speed - ref)) < 2e-6
    if out["png"] is not None:
        assert os.path.getsize(out["png"]) > 0


def test_profile_rows(tmp_path):
    from ns_solver_jax_differential_distributed_amd import NS_solver as S
    from ns_solver_jax_differential_distributed_amd.solver_profiler import profile_solver_kernels
    cfg = S.load_config(_cfg(256))
    solver = S.SingleCoreNSSolver(S.make_grid(cfg.N, cfg.L), S.Phys(cfg.nu, cfg.mu0, cfg.sigma0, cfg.C_force),
                                  S.Time(cfg.dt, cfg.steps))
    rows = profile_solver_kernels(solver, None, theta=cfg.theta0, warmup=1, runs=3)
    labels = [r.label for r in rows]
    assert labels == ["single: forcing", "single: advection_diffusion", "single: pressure_solve",
                      "single: projection", "single: full_step", "single: full_rollout",
                      "single: row_pass (fused)", "single: column_pass (fused)"]
    for r in rows:
        assert 0.0 < r.min_sec <= r.mean_sec <= r.max_sec
    # a rollout of 320 steps costs more than one step
    assert rows[5].mean_sec > rows[4].mean_sec
    with pytest.raises(ValueError):
        profile_solver_kernels(solver, object())


def test_runner_profile_flag(tmp_path):
    from ns_solver_jax_differential_distributed_amd import runner
    out = runner.main(["--config", _cfg(128), "--out", str(tmp_path), "--profile", "--profile-warmup", "0",
                       "--profile-runs", "2"])
    assert os.path.exists(out["npy"])
