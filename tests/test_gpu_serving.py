"""Continuous-batching server: results alone and in a mix are identical,
gradients match the batched adjoint, admission respects the HBM budget."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def M():
    import ns_solver_jax_differential_distributed_amd as pkg
    from ns_solver_jax_differential_distributed_amd import serving, engine
    return pkg, serving, engine


def _req(serving, N, steps, seed, grad=False):
    from ns_solver_jax_differential_distributed_amd import workloads as W
    L = 2 * math.pi
    th = W.theta_batch(1, seed=seed)[0]
    nu = float(W.nu_batch(1, seed=seed)[0])
    u0 = torch.from_numpy(W.initial_velocity(N, seed=seed, kmax=4)).cuda()
    dt = W.cfl_dt(u0.cpu().numpy(), L, 0.3, nu)
    return serving.SolveRequest(N=N, steps=steps, theta=th, nu=nu, dt=dt, u0=u0, want_grad=grad)


def test_served_request_equals_plain_solver_and_is_mix_independent():
    pkg, serving, _ = M()
    L = 2 * math.pi
    target = _req(serving, 64, 37, seed=1)
    # alone
    srv = serving.NSServer(chunk=8, max_steps=64, capacity={32: 4, 64: 4, 128: 4})
    alone = srv.run([target], realtime=False)[0]
    # plain per-request path
    s = pkg.SingleCoreNSSolver(pkg.make_grid(64, L), pkg.Phys(target.nu, 0.5, 0.1, 1.0),
                               pkg.Time(target.dt, target.steps))
    J, uf = s.rollout(target.u0, target.theta)
    assert torch.equal(alone.u_final, uf) and alone.J == float(J)
    # in a mix of sizes, step counts, with requests arriving mid-flight
    mix = [_req(serving, N, st, seed=10 + i) for i, (N, st) in
           enumerate([(32, 20), (64, 50), (128, 33), (64, 12), (32, 61), (128, 48), (64, 29)])]
    t2 = _req(serving, 64, 37, seed=1)
    reqs = mix[:3] + [t2] + mix[3:]
    for i, r in enumerate(reqs):
        r.arrival = 0.002 * i
    srv2 = serving.NSServer(chunk=8, max_steps=64, capacity={32: 2, 64: 3, 128: 2})
    res = {r.rid: r for r in srv2.run(reqs, realtime=True)}
    got = res[t2.rid]
    assert torch.equal(got.u_final, alone.u_final) and got.J == alone.J
    assert len(res) == len(reqs)
    for r in reqs:
        one = pkg.SingleCoreNSSolver(pkg.make_grid(r.N, L), pkg.Phys(r.nu, 0.5, 0.1, 1.0),
                                     pkg.Time(r.dt, r.steps))
        Jr, ur = one.rollout(r.u0, r.theta)
        assert torch.equal(res[r.rid].u_final, ur) and res[r.rid].J == float(Jr)


def test_shared_streams_and_timed_chunks_give_the_same_results():
    """streams="shared" (small buckets, large buckets and gradient groups on
    three streams) with chunks sized by GPU time: a mix of forward and
    gradient requests equals the plain per-request results bit for bit."""
    pkg, serving, engine = M()
    L = 2 * math.pi
    reqs = [_req(serving, N, st, seed=40 + i, grad=g) for i, (N, st, g) in
            enumerate([(32, 20, False), (64, 50, True), (256, 33, False), (64, 12, False),
                       (512, 26, False), (256, 40, True), (32, 61, True)])]
    for i, r in enumerate(reqs):
        r.arrival = 0.001 * i
    srv = serving.NSServer(chunk=None, streams="shared", max_steps=64,
                           capacity={32: 2, 64: 2, 256: 2, 512: 1})
    assert srv.chunk_for(512, 1) >= 2 and srv.chunk_for(32, 2) <= 64
    res = {r.rid: r for r in srv.run(reqs, realtime=True)}
    assert len(res) == len(reqs)
    for r in reqs:
        one = pkg.SingleCoreNSSolver(pkg.make_grid(r.N, L), pkg.Phys(r.nu, 0.5, 0.1, 1.0),
                                     pkg.Time(r.dt, r.steps))
        if r.want_grad:
            Jr, ur, g, _ = one.rollout_and_grad(r.u0, r.theta)
            assert np.array_equal(res[r.rid].dtheta, g[:2].cpu().numpy())
        else:
            Jr, ur = one.rollout(r.u0, r.theta)
        assert torch.equal(res[r.rid].u_final, ur) and res[r.rid].J == float(Jr)


def test_served_gradients_equal_single_request_gradients():
    pkg, serving, engine = M()
    L = 2 * math.pi
    reqs = [_req(serving, 32, st, seed=40 + i, grad=True) for i, st in enumerate([30, 17, 44])]
    reqs.append(_req(serving, 32, 25, seed=50))
    srv = serving.NSServer(chunk=8, max_steps=64, grad_group=4, grad_ckpt_every=7)
    res = {r.rid: r for r in srv.run(reqs, realtime=False)}
    for r in reqs[:3]:
        eng = engine.RolloutEngine(32, L, 1)
        J, uf, dth, du0 = eng.rollout_and_grad(r.u0, r.nu, r.dt, r.theta, r.steps, 7)
        out = res[r.rid]
        assert out.J == float(J[0]) and torch.equal(out.u_final, uf[0])
        assert np.array_equal(out.dtheta, dth[0].cpu().numpy())
        assert torch.equal(out.du0, du0[0])


def test_hbm_budget_bounds_admission():
    _, serving, _ = M()
    per = serving._Bucket.slot_bytes(64, 64, 4)
    srv = serving.NSServer(chunk=8, max_steps=64, capacity={64: 8}, hbm_budget=3 * per)
    reqs = [_req(serving, 64, 20 + i, seed=60 + i) for i in range(6)]
    out = srv.run(reqs, realtime=False)
    assert len(out) == 6
    assert srv.buckets[64].cap == 2          # halved until it fit the budget
    assert srv.budget.used <= srv.budget.limit


def test_a_request_that_blows_up_is_flagged_and_leaves_the_others_untouched():
    """One request with a NaN in u0 and one whose dt is far beyond the
    explicit-Euler limit, among ordinary ones and a gradient request in the
    same buckets: the two come back with ``finite`` False (J and, for the
    gradient, dJ/dtheta non-finite), every other result equals the same mix
    without them bit for bit, and nothing hangs or faults."""
    _, serving, _ = M()
    good = [_req(serving, N, st, seed=40 + i, grad=(i == 2)) for i, (N, st) in
            enumerate([(64, 30), (64, 44), (64, 26), (128, 21)])]
    nan_req = _req(serving, 64, 35, seed=50)
    u = nan_req.u0.clone()
    u[0, 7, 9] = float("nan")
    nan_req.u0 = u
    big_dt = _req(serving, 64, 40, seed=51, grad=True)
    big_dt.dt = 50.0 * big_dt.dt

    def run(reqs):
        srv = serving.NSServer(chunk=8, max_steps=64, capacity={64: 4, 128: 2})
        return {r.rid: r for r in srv.run(reqs, realtime=False)}

    clean = run([_req(serving, N, st, seed=40 + i, grad=(i == 2)) for i, (N, st) in
                 enumerate([(64, 30), (64, 44), (64, 26), (128, 21)])])
    mixed = run(good[:2] + [nan_req] + good[2:] + [big_dt])
    assert not mixed[nan_req.rid].finite and not mixed[big_dt.rid].finite
    assert not math.isfinite(mixed[nan_req.rid].J)
    for a, b in zip(sorted(clean.values(), key=lambda r: r.rid),
                    [mixed[r.rid] for r in good]):
        assert b.finite
        assert b.J == a.J and torch.equal(b.u_final, a.u_final)
        if a.dtheta is not None:
            assert np.array_equal(a.dtheta, b.dtheta) and torch.equal(a.du0, b.du0)


def test_checkpoint_policy_leaves_gradients_bit_identical_and_close_releases():
    """A gradient group keeps every state or checkpoints every k steps
    depending on keep_states_fraction; the reverse sweep recomputes the same
    float32 states either way, so results are bit-identical.  close() keeps
    returned results valid and refuses further requests."""
    _, serving, _ = M()
    out = []
    for frac in (0.0, 1.0):
        reqs = [_req(serving, 64, st, seed=70 + i, grad=True) for i, st in enumerate([41, 23, 64])]
        with serving.NSServer(chunk=8, max_steps=64, grad_group=4, grad_ckpt_every=6,
                              keep_states_fraction=frac) as srv:
            res = sorted(srv.run(reqs, realtime=False), key=lambda r: r.rid)
        assert srv.closed and not srv.buckets and not srv.lanes
        with pytest.raises(RuntimeError):
            srv.submit(_req(serving, 64, 5, seed=1))
        out.append(res)
    for a, b in zip(*out):
        assert a.J == b.J and np.array_equal(a.dtheta, b.dtheta)
        assert torch.equal(a.u_final, b.u_final) and torch.equal(a.du0, b.du0)
