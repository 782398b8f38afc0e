"""Batched serving (lbx_serve_*): each run's result is the same alone and in a
mix (bitwise), FP64 serving matches the oracle, admission respects the
arena, and bad requests are refused."""
import numpy as np
import pytest

from lambrex_amd import lbx, workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

MIX = [  # dims, steps, tau, amplitude
    ((32, 32, 32), 17, 0.55, 1e-2),
    ((48, 48, 48), 9, 1.2, 3e-3),
    ((16, 24, 20), 31, 0.8, 1e-4),
    ((64, 64, 64), 5, 0.9, 5e-3),
    ((20, 16, 12), 1, 1.5, 2e-3),
    ((96, 8, 4), 12, 0.6, 1e-3),
    ((300, 5, 3), 7, 0.7, 1e-3),  # rows longer than a work item; last item partial
]


def alone_plain(ctx, dims, steps, tau, amp, precision):
    """The same run through the single-run API (lbx_sim + lbx_iterate)."""
    sim = lbx.Sim(ctx, dims, precision=precision)
    sim.define_level(0, workloads.chop(dims, 64))
    sim.set_tau(0, tau)
    rho, u = workloads.plane_pulse(dims, amp, axis=0, plane=dims[0] // 2)
    sim.init_from_fields(rho, u)
    sim.iterate(steps)
    f = sim.get_f()
    sim.close()
    return f


def bitwise_equal(a, b):
    return bool((np.ascontiguousarray(a).view(np.uint64) == np.ascontiguousarray(b).view(np.uint64)).all())


@pytest.mark.parametrize("precision", [lbx.LBX_FP32_SHIFTED, lbx.LBX_FP64])
def test_mix_equals_each_run_alone(ctx, precision):
    srv = lbx.Server(ctx, precision=precision, arena_bytes=1 << 30)
    outs = {}
    for dims, steps, tau, amp in MIX:
        rid, o = srv.submit(dims, steps, tau, amplitude=amp, axis=0, want_f=True, want_rho=True)
        outs[rid] = (dims, steps, tau, amp, o)
    srv.drain()
    res = {r["id"]: r for r in srv.poll()}
    srv.close()
    assert sorted(res) == sorted(outs)
    for rid, (dims, steps, tau, amp, o) in outs.items():
        f_alone = alone_plain(ctx, dims, steps, tau, amp, precision)
        assert bitwise_equal(o["f"], f_alone), (dims, steps)
        assert res[rid]["steps"] == steps and res[rid]["cells"] == np.prod(dims)
        assert abs(res[rid]["mass"] - o["rho"].sum()) <= 1e-9 * res[rid]["mass"]


def test_fp64_serving_matches_oracle(ctx):
    dims, steps, tau = (12, 10, 8), 6, 0.7
    rho, u = workloads.random_state(dims, seed=31, rho_sigma=1e-2, u_amp=0.02)
    srv = lbx.Server(ctx, precision=lbx.LBX_FP64, arena_bytes=64 << 20)
    rid, o = srv.submit(dims, steps, tau, rho=rho, u=u, want_f=True)
    srv.drain()
    srv.poll()
    srv.close()
    s = orc.Sim(*dims, tau, boxes=[([0, 0, 0], [d - 1 for d in dims])])
    s.init(np.ascontiguousarray(rho.transpose(2, 1, 0)).reshape(-1),
           np.ascontiguousarray(u.transpose(3, 2, 1, 0)).reshape(-1))
    s.iterate(steps)
    assert bitwise_equal(o["f"], s.get_f())
    s.close()


def test_admission_respects_the_arena(ctx):
    """Arena for two 32^3 FP32 runs: never more than two active, FIFO
    admission, every request completes."""
    run_bytes = 2 * 15 * 32 ** 3 * 4
    srv = lbx.Server(ctx, precision=lbx.LBX_FP32_SHIFTED, arena_bytes=2 * run_bytes + 4096)
    ids = [srv.submit((32, 32, 32), 20 + 3 * k, 0.8, amplitude=1e-3)[0] for k in range(6)]
    max_active = 0
    results = []
    for _ in range(10000):
        active, queued = srv.progress(4)
        max_active = max(max_active, active)
        results += srv.poll()
        if len(results) == len(ids):
            break
    srv.close()
    assert len(results) == len(ids)
    assert max_active <= 2
    by_id = {r["id"]: r for r in results}
    admits = [by_id[i]["t_admit"] for i in ids]
    assert admits == sorted(admits)
    for r in results:
        assert r["t_submit"] <= r["t_admit"] <= r["t_done"]


def test_bad_requests_are_refused(ctx):
    srv = lbx.Server(ctx, precision=lbx.LBX_FP32_SHIFTED, arena_bytes=16 << 20)
    with pytest.raises(lbx.LbxError) as e:
        srv.submit((160, 128, 128), 5, 0.8)  # FP64 populations exceed the 256 MB staging buffer
    assert e.value.code == lbx.LBX_E_ARG
    with pytest.raises(lbx.LbxError) as e:
        srv.submit((8, 8, 8), 0, 0.8)
    assert e.value.code == lbx.LBX_E_ARG
    with pytest.raises(lbx.LbxError) as e:
        srv.submit((128, 128, 128), 5, 0.8)  # 252 MB > 16 MB arena
    assert e.value.code == lbx.LBX_E_NOMEM
    for dims, steps, tau, plane in (((1 << 30, 1 << 30, 1 << 30), 5, 0.8, 0),  # cell count overflows int64
                                    ((8, 8, 8), 5, -0.5, 4),                    # omega = 1/(tau + 1/2) infinite
                                    ((8, 8, 8), 5, float("nan"), 4),
                                    ((8, 8, 8), 5, 0.8, 8)):                    # pulse plane outside the run
        with pytest.raises(lbx.LbxError) as e:
            srv.submit(dims, steps, tau, plane=plane)
        assert e.value.code == lbx.LBX_E_ARG, (dims, tau, plane)
    srv.close()


def test_shortest_first_admission(ctx):
    """LBX_SERVE_SHORTEST_FIRST: with one run active at a time, queued runs
    are admitted in order of work (cells x steps, ties by arrival), and each
    run's populations are the same bits as under FIFO admission."""
    reqs = [((32, 32, 32), 40), ((16, 16, 16), 10), ((24, 24, 24), 20), ((16, 16, 16), 10), ((32, 32, 32), 5)]
    out = {}
    for policy in (lbx.LBX_SERVE_FIFO, lbx.LBX_SERVE_SHORTEST_FIRST):
        srv = lbx.Server(ctx, precision=lbx.LBX_FP64, arena_bytes=1 << 28, max_active=1, policy=policy)
        subs = [srv.submit(d, n, 0.8, amplitude=1e-3, want_f=True) for d, n in reqs]
        srv.drain()
        res = {r["id"]: r for r in srv.poll()}
        srv.close()
        assert len(res) == len(reqs)
        admit_order = sorted(range(len(reqs)), key=lambda k: res[subs[k][0]]["t_admit"])
        out[policy] = (admit_order, [o["f"] for _, o in subs])
    assert out[lbx.LBX_SERVE_FIFO][0] == [0, 1, 2, 3, 4]
    work = [np.prod(d) * n for d, n in reqs]
    assert out[lbx.LBX_SERVE_SHORTEST_FIRST][0] == sorted(range(len(reqs)), key=lambda k: (work[k], k))
    for a, b in zip(out[lbx.LBX_SERVE_FIFO][1], out[lbx.LBX_SERVE_SHORTEST_FIRST][1]):
        assert bitwise_equal(a, b)


def test_shortest_first_with_wait_limit_serves_overdue_runs_first(ctx):
    """With a wait limit every queued run has exceeded, shortest-first admits
    in arrival order (the overdue rule), one run active at a time."""
    reqs = [((32, 32, 32), 40), ((16, 16, 16), 10), ((24, 24, 24), 20), ((16, 16, 16), 5)]
    srv = lbx.Server(ctx, precision=lbx.LBX_FP32_SHIFTED, arena_bytes=1 << 28, max_active=1,
                     policy=lbx.LBX_SERVE_SHORTEST_FIRST, max_wait_s=1e-9)
    ids = [srv.submit(d, n, 0.8, amplitude=1e-3)[0] for d, n in reqs]
    srv.drain()
    res = {r["id"]: r for r in srv.poll()}
    srv.close()
    admits = [res[i]["t_admit"] for i in ids]
    assert admits == sorted(admits)


def test_unknown_policy_is_refused(ctx):
    srv = lbx.Server(ctx, precision=lbx.LBX_FP32_SHIFTED, arena_bytes=16 << 20)
    with pytest.raises(lbx.LbxError) as e:
        srv.set_policy(7)
    assert e.value.code == lbx.LBX_E_ARG
    for bad in (-1.0, float("nan"), float("inf")):
        with pytest.raises(lbx.LbxError) as e:
            srv.set_policy(lbx.LBX_SERVE_SHORTEST_FIRST, bad)
        assert e.value.code == lbx.LBX_E_ARG
    srv.close()
