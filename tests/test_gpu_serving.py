"""Batched serving (lbx_serve_*): each run's result is the same alone and in a
mix (bitwise), refined (two-level) runs included, FP64 serving matches the
oracle, admission respects the arena, and bad requests are refused."""
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


def fine_boxes_in(lo, hi, edge):
    """Fine boxes (fine index space) tiling fine cells lo..hi in cubes of `edge`."""
    out = []
    for z in range(lo[2], hi[2] + 1, edge):
        for y in range(lo[1], hi[1] + 1, edge):
            for x in range(lo[0], hi[0] + 1, edge):
                out.append(((x, y, z), (min(x + edge, hi[0] + 1) - 1, min(y + edge, hi[1] + 1) - 1,
                                        min(z + edge, hi[2] + 1) - 1)))
    return out


REFINED = [  # dims, level-0 max grid size (0: one box), fine region (fine index space), steps, tau, amplitude
    ((32, 24, 16), 8, ((16, 8, 8), (47, 39, 23)), 6, 0.8, 1e-2),
    ((24, 24, 24), 0, ((8, 8, 8), (39, 31, 39)), 9, 0.7, 3e-3),
]


def alone_refined(ctx, dims, mgs, region, steps, tau, amp, precision):
    """The same two-level run through lbx_sim + lbx_iterate: level-0 f and sums."""
    boxes = workloads.chop(dims, mgs) if mgs else [((0, 0, 0), tuple(d - 1 for d in dims))]
    fine = fine_boxes_in(*region, 16)
    sim = lbx.Sim(ctx, dims, precision=precision)
    try:
        sim.define_level(0, boxes)
        sim.set_tau(0, tau)
        rho, u = workloads.plane_pulse(dims, amp, axis=0, plane=dims[0] // 2)
        sim.init_from_fields(rho, u)
        sim.define_level(1, fine)
        sim.set_tau(1, 2 * (tau - 0.5) + 0.5)
        sim.fill_from_coarse(1)
        sim.set_fine_mask(fine)
        sim.iterate(steps)
        return sim.get_f(0), sim.conservation(0), sim.get_hydro(0)[0]
    finally:
        sim.close()


@pytest.mark.parametrize("precision", [lbx.LBX_FP32_SHIFTED, lbx.LBX_FP64])
def test_refined_runs_in_a_mix_equal_each_run_alone(ctx, precision):
    """Two-level runs served beside single-level ones: each refined run's
    level-0 populations, density and sums equal the same hierarchy run
    alone through lbx_iterate, bit for bit; the single-level runs still
    equal theirs."""
    srv = lbx.Server(ctx, precision=precision, arena_bytes=1 << 30)
    singles, refined = {}, {}
    for k, (dims, steps, tau, amp) in enumerate(MIX[:4]):
        rid, o = srv.submit(dims, steps, tau, amplitude=amp, axis=0, want_f=True)
        singles[rid] = (dims, steps, tau, amp, o)
        if k < len(REFINED):
            dims_r, mgs, region, steps_r, tau_r, amp_r = REFINED[k]
            boxes = workloads.chop(dims_r, mgs) if mgs else None
            rid, o = srv.submit(dims_r, steps_r, tau_r, amplitude=amp_r, axis=0, want_f=True, want_rho=True,
                                boxes=boxes, fine_boxes=fine_boxes_in(*region, 16))
            refined[rid] = (REFINED[k], o)
    srv.drain()
    res = {r["id"]: r for r in srv.poll()}
    stats = srv.stats()
    srv.close()
    assert sorted(res) == sorted(list(singles) + list(refined))
    assert all(r["status"] == lbx.LBX_OK for r in res.values())
    for rid, (dims, steps, tau, amp, o) in singles.items():
        assert bitwise_equal(o["f"], alone_plain(ctx, dims, steps, tau, amp, precision)), (dims, steps)
    lu = sum(np.prod(d) * st for d, st, _, _, _ in singles.values())
    for rid, (case, o) in refined.items():
        dims, mgs, region, steps, tau, amp = case
        f0, sums, rho0 = alone_refined(ctx, dims, mgs, region, steps, tau, amp, precision)
        assert bitwise_equal(o["f"], f0), case
        assert bitwise_equal(o["rho"], rho0), case
        assert bitwise_equal(np.array([res[rid]["mass"], *res[rid]["momentum"]]), sums), case
        assert res[rid]["steps"] == steps and res[rid]["cells"] == np.prod(dims)
        fine_cells = np.prod([h - l + 1 for l, h in zip(*region)])
        lu += (np.prod(dims) + 2 * fine_cells) * steps
    assert stats["lattice_updates"] == lu


@pytest.mark.parametrize("precision", [lbx.LBX_FP32_SHIFTED, lbx.LBX_FP64])
def test_refined_run_at_ratio_3_equals_it_alone(ctx, precision):
    """A two-level run at refinement ratio 3 (ref_ratio = 3: fine boxes in
    the index space 3 * dims, 3 fine substeps per coarse step) served beside
    a single-level run equals the same hierarchy through lbx_iterate alone,
    bit for bit, and the lattice updates count 3 fine substeps."""
    dims, steps, tau, amp = (18, 12, 12), 5, 0.8, 1e-2
    fine = fine_boxes_in((18, 9, 0), (35, 26, 35), 9)  # coarse x in [6, 12), y in [3, 9), all z
    srv = lbx.Server(ctx, precision=precision, arena_bytes=256 << 20)
    single, os_ = srv.submit((16, 12, 10), 7, 0.7, amplitude=1e-3, want_f=True)
    rid, o = srv.submit(dims, steps, tau, amplitude=amp, axis=0, want_f=True, want_rho=True, fine_boxes=fine,
                        ref_ratio=3)
    srv.drain()
    res = {r["id"]: r for r in srv.poll()}
    stats = srv.stats()
    srv.close()
    assert res[rid]["status"] == lbx.LBX_OK and res[single]["status"] == lbx.LBX_OK
    sim = lbx.Sim(ctx, dims, precision=precision)
    try:
        sim.define_level(0, [((0, 0, 0), tuple(d - 1 for d in dims))])
        sim.set_tau(0, tau)
        rho, u = workloads.plane_pulse(dims, amp, axis=0, plane=dims[0] // 2)
        sim.init_from_fields(rho, u)
        sim.set_ref_ratio(3)
        sim.define_level(1, fine)
        sim.set_tau(1, 3 * (tau - 0.5) + 0.5)
        sim.fill_from_coarse(1)
        sim.set_fine_mask(fine)
        sim.iterate(steps)
        f0, rho0 = sim.get_f(0), sim.get_hydro(0)[0]
    finally:
        sim.close()
    assert bitwise_equal(o["f"], f0)
    assert bitwise_equal(o["rho"], rho0)
    assert bitwise_equal(os_["f"], alone_plain(ctx, (16, 12, 10), 7, 0.7, 1e-3, precision))
    fine_cells = sum(np.prod([h - l + 1 for l, h in zip(a, b)]) for a, b in fine)
    assert stats["lattice_updates"] == 16 * 12 * 10 * 7 + (np.prod(dims) + 3 * fine_cells) * steps


def test_refined_request_with_bad_boxes_is_delivered_as_an_error(ctx):
    """A fine box that is not aligned to ratio 2 fails when the run is set
    up: the request comes back with an error status and the server goes on
    with the others."""
    srv = lbx.Server(ctx, precision=lbx.LBX_FP64, arena_bytes=64 << 20)
    bad, _ = srv.submit((16, 16, 16), 3, 0.8, amplitude=1e-3, fine_boxes=[((1, 0, 0), (8, 7, 7))])
    good, o = srv.submit((16, 16, 16), 3, 0.8, amplitude=1e-3, want_f=True)
    srv.drain()
    res = {r["id"]: r for r in srv.poll()}
    srv.close()
    assert res[bad]["status"] == lbx.LBX_E_ARG
    assert res[good]["status"] == lbx.LBX_OK
    assert bitwise_equal(o["f"], alone_plain(ctx, (16, 16, 16), 3, 0.8, 1e-3, lbx.LBX_FP64))


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
    for ratio in (1, 5, -2):  # refinement ratio outside 2..4
        with pytest.raises(lbx.LbxError) as e:
            srv.submit((8, 8, 8), 5, 0.8, fine_boxes=[((0, 0, 0), (7, 7, 7))], ref_ratio=ratio)
        assert e.value.code == lbx.LBX_E_ARG, ratio
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
