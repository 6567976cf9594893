"""Batched solves, multi-seed search and the serving loop: every result must
equal the single-instance solve of the same request (bit-exact makespan,
sequence and expanded count), whatever else shares the device."""
import numpy as np
import pytest

import beamsearch_jssp_gpu_amd as bs
from testlib import orc_gpu, splitmix_instance

pytestmark = pytest.mark.gpu


def _b(i):
    return bs.JSSPInstance(i.machine, i.duration, i.name)


def _mixed(seed, count, ns=(10, 15, 20, 30, 50), ms=(5, 10, 15, 20), ks=(32, 64, 128), phases=2):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        n, m, K = int(rng.choice(ns)), int(rng.choice(ms)), int(rng.choice(ks))
        mode = int(rng.choice([bs.MODE_ICH_DET, bs.MODE_PMS_GPU]))
        inst = splitmix_instance(1000 + seed * 100 + t, n, m, 1, 99)
        out.append((inst, dict(beam_width=K, mode=mode, phases=phases, best_known=bs.NO_MAKESPAN, seed=7 + t)))
    return out


def _same(a, b):
    assert a.makespan == b.makespan and a.found == b.found and a.expanded_states == b.expanded_states
    if a.found:
        np.testing.assert_array_equal(a.sequence, b.sequence)


@pytest.mark.parametrize("phases", [2, 1])
def test_batch_equals_single_solves(engine, phases):
    """Uneven instances batched (several survivors per rank wave, ties of
    ICH_DET included) equal their single solves, two-phase and single-phase."""
    jobs = _mixed(1, 24, phases=phases)
    print("batch of", len(jobs), "phases", phases, flush=True)
    singles = [engine.solve(_b(i), **p) for i, p in jobs]
    for use_graph in (True, False):
        got = engine.solve_batch([_b(i) for i, _ in jobs], [p for _, p in jobs], use_graph=use_graph)
        for g, s in zip(got, singles):
            _same(g, s)
    # and against the oracle for a few
    for (inst, p), g in list(zip(jobs, got))[:4]:
        w = orc_gpu(inst, p["beam_width"], p["mode"], p["best_known"], p["seed"], trace=False)
        assert g.makespan == w["makespan"] and g.expanded_states == w["expanded"]
        np.testing.assert_array_equal(g.sequence, w["sequence"])


def test_batch_wide_beams_equal_single_solves(engine):
    """Beams of 1024 (one histogram copy; the throughput rank blocks stage
    the op table in LDS) and 2048 (two copies; op table read from memory)
    batched equal their single solves."""
    jobs = _mixed(5, 8, ns=(10, 15, 20), ms=(5, 10), ks=(1024, 2048))
    print("batch of", len(jobs), "K", sorted({p["beam_width"] for _, p in jobs}), flush=True)
    singles = [engine.solve(_b(i), **p) for i, p in jobs]
    got = engine.solve_batch([_b(i) for i, _ in jobs], [p for _, p in jobs])
    for g, s in zip(got, singles):
        _same(g, s)


def _shape(lanes, n):
    return lanes if (lanes == bs.LANES_FUSED or lanes >= n) else 64


def test_batch_lane_shapes_do_not_change_results(engine, tmp_path):
    """Each score shape (fused into the rank kernel, or 16 / 32 / 64 lanes per
    parent in a separate kernel) gives the same beams."""
    jobs = _mixed(2, 9, ns=(10, 15, 20, 30))
    want = [engine.solve(_b(i), **p) for i, p in jobs]
    for lanes in (bs.LANES_FUSED, 16, 32, 64):
        table = tmp_path / f"tune{lanes}.txt"
        with open(table, "w") as f:
            for inst, p in jobs:
                kb = int(np.ceil(np.log2(p["beam_width"])))
                f.write(f"{inst.n} {inst.m} {kb} {_shape(lanes, inst.n)}\n")
        eng = bs.Engine(0)
        eng.set_autotune(True, str(table))
        got = eng.solve_batch([_b(i) for i, _ in jobs], [p for _, p in jobs])
        for g, s in zip(got, want):
            _same(g, s)
        for inst, p in jobs:
            assert eng.tuned_lanes(inst.n, inst.m, p["beam_width"]) == _shape(lanes, inst.n)
        eng.close()


def test_batch_rank_shapes_and_filter_caps_do_not_change_results(engine, tmp_path):
    """Each rank block shape (8-wave blocks ranking four survivors per wave or
    one, 16-wave blocks) and filter block cap (16, 64, 256) the autotuner can
    choose gives the same beams, with the fused and a separate score shape."""
    jobs = _mixed(12, 10, ns=(10, 20, 30, 50), ks=(64, 256, 1024))
    want = [engine.solve(_b(i), **p) for i, p in jobs]
    for lanes in (bs.LANES_FUSED, 64):
        for rs, waves, nb in ((0, 8, 4), (1, 8, 1), (2, 16, 1)):
            for fc, blocks in ((1, 16), (3, 256)):
                table = tmp_path / f"tune_{lanes}_{rs}_{fc}.txt"
                with open(table, "w") as f:
                    for inst, p in jobs:
                        kb = int(np.ceil(np.log2(p["beam_width"])))
                        f.write(f"{inst.n} {inst.m} {kb} {_shape(lanes, inst.n)} {rs} {fc}\n")
                eng = bs.Engine(0)
                eng.set_autotune(True, str(table))
                print("shape", lanes, rs, fc, flush=True)
                got = eng.solve_batch([_b(i) for i, _ in jobs], [p for _, p in jobs])
                for g, w in zip(got, want):
                    _same(g, w)
                inst, p = jobs[0]
                sh = eng.tuned_shape(inst.n, inst.m, p["beam_width"])
                assert sh == dict(lanes=_shape(lanes, inst.n), rank_waves=waves, rank_batch=nb, filter_blocks=blocks)
                eng.close()


def test_autotuner_records_a_whole_shape(tmp_path):
    """A batch of classes not yet seen is tuned: every class gets a score
    shape, a rank block shape and a filter cap, written to the table file and
    read back by a new engine; the results equal the single solves."""
    jobs = _mixed(13, 6, ns=(15, 30), ms=(5, 10), ks=(256,))
    table = tmp_path / "tuned.txt"
    eng = bs.Engine(0)
    eng.set_autotune(True, str(table))
    got = eng.solve_batch([_b(i) for i, _ in jobs], [p for _, p in jobs])
    for (inst, p), g in zip(jobs, got):
        _same(g, eng.solve(_b(inst), **p))
        sh = eng.tuned_shape(inst.n, inst.m, p["beam_width"])
        assert sh is not None and sh["rank_waves"] in (8, 16) and sh["rank_batch"] in (1, 4)
        assert sh["filter_blocks"] in (16, 64, 256)
    eng.close()
    lines = [l.split() for l in open(table).read().splitlines()]
    assert lines and all(len(l) == 6 for l in lines)
    eng2 = bs.Engine(0)
    eng2.set_autotune(True, str(table))
    inst, p = jobs[0]
    eng2.solve_batch([_b(inst)], [p])  # loads the table, tunes nothing new
    assert eng2.tuned_shape(inst.n, inst.m, p["beam_width"]) is not None
    eng2.close()


def test_batch_repeated_runs_identical(engine):
    jobs = _mixed(3, 12)
    a = engine.solve_batch([_b(i) for i, _ in jobs], [p for _, p in jobs])
    engine.solve(_b(jobs[0][0]), **jobs[0][1])  # eager work in between
    b = engine.solve_batch([_b(i) for i, _ in jobs], [p for _, p in jobs])
    c = engine.solve_batch([_b(i) for i, _ in jobs], [p for _, p in jobs])
    for x, y, z in zip(a, b, c):
        _same(x, y)
        _same(x, z)


def test_multiseed_shared_bound(engine):
    inst = splitmix_instance(11, 15, 10, 1, 99)
    S, E, K = 6, 2, 64
    got = engine.solve(_b(inst), K, mode=bs.MODE_PMS_GPU, seed=100, n_seeds=S, n_passes=E)
    # sequential restatement of the pass structure with the oracle
    bound, best, best_idx, exp = bs.NO_MAKESPAN, None, -1, 0
    for e in range(E):
        pass_bound = bound
        for i in range(S):
            r = orc_gpu(inst, K, bs.MODE_PMS_GPU, pass_bound, 100 + e * S + i, trace=False)
            exp += r["expanded"]
            if r["found"] and r["makespan"] < bound:
                bound, best, best_idx = r["makespan"], r["sequence"], e * S + i
    assert got.makespan == bound
    assert got.expanded_states == exp
    assert got.best_seed_index == best_idx
    np.testing.assert_array_equal(got.sequence, best)


def test_serving_alone_equals_in_mix(engine):
    jobs = _mixed(4, 20, ks=(64, 128, 256))
    alone = [engine.solve(_b(i), **p) for i, p in jobs]
    srv = bs.Server(engine, hbm_budget_bytes=256 << 20, max_slots=8)
    tickets = []
    for k, (inst, p) in enumerate(jobs):
        tickets.append(srv.submit(_b(inst), p["beam_width"], p["mode"], p["best_known"], p["seed"]))
        if k % 3 == 2:
            srv.step(8)  # requests arrive while others are resident
    srv.drain()
    for t, a in zip(tickets, alone):
        _same(srv.poll(t), a)
        sub, adm, done = srv.times(t)
        assert 0 <= sub <= adm <= done
    srv.close()


def test_serving_wide_beams_equal_alone(engine):
    """Served requests with beams of 1024 and 2048 (one and two histogram
    copies, whatever rank block shape the server picks as the load changes)
    equal their single solves."""
    jobs = _mixed(6, 10, ns=(10, 15, 20), ms=(5, 10), ks=(256, 1024, 2048))
    alone = [engine.solve(_b(i), **p) for i, p in jobs]
    srv = bs.Server(engine, hbm_budget_bytes=1 << 30, max_slots=16)
    tickets = []
    for k, (inst, p) in enumerate(jobs):
        tickets.append(srv.submit(_b(inst), p["beam_width"], p["mode"], p["best_known"], p["seed"]))
        if k % 4 == 3:
            srv.step(8)
    srv.drain()
    for t, a in zip(tickets, alone):
        _same(srv.poll(t), a)
    srv.close()


def test_serving_deep_queue_equals_alone(engine):
    """Requests queued at once: those admitted behind eight or more waiting
    take the batched rank form, the last ones the per-survivor form; both
    run in the same steps and every result equals its single solve."""
    jobs = _mixed(6, 16, ks=(64, 128, 256, 512))
    alone = [engine.solve(_b(i), **p) for i, p in jobs]
    srv = bs.Server(engine, hbm_budget_bytes=512 << 20, max_slots=16)
    tickets = [srv.submit(_b(i), p["beam_width"], p["mode"], p["best_known"], p["seed"]) for i, p in jobs]
    srv.drain()
    for t, a in zip(tickets, alone):
        _same(srv.poll(t), a)
    srv.close()


def test_serving_budget_and_cancel(engine):
    inst = splitmix_instance(5, 20, 10, 1, 99)
    srv = bs.Server(engine, hbm_budget_bytes=4 << 20, max_slots=2)
    with pytest.raises(bs.JsspError) as e:
        srv.submit(_b(inst), 4096, bs.MODE_ICH_DET)
    assert e.value.code == 6
    t1 = srv.submit(_b(inst), 32, bs.MODE_ICH_DET, seed=1)
    t2 = srv.submit(_b(inst), 32, bs.MODE_ICH_DET, seed=2)
    t3 = srv.submit(_b(inst), 32, bs.MODE_ICH_DET, seed=3)
    srv.cancel(t2)
    srv.drain()
    assert srv.poll(t1) is not None and srv.poll(t3) is not None
    with pytest.raises(bs.JsspError) as e:
        srv.poll(t2)
    assert e.value.code == 9
    want = engine.solve(_b(inst), 32, mode=bs.MODE_ICH_DET, seed=3)
    _same(srv.poll(t3), want)
    srv.close()


def test_serving_steps_are_pipelined(engine):
    """A serve step returns with its levels still running and collects what
    the previous step finished: a request whose last level runs in one step
    is reported by the next call; a step of no levels waits for everything
    enqueued.  Results equal the plain solve either way."""
    inst = splitmix_instance(77, 10, 5, 1, 99)  # 50 levels
    p = dict(beam_width=64, mode=bs.MODE_PMS_GPU, phases=2, best_known=bs.NO_MAKESPAN, seed=3)
    alone = engine.solve(_b(inst), **p)
    srv = bs.Server(engine, hbm_budget_bytes=64 << 20, max_slots=4)
    t = srv.submit(_b(inst), 64, bs.MODE_PMS_GPU, seed=3)
    assert srv.step(64) == 1  # every level enqueued; nothing collected yet
    assert srv.poll(t) is None
    assert srv.step(64) == 0  # collects the previous step
    _same(srv.poll(t), alone)
    t2 = srv.submit(_b(inst), 64, bs.MODE_PMS_GPU, seed=3)
    srv.step(64)
    assert srv.step(0) == 0  # no levels: waits and collects
    _same(srv.poll(t2), alone)
    srv.close()


def test_large_snapshot_batch_and_serving(engine):
    """Instances whose snapshot exceeds 128 words or that have more than 64
    jobs take the large rank-kernel variant (register-limited in the batch
    and server form); mixed with small ones, each result equals its single
    solve, in a batch and through the server."""
    rng = np.random.default_rng(9)
    jobs = []
    for t in range(10):
        n, m = [(70, 5), (100, 8), (60, 10), (20, 10), (30, 15)][t % 5]
        K = int(rng.choice([64, 256]))
        mode = [bs.MODE_ICH_DET, bs.MODE_PMS_GPU][t % 2]
        inst = splitmix_instance(5000 + t, n, m, 1, 99)
        jobs.append((inst, dict(beam_width=K, mode=mode, phases=2, best_known=bs.NO_MAKESPAN, seed=11 + t)))
    print("large-snapshot batch of", len(jobs), flush=True)
    singles = [engine.solve(_b(i), **p) for i, p in jobs]
    got = engine.solve_batch([_b(i) for i, _ in jobs], [p for _, p in jobs])
    for g, s in zip(got, singles):
        _same(g, s)
    srv = bs.Server(engine, hbm_budget_bytes=512 << 20, max_slots=16)
    tickets = [srv.submit(_b(i), p["beam_width"], p["mode"], p["best_known"], p["seed"]) for i, p in jobs]
    srv.drain()
    for t, s in zip(tickets, singles):
        _same(srv.poll(t), s)
    srv.close()
