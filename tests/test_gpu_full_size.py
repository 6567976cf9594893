"""Parity at the benchmark shapes, against the oracle's committed results
(tests/golden/oracle_full_size.json, written in the build container by
tests/golden/make_oracle_goldens.py from oracle/oracle.cpp, whose survivor
order is pinned by tests/test_oracle.py):

* 50x20 K=4096 (about 1.6e8 children per solve), ICH_DET and PMS_GPU, single-
  and two-phase: makespan, sequence, expanded count, and digests of the
  parent counts and of every level's survivor trail (parent, job) -- the
  beams themselves, level by level;
* the 16-beam multi-seed search: every beam equals the oracle's solve of its
  seed, and the shared-bound result is their minimum;
* the 256-instance batch: 16 members equal the oracle, and every member
  equals its single solve.

The single- and two-phase paths share the selection kernels (the filter and
the exact selection digit), so their agreement alone would not catch a
selection bug; the oracle comparison does.  Size-independent properties are
checked too: the sequence re-simulates to the makespan and respects the
lower bound, and repeated solves are bit-identical."""
import hashlib
import json
import os

import numpy as np
import pytest

import beamsearch_jssp_gpu_amd as bs
from beamsearch_jssp_gpu_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_full_size.json")))
MODES = {"ICH_DET": bs.MODE_ICH_DET, "PMS_GPU": bs.MODE_PMS_GPU}


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()[:16]


def trail_digest(trail, pcount, K) -> str:
    parts = [trail[l * K: l * K + int(pcount[l + 1])] for l in range(len(pcount) - 2)]
    return sha(np.concatenate(parts))


def inst_of(spec):
    return bs.generate_instance(spec["gen_seed"], spec["n"], spec["m"], W.DMIN, W.DMAX)


def check_gold(r, g, what):
    assert r.makespan == g["makespan"], what
    assert r.found == g["found"], what
    assert r.expanded_states == g["expanded"], what
    if g["found"]:
        assert sha(r.sequence) == g["sequence_sha"], what


def _same(a, b):
    assert a.makespan == b.makespan and a.found == b.found and a.expanded_states == b.expanded_states
    if a.found:
        np.testing.assert_array_equal(a.sequence, b.sequence)


@pytest.mark.parametrize("phases", [2, 1])
@pytest.mark.parametrize("mode", ["ICH_DET", "PMS_GPU"])
def test_50x20_k4096_matches_oracle(engine, mode, phases):
    spec, g = GOLD["headline"]["spec"], GOLD["headline"][mode]
    inst = inst_of(spec)
    print("50x20 K=4096", mode, "phases", phases, flush=True)
    r = engine.solve(inst, spec["K"], mode=MODES[mode], phases=phases, seed=spec["seed"], trace=True)
    check_gold(r, g, (mode, phases))
    np.testing.assert_array_equal(r.sequence, np.array(g["sequence"], np.int32))
    assert sha(r.pcount) == g["pcount_sha"]
    assert trail_digest(r.trail, r.pcount, spec["K"]) == g["trail_sha"]
    assert bs.validate_sequence(inst, r.sequence) == r.makespan
    assert r.makespan >= bs.lower_bound(inst)
    assert r.levels == spec["n"] * spec["m"]
    again = engine.solve(inst, spec["K"], mode=MODES[mode], phases=phases, seed=spec["seed"])
    _same(r, again)


def test_16_beams_match_oracle(engine):
    ms = GOLD["multiseed"]
    spec, beams = ms["spec"], ms["beams"]
    inst = inst_of(spec)
    S = len(beams)
    print("16 beams", flush=True)
    got = engine.solve(inst, spec["K"], mode=bs.MODE_PMS_GPU, seed=spec["seeds"][0], n_seeds=S)
    singles = [engine.solve(inst, spec["K"], mode=bs.MODE_PMS_GPU, seed=s) for s in spec["seeds"]]
    for i, (r, g) in enumerate(zip(singles, beams)):
        check_gold(r, g, ("beam", i))
    best = min(range(S), key=lambda i: (beams[i]["makespan"] if beams[i]["found"] else bs.NO_MAKESPAN, i))
    assert got.makespan == beams[best]["makespan"]
    assert got.best_seed_index == best
    assert got.expanded_states == sum(b["expanded"] for b in beams)
    assert sha(got.sequence) == beams[best]["sequence_sha"]


def test_16_beams_two_passes_share_the_bound(engine):
    """Two passes of the 16 beams: pass 2 prunes with pass 1's best (the
    device atomic), which equals running each pass-2 beam alone with that
    bound."""
    ms = GOLD["multiseed"]
    spec = ms["spec"]
    inst = inst_of(spec)
    S = len(ms["beams"])
    bound = min(b["makespan"] for b in ms["beams"])
    seeds = np.array([spec["seeds"], [s + 100 for s in spec["seeds"]]], np.int64)
    got = engine.solve_seeded(inst, spec["K"], seeds, mode=bs.MODE_PMS_GPU)
    second = [engine.solve(inst, spec["K"], mode=bs.MODE_PMS_GPU, seed=int(s), best_known=bound) for s in seeds[1]]
    want = min([bound] + [r.makespan for r in second])
    assert got.makespan == want
    assert got.expanded_states == sum(b["expanded"] for b in ms["beams"]) + sum(r.expanded_states for r in second)
    assert got.best_seed_index >= S or want == bound


def test_256_instance_batch_matches_oracle(engine):
    specs = W.batch_specs()
    insts = [inst_of(s) for s in specs]
    params = [dict(beam_width=s["K"], mode=bs.MODE_PMS_GPU, phases=2, best_known=bs.NO_MAKESPAN, seed=s["seed"])
              for s in specs]
    print("batch of 256", flush=True)
    got = engine.solve_batch(insts, params)
    for g in GOLD["batch"]:
        assert g["spec"] == specs[g["index"]]
        check_gold(got[g["index"]], g, ("batch member", g["index"]))
    for i, (inst, p) in enumerate(zip(insts, params)):
        _same(got[i], engine.solve(inst, **p))
    # the graph-free step sequence gives the same results
    nog = engine.solve_batch(insts[:32], params[:32], use_graph=False)
    for i in range(32):
        _same(nog[i], got[i])
