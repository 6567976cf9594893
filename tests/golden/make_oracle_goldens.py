"""Writes tests/golden/oracle_full_size.json: the oracle's results (CPU
restatement of the GPU reference's semantics, oracle/oracle.cpp) at the
benchmark shapes, so the GPU tests compare the HIP path with the oracle
there without running it (a 50x20 K=4096 operation-level solve takes the
oracle ~15 s).

    python tests/golden/make_oracle_goldens.py      (about 1-2 minutes)

Cases (workload specs from beamsearch_jssp_gpu_amd/workloads.py; instances
from the Python statement of the generator in tests/testlib.py):
  headline   50x20 K=4096, instance seed 1, solve seed 12345, ICH_DET and
             PMS_GPU: makespan, expanded count, the sequence, and SHA-256
             digests of the parent counts and of every level's survivor trail
  multiseed  16 PMS_GPU beams on that instance, solve seeds 1..16
  batch      16 members of the 256-instance batch (PMS_GPU, seed 12345)
  serving    the first 10 requests of the serving mix with K >= 1024
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.dirname(HERE))

from beamsearch_jssp_gpu_amd import workloads as W  # noqa: E402
from testlib import orc_gpu, splitmix_instance  # noqa: E402

ICH_DET, PMS_GPU = 1, 2
NO = 500000


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()[:16]


def trail_digest(trail, pcount, K) -> str:
    """Digest of the written trail entries: level l's first pcount[l+1]."""
    parts = [trail[l * K: l * K + int(pcount[l + 1])] for l in range(len(pcount) - 2)]
    return sha(np.concatenate(parts) if parts else np.zeros(0, np.int32))


def solve(spec, mode=PMS_GPU, trace=False, seed=None):
    inst = splitmix_instance(spec["gen_seed"], spec["n"], spec["m"], W.DMIN, W.DMAX)
    r = orc_gpu(inst, spec["K"], mode, NO, spec["seed"] if seed is None else seed, trace=trace)
    out = {"makespan": r["makespan"], "found": r["found"], "expanded": r["expanded"],
           "sequence_sha": sha(r["sequence"]) if r["found"] else None}
    if trace:
        out["pcount_sha"] = sha(r["pcount"])
        out["trail_sha"] = trail_digest(r["trail"], r["pcount"], spec["K"])
    return out, r


def main():
    t0 = time.time()
    g = {"about": "oracle (oracle/oracle.cpp orc_gpu_solve) results at the benchmark shapes; "
                  "written by tests/golden/make_oracle_goldens.py"}
    h = dict(W.HEADLINE)
    g["headline"] = {"spec": h}
    for name, mode in (("ICH_DET", ICH_DET), ("PMS_GPU", PMS_GPU)):
        out, r = solve(h, mode, trace=True)
        out["sequence"] = [int(v) for v in r["sequence"]]
        g["headline"][name] = out
        print(f"headline {name}: {out['makespan']} ({time.time() - t0:.0f} s)", flush=True)
    ms = dict(W.MULTISEED)
    beams = []
    for s in ms["seeds"]:
        out, _ = solve(dict(ms, seed=s), PMS_GPU)
        beams.append(dict(out, seed=s))
    g["multiseed"] = {"spec": ms, "beams": beams}
    print(f"multiseed ({time.time() - t0:.0f} s)", flush=True)
    specs = W.batch_specs()
    g["batch"] = [dict(solve(specs[i])[0], index=i, spec=specs[i]) for i in W.BATCH_GOLDEN_MEMBERS]
    print(f"batch ({time.time() - t0:.0f} s)", flush=True)
    sspecs, _ = W.serve_specs()
    g["serving"] = [dict(solve(sspecs[i])[0], index=i, spec=sspecs[i]) for i in W.serve_golden_members(sspecs)]
    print(f"serving ({time.time() - t0:.0f} s)", flush=True)
    with open(os.path.join(HERE, "oracle_full_size.json"), "w") as f:
        json.dump(g, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
