"""Writes rt16_oracle_3steps.json: the oracle's result for 3 steps of the
three-level 16^3 Rayleigh-Taylor hierarchy, as hashes.

The problem has the shape of the 3D Rayleigh-Taylor workload at a twelfth
of its size: 360 leaves of 16^3 (8 / 96 / 256 on levels 1-3 over a
2 x 2 x 6 level-1 grid, z in [0, 3]; level 2 where z meets (0.5, 2.5),
level 3 where it meets (1.25, 1.75), so two level jumps with every guard
layer over a leaf at most one level finer), 8 mass scalars, gravity g = -1
along z, periodic x / y, reflecting z, 10 guard cells, fine/coarse flux
correction every step.  The file holds the SHA-256 of the generated input
state (so a run can tell that it starts from the same bytes), the dt of
each step as hex floats, and per leaf the SHA-256 of its interior after the
3 steps (all unk variables, C order).  A device run is bit-exact with the oracle exactly
when its dt sequence and every leaf hash agree.

Run from the repository root:  python tests/golden/make_rt16_oracle.py
"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
for p in (ROOT / "flash-x_amd", ROOT / "oracle"):
    sys.path.insert(0, str(p))

from sparkhip import problems  # noqa: E402

OUT = Path(__file__).resolve().parent / "rt16_oracle_3steps.json"
KW = dict(nb0=(2, 2, 6), n=16, lmax=3, slabs=((0.5, 2.5), (1.25, 1.75)), nscalars=8)
NSTEPS = 3


def problem():
    return problems.rt_amr(**KW)


def input_hash(prob):
    h = hashlib.sha256()
    for a in prob.arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def interior_hashes(prob, arrays):
    g = prob.cfg["nguard"]
    return [hashlib.sha256(np.ascontiguousarray(a[g:-g, g:-g, g:-g]).tobytes()).hexdigest() for a in arrays]


def main():
    import pyoracle
    prob = problem()
    rec = {"problem": "problems.rt_amr(" + ", ".join(f"{k}={v!r}" for k, v in KW.items()) + ")",
           "leaves_per_level": [int(x) for x in prob.info["leaves_per_level"]],
           "steps": NSTEPS, "input_sha256": input_hash(prob)}
    ref = [a.copy() for a in prob.arrays]
    dts, nerr = pyoracle.amr_advance(prob.cfg, prob.blocks, prob.info["nb0"], prob.info["origins"], ref, NSTEPS,
                                     flux_correct=True, bc=prob.info["bc"])
    assert nerr == 0
    rec["dts"] = [float(d).hex() for d in dts]
    rec["interior_sha256"] = interior_hashes(prob, ref)
    OUT.write_text(json.dumps(rec, indent=1) + "\n")
    print(f"wrote {OUT}: dts {[float(d) for d in dts]}")


if __name__ == "__main__":
    main()
