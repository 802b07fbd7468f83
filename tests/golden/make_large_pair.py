"""Writes large_pair_200k.json: the oracle's CPU-route clip of the
200000 x 20000 pair of the large-pair shape (gen_single_pair kind 1, seed 0)
in the edge-scaled frame (oracle mode 5), as component sizes and the SHA-256
of the result vertices (float64 x, y in order), plus its traced area and the
boundary-integral area.  Takes about two and a half minutes."""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, "..", "..", "gcmf-gh_amd"))

import numpy as np  # noqa: E402

import gcmf_amd as G  # noqa: E402
import oracle_py as O  # noqa: E402


def main():
    P, Q = G.gen_single_pair(1, 0, 200000, 20000)
    status, comps, extra = O.clip_cpu(P, Q, 5)
    xy = np.concatenate(comps) if comps else np.zeros((0, 2))
    allxy = np.concatenate([P, Q])
    ox = 0.5 * (allxy[:, 0].min() + allxy[:, 0].max())
    oy = 0.5 * (allxy[:, 1].min() + allxy[:, 1].max())
    out = {"pair": "gen_single_pair(1, 0, 200000, 20000), oracle mode 5 (normalised, "
                   "edge-scaled, clean-up on)",
           "status": int(status), "comp_sizes": [len(c) for c in comps],
           "xy_sha256": hashlib.sha256(np.ascontiguousarray(xy, np.float64).tobytes()).hexdigest(),
           "n_inter": int(extra["n_inter"]),
           "area_result": O.comps_area(comps, ox, oy),
           "area_boundary": O.area_boundary(P, Q)}
    with open(os.path.join(HERE, "large_pair_200k.json"), "w") as f:
        json.dump(out, f)
    print(out["status"], len(out["comp_sizes"]), sum(out["comp_sizes"]), out["xy_sha256"])


if __name__ == "__main__":
    main()
