"""Writes tests/golden/batched_pairs.npz: pairs taken from the batched clip
workload (gen_pair_layers(200000, seed=0): log-normal vertex counts, median
64, sigma 1.5, 8..50000, 70 % overlapping) that exercise rare paths:

  p4596   the trace of the GPU route does not close (status -1 in the oracle)
  p156581 two BOUNCING intersection pairs, both split in step 5, among 4518
          crossings
  p73779  one BOUNCING intersection pair (not split) among 3670 crossings
  p176719 two split BOUNCING pairs among 2858 crossings

Each pair is stored as <name>_P / <name>_Q (float64 [n, 2]).  Generating the
whole workload takes about a minute and a half on one core.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "gcmf-gh_amd"))

import gcmf_amd as G  # noqa: E402


def main():
    P, Q = G.gen_pair_layers(200000, seed=0)
    out = {}
    for k in (4596, 156581, 73779, 176719):
        out["p%d_P" % k] = P.ring(k)
        out["p%d_Q" % k] = Q.ring(k)
    np.savez_compressed(os.path.join(HERE, "batched_pairs.npz"), **out)


if __name__ == "__main__":
    main()
