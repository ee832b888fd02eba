"""Writes c4_problems.json: the batched workload's 512 problem shapes, fixed by seed.

R log-uniform in [1K, 256K], Q log-uniform in [16, 4096], dim uniform over {32, 64, 128, 256, 512},
k uniform over {1, 8, 16, 64, 128}; each problem's data is uniform [0, 1) from the library's seeded
generator with the seeds listed here.
"""
import json
import os

import numpy as np

SEED = 0x6B4E4E04


def make(n=512, seed=SEED):
    rng = np.random.default_rng(seed)
    out = []
    for j in range(n):
        R = int(round(np.exp(rng.uniform(np.log(1024), np.log(262144)))))
        Q = int(round(np.exp(rng.uniform(np.log(16), np.log(4096)))))
        D = int(rng.choice([32, 64, 128, 256, 512]))
        k = int(rng.choice([1, 8, 16, 64, 128]))
        out.append({"ref_nb": R, "query_nb": Q, "dim": D, "k": k,
                    "ref_seed": seed * 1000 + 2 * j, "query_seed": seed * 1000 + 2 * j + 1})
    return out


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c4_problems.json")
    with open(path, "w") as f:
        json.dump({"seed": SEED, "problems": make()}, f, indent=0)
    print(path)
