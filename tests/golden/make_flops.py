"""Writes tests/golden/flops_per_de.json: mean ND-FLOP v1 (and transcendental ops) per
scene_de evaluation over each workload's actual march and shadow evaluation points,
measured by the oracle's counting scalar on every 16th pixel in x and y.

ND-FLOP v1: add/sub/mul/min/max/div/sqrt = 1, a*b+c = 2, compare/select/abs/neg = 0,
transcendentals counted apart; uniform-only subexpressions are hoisted.

Run: python tests/golden/make_flops.py   (CPU only; takes a few seconds)
"""
import ctypes as C
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numericdream_amd as nd  # noqa: E402
import oracle  # noqa: E402


def main():
    L = oracle.lib()
    L.ndo_flops_render.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_void_p]
    res = {"convention": "ND-FLOP v1", "sample_step_px": 16, "workloads": {}}
    for cfg, jobs, shadow_steps in ((1, [0], 0), (2, [0, 1, 2, 3], 128), (3, [0], 40), (4, list(range(8)), 0),
                                    (5, [0, 1], 0)):
        acc = np.zeros(6)
        for j in jobs:
            u = nd.config_uniforms(cfg, j)
            W, H = int(u.resolution[0]), int(u.resolution[1])
            out = np.zeros(6)
            L.ndo_flops_render(C.addressof(u), None, shadow_steps, 0, 0, W, H, 16, out.ctypes.data)
            for w in (0, 1):
                n = out[3 * w]
                acc[3 * w] += n
                acc[3 * w + 1] += out[3 * w + 1] * n
                acc[3 * w + 2] += out[3 * w + 2] * n
        entry = {}
        for w, name in ((0, "march"), (1, "shadow")):
            n = acc[3 * w]
            entry[name] = {"evals_sampled": int(n), "flops_per_de": (acc[3 * w + 1] / n) if n else 0.0,
                           "tops_per_de": (acc[3 * w + 2] / n) if n else 0.0}
        res["workloads"][str(cfg)] = entry
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "flops_per_de.json")
    with open(path, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
    print(json.dumps(res, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
