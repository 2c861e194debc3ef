"""Write tests/golden/sfc64_counter_vectors.json from numpy's SFC64.

numpy's SFC64 is an implementation independent of this repository and of the
reference; the reference's own KAT (hoomd-rand/src/sfc.rs:128-168) was produced
with it.  Each vector sets the raw generator state to the Counter layout of
hoomd-rand/src/counter.rs:160-167, i.e. (step, substep << 32 | seed, index, 0),
discards the 12 warm-up outputs of sfc.rs:62-71 and records the next 32 raw
64-bit outputs.
"""
import json
import pathlib

import numpy as np

OUT = pathlib.Path(__file__).with_name("sfc64_counter_vectors.json")


def stream(a, b, c, ctr, n, warm=12):
    g = np.random.SFC64()
    g.state = {"bit_generator": "SFC64",
               "state": {"state": np.array([a, b, c, ctr], dtype=np.uint64)},
               "has_uint32": 0, "uinteger": 0}
    raw = g.random_raw(warm + n)
    return [int(x) for x in raw[warm:]]


def main():
    rng = np.random.default_rng(20240601)
    cases = []
    fixed = [(0, 0, 0, 0), (1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 0, 0, 1),
             (0, 0, 0, 2**64 - 1), (2**64 - 1, 2**32 - 1, 2**32 - 1, 2**64 - 1),
             (42, 0, 42, 0), (42, 3, 42, 123456), (1000, 7, 42, 2**64 - 1)]
    for step, substep, seed, index in fixed:
        cases.append((step, substep, seed, index))
    while len(cases) < 64:
        step = int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2))
        substep = int(rng.integers(0, 2**32))
        seed = int(rng.integers(0, 2**32))
        index = int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2))
        cases.append((step, substep, seed, index))
    vectors = []
    for step, substep, seed, index in cases:
        out = stream(step, (substep << 32) | seed, index, 0, 32)
        vectors.append({"step": step, "substep": substep, "seed": seed, "index": index,
                        "draws": [str(x) for x in out]})
    # seed_from_u64(0) = initialize(0, 0, 0, 1): deep value after 2^17 outputs
    deep = stream(0, 0, 0, 1, 2**17)[-1]
    OUT.write_text(json.dumps({"generator": "numpy.random.SFC64 " + np.__version__,
                               "vectors": vectors, "deep_seed0_2pow17": str(deep)}, indent=0))


if __name__ == "__main__":
    main()
