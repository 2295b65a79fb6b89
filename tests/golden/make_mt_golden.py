"""Write tests/golden/mt_draws.json from the reference's own RNG source
(/root/reference/mersenne_inline.c, compiled unmodified into oracle/_ref by
oracle/Makefile): 64 draws after the reference's warm-up of 543210 draws
(hmc.c:877) for the seeds used by the tests and the benchmark."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

WARMUP = 543210
SEEDS = [1001, 2001, 54357345, 4354365264]


def main():
    oracle.build()
    out = {"warmup": WARMUP, "source": "reference mersenne_inline.c via oracle/_ref",
           "draws": {}}
    for s in SEEDS:
        out["draws"][str(s)] = [float(v) for v in oracle.ref_mt_draws(s, 64, skip=WARMUP)]
    with open(os.path.join(HERE, "mt_draws.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
