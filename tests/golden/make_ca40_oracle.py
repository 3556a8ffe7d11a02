"""Runs the FP64 CPU oracle on the Ca-40 40^3 HF+BCS workload to convergence
(|dE| < 1e-6 MeV) from the seeded HO guess and writes the converged
energies, single-particle energies and observables as a JSON fixture.

    python tests/golden/make_ca40_oracle.py [out.json]

Takes ~10-30 min on 8 cores.  The guess is built with the library's HO
generator (pinned against the oracle's own in the CPU tests)."""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import _oracle as O  # noqa: E402
from nuclearcgm_amd import config as C  # noqa: E402

import bench  # noqa: E402


def main(out):
    cfg = C.workload("ca40")
    cfg.max_iter = 600
    guess = bench.make_guess(cfg)
    t = time.time()
    pn, pp, en, ep, res, trace = O.scf_run(cfg, *guess)
    s = res.summary()
    s.update({"eps_n": list(map(float, en)), "eps_p": list(map(float, ep)), "wall_s": time.time() - t,
              "trace_E_int": [float(x) for x in trace[:, 0]], "workload": "ca40", "seed_guess": 7,
              "V0": cfg.pair_n.V0, "threads": os.cpu_count()})
    with open(out, "w") as fh:
        json.dump(s, fh, indent=1)
    print(s["E_int"], s["iterations"], s["wall_s"])


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "ca40_oracle.json"))
