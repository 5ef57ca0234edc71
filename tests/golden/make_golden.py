"""Writes reference_outputs.npz: every case of cases.py run through the
reference's own CPU library (CPU-4State-Double / CPU-Double, built from
/root/reference sources by oracle/ref.mk into oracle/_ref), with the
problem inputs stored beside the outputs.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "beagle-lib_amd"))
sys.path.insert(0, HERE)

from beagle_amd import api, synth  # noqa: E402
import cases  # noqa: E402


def main():
    ref = api.Library(os.path.join(REPO, "oracle", "_ref", "libhmsbeagle.so"), "beagle")
    arrays = {}
    for name, (_, opts, parts) in cases.CASES.items():
        prob = cases.problem(name)
        out = cases.run(ref, prob, opts, parts, extra_flags=api.FLAG_PROCESSOR_CPU)
        for k, v in synth.problem_arrays(prob, name + "/in/").items():
            arrays[k] = v
        for k, v in out.items():
            arrays[name + "/out/" + k] = np.asarray(v)
        print(name, "logL", out["logL"][0])
    np.savez_compressed(os.path.join(HERE, "reference_outputs.npz"), **arrays)


if __name__ == "__main__":
    main()
