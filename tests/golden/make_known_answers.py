"""Writes known_answers.json: inputs and published answers of the
reference's example programs, as data.

  tinytest  examples/tinytest/tinytest.cpp:19-21 (sequences), :209 (rates),
            :234-248 (JC69 eigen-system), :261 (edge lengths), :288-291
            (operations), :332 (PAUP logL -1498.89812)
  hmctest5  examples/hmctest/hmctest5.cpp:25-27 (sequences), :275-276
            (rates), :289 (frequencies), :374-385 (Q), :424 (edge
            lengths), :491-494 and :524-529 (operations), :589 (R logL
            -18.04619478977292), :675-679 (R per-site dlogL/dt table)
  hmcGapTest examples/hmctest/hmcGapTest.cpp:25-27 (sequences with gaps),
            :368 (Q), :406 (edge lengths), same model and tree as
            hmctest5, :564 (R logL -16.18744865764452)

The R values were computed (examples/hmctest/hmctest.R:95-105,
hmcGaptest.R:97-107) with an F81 eigen-system written to 7 digits
(eigenvalues -1.428571) and Q = evec diag(eval) ivec, so both fixtures
carry that eigen-system.

The sequences of tinytest are read from the reference file (they are
long); the rest is transcribed.

    python tests/golden/make_known_answers.py
"""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/examples/tinytest/tinytest.cpp"


def tinytest_sequences():
    src = open(REF).read().splitlines()
    seqs = {}
    for line in src[18:21]:   # lines 19-21
        m = re.match(r'char \*(\w+) = \(char\*\)"([^"]*)";', line.strip())
        seqs[m.group(1)] = m.group(2)
    return seqs


R_EIGEN = {
    "evec": [[0.9819805, 0.040022305, 0.04454354, -0.5],
             [-0.1091089, -0.002488732, 0.81606029, -0.5],
             [-0.1091089, -0.896939683, -0.11849713, -0.5],
             [-0.1091089, 0.440330814, -0.56393254, -0.5]],
    "ivec": [[0.9165151, -0.3533241, -0.1573578, -0.4058332],
             [0.0, 0.2702596, -0.8372848, 0.5670252],
             [0.0, 0.8113638, -0.2686725, -0.5426913],
             [-0.2, -0.6, -0.4, -0.8]],
    "evals": [-1.428571, -1.428571, -1.428571, 0.0],
}


def main():
    s = tinytest_sequences()
    data = {
        "tinytest": {
            "sequences": [s["human"], s["chimp"], s["gorilla"]],
            "tips_as": "partials",
            "rates": [0.03338775, 0.25191592, 0.82026848, 2.89442785],
            "model": "JC69",
            "tree": {"children": [[0, 1], [2, 3]], "lengths": [0.1, 0.1, 0.2, 0.1]},
            "logL": -1498.89812, "logL_digits": 5, "source": "PAUP"},
        "hmctest5": {
            "sequences": ["GAGT", "GAGG", "AAAT"],
            "tips_as": "states",
            "rates": [0.14251623900062188, 1.857483760999378],
            "model": "eigen", "eigen": R_EIGEN, "freqs": [0.1, 0.3, 0.2, 0.4],
            "tree": {"children": [[0, 1], [2, 3]], "lengths": [0.6, 0.6, 1.3, 0.7]},
            "logL": -18.04619478977292, "source": "R",
            "gradient_nodes": [1, 0, 2, 3],
            "site_dlogL_dt": [[-0.248521, -0.194621, -0.248521, 0.36811],
                              [-0.248521, -0.194621, -0.248521, 0.114741],
                              [0.221279, -0.171686, 0.221279, -0.00658093],
                              [0.22128, -0.171686, 0.22128, -0.00658095]],
            "gradient_digits": 6},
        "hmcGapTest": {
            "sequences": ["-AGTC", "GAGGC", "AA-TC"],
            "tips_as": "states",
            "rates": [0.14251623900062188, 1.857483760999378],
            "model": "eigen", "eigen": R_EIGEN, "freqs": [0.1, 0.3, 0.2, 0.4],
            "tree": {"children": [[0, 1], [2, 3]], "lengths": [0.6, 0.6, 1.3, 0.7]},
            "logL": -16.18744865764452, "source": "R"},
    }
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
