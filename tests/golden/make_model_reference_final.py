"""Writes tests/golden/model_reference_final.json: the last row of each trace
fixture that the reference's tests/model_reference_test.rs compares against
(tests/fixtures/<name>_ref.npy, TVB traces of two uncoupled nodes, layout
[steps][nvar * nnodes] variable-major). Only the data is kept: the final row,
the fixture's shape and its file name. Run from the repository root, in a
container that holds the reference checkout:

    python tests/golden/make_model_reference_final.py /root/reference
"""
import json
import os
import sys

import numpy as np

FIXTURES = ["epileptor_codim3_ref.npy", "epileptor_codim3_slowmod_ref.npy", "epileptor_rs_ref.npy",
            "zerlaut_first_ref.npy", "zerlaut_second_ref.npy", "kionex_ref.npy"]


def main(ref_root: str) -> None:
    out = {}
    for name in FIXTURES:
        a = np.load(os.path.join(ref_root, "tests", "fixtures", name), allow_pickle=False)
        out[name] = {"shape": list(a.shape), "final_row": [float(v) for v in a[-1]]}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "model_reference_final.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
