"""Writes tests/golden/model_fixture_traces.npz: the TVB traces the reference
keeps for its single-node model fixtures (tests/fixtures/<name>_ref.npy,
[steps][nvar] of one node, float32), for the models this library runs on the
multi-subnetwork path whose traces the reference's own tests do not consume.
Only the data is kept (one array per fixture, named after it). Run from the
repository root, in a container that holds the reference checkout:

    python tests/golden/make_model_fixture_traces.py /root/reference
"""
import os
import sys

import numpy as np

FIXTURES = ["coombes_byrne_ref.npy", "gast_sd_ref.npy", "gast_sf_ref.npy", "larter_breakspear_ref.npy",
            "epileptor2d_ref.npy", "epileptor_ref.npy", "rww_exc_inh_ref.npy", "deco_balanced_ref.npy",
            "zetterberg_jansen_ref.npy", "reduced_fhn_ref.npy", "reduced_hr_ref.npy", "dumont_gutkin_ref.npy",
            "sup_hopf_ref.npy", "hopfield_ref.npy", "coombes_byrne2d_ref.npy", "linear_ref.npy"]


def main(ref_root: str) -> None:
    out = {}
    for name in FIXTURES:
        a = np.load(os.path.join(ref_root, "tests", "fixtures", name), allow_pickle=False)
        out[name[:-len(".npy")]] = np.ascontiguousarray(a, np.float32)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "model_fixture_traces.npz")
    np.savez_compressed(path, **out)
    print(path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
