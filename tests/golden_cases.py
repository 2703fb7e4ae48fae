"""Loads the golden fixtures written by tests/golden/make_fixtures.py (the reference's
own outputs) and maps their namelist values onto the oracle's and the library's
parameters."""
import glob
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    return sorted(os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["params"] = json.loads(str(d.pop("params_json")))
    return d


def oracle_params(pyoracle, p):
    op = pyoracle.OracleParams()
    for k, v in p.items():
        if hasattr(op, k):
            setattr(op, k, type(getattr(op, k))(v))
    return op
