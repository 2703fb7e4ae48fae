"""Loads the golden fixtures written by tests/golden/make_fixtures.py (the reference's
own outputs) and maps their namelist values onto the oracle's and the library's
parameters."""
import glob
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names(kind="run"):
    """Fixture names of one kind: "run" (reference time-loop runs) or "statistics"."""
    out = []
    for p in sorted(glob.glob(os.path.join(GOLDEN, "*.npz"))):
        z = np.load(p, allow_pickle=False)
        k = json.loads(str(z["params_json"])).get("kind", "run")
        if k == kind:
            out.append(os.path.splitext(os.path.basename(p))[0])
    return out


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
