"""Loading of the golden fixtures (tests/golden/*.npz) into CcdProblem objects."""
import glob
import os

import numpy as np

from ccd_amd.data import CcdProblem

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    """Problem fixtures (every npz holding a data set; cv160_oracle.npz holds fits only)."""
    out = []
    for p in glob.glob(os.path.join(GOLDEN, "*.npz")):
        with np.load(p, allow_pickle=False) as z:
            if "model" in z.files:
                out.append(os.path.basename(p)[:-4])
    return sorted(out)


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    p = CcdProblem(int(z["model"]), int(z["K"]), int(z["J"]), z["col_ptr"], z["rows"], z["vals"],
                   z["fmt"], z["y"], z["offs"], z["strata"], has_offset=bool(z["has_offset"]))
    extra = {k: z[k] for k in z.files if k not in (
        "model", "K", "J", "col_ptr", "rows", "vals", "fmt", "y", "offs", "strata", "has_offset")}
    return p.validate(), extra
