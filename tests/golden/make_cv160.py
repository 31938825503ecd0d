"""Writes tests/golden/cv160_oracle.npz: the CPU oracle's fits 0, 87 and 159 of the
batched cross-validation benchmark (conditional logistic, 200,000 strata, 20,000
Zipf covariates, 10 folds x 16 Laplace variances 1e-4..1, seed 123), with their
held-out scores, and a fingerprint of the data set they were fitted on.  The
weakest-prior fit takes several minutes on one core, too long to recompute inside
a GPU test.

    python tests/golden/make_cv160.py
"""
import os
import sys
import zlib
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

FITS = (0, 87, 159)
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cv160_oracle.npz")


def fingerprint(p):
    """CRC-32 of the arrays that define the data set."""
    crc = 0
    for a in (p.col_ptr, p.rows, p.y, p.strata):
        crc = zlib.crc32(np.ascontiguousarray(a).tobytes(), crc)
    return np.array([p.K, p.J, p.nnz, crc], dtype=np.int64)


def main():
    from ccd_amd import synth
    from oracle import oracle as orc
    from tests.test_gpu_fullsize import headline_cv_specs
    p, _ = synth.clr_strata()
    specs, held = headline_cv_specs(p)
    with ThreadPoolExecutor(len(FITS)) as ex:
        refs = list(ex.map(lambda i: orc.fit(p, specs[i], heldout=held[i]), FITS))
    np.savez_compressed(
        OUT, fits=np.array(FITS), fingerprint=fingerprint(p),
        beta=np.array([b for b, _ in refs]),
        log_likelihood=np.array([i["log_likelihood"] for _, i in refs]),
        log_prior=np.array([i["log_prior"] for _, i in refs]),
        pred_log_likelihood=np.array([i["pred_log_likelihood"] for _, i in refs]),
        iterations=np.array([i["iterations"] for _, i in refs]),
        return_flag=np.array([i["return_flag"] for _, i in refs]))
    for f, (b, i) in zip(FITS, refs):
        print(f, i["iterations"], i["return_flag"], i["log_likelihood"], int((b != 0).sum()))


if __name__ == "__main__":
    main()
