"""Writes batch256_oracle.json: the lowest eigenvalues of the 256 uneven sd-shell systems
of the batched workload (synth.random_requests(256, seed 17): Nv, Zv uniform in 1..6,
2Jz from the allowed range, an interaction of its own seed), computed from the oracle's H
(oracle.System.h_upper, pinned against the reference's own .HH rows and products) with
numpy's dense eigvalsh up to n = 1500 and scipy's Lanczos (eigsh, tol 1e-12) above.
Needs no reference tree; it only saves the GPU test the few CPU minutes this takes.
"""
import json
import os
import sys
import time

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import cosmo_amd  # noqa: E402  (System gives the dimension, as in the workload generator)
import oracle  # noqa: E402
from cosmo_amd import synth  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch256_oracle.json")


def lowest(xml, k):
    o = oracle.System(xml)
    rp, c, v = o.h_upper()
    n = len(rp) - 1
    if n <= 1500:
        return np.linalg.eigvalsh(oracle.dense_from_upper(n, rp, c, v))[:k]
    rows = np.repeat(np.arange(n), np.diff(rp))
    U = sp.csr_matrix((v, (rows, c)), shape=(n, n))
    H = (U + sp.triu(U, 1).T).tocsr()
    return np.sort(sla.eigsh(H, k=min(k + 2, n - 1), which="SA", tol=1e-12)[0])[:k]


def main():
    reqs, _ = synth.random_requests(cosmo_amd, 256, 17)
    out = []
    t0 = time.time()
    for i, (shell, nv, zv, twojz, xml, s) in enumerate(reqs):
        k = min(5, s.dim)
        out.append(dict(nv=nv, zv=zv, twojz=twojz, n=int(s.dim), evals=[float(x) for x in lowest(xml, k)]))
        if i % 32 == 31:
            print("%d systems, %.0f s" % (i + 1, time.time() - t0), flush=True)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
