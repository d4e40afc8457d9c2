"""Writes tests/golden/3elt.graph from the reference's only in-repo graph,
mtx/3elt.mtx (Matrix Market, pattern symmetric), with the conversion
semantics of the reference's mtx_to_graph.c (A union A^T, diagonal dropped,
rows ascending, 1-based ids, unweighted METIS file), and 3elt_k8.npz with the
oracle's level-0 result on it (checked against closed forms by the tests).

Run in the build container only: python tests/golden/make_3elt.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from cumetis_amd import graphs  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main(src="/root/reference/mtx/3elt.mtx"):
    rows, cols = [], []
    n = None
    with open(src) as f:
        for ln in f:
            if ln.startswith("%"):
                continue
            t = ln.split()
            if n is None:
                n = int(t[0])
                continue
            rows.append(int(t[0]) - 1)
            cols.append(int(t[1]) - 1)
    g = graphs.from_edges(n, np.array(rows), np.array(cols))
    graphs.write_metis(str(HERE / "3elt.graph"), g)
    x = g.xadj.astype(np.int64)
    w = np.ones(g.m, np.int32)
    T = O.threshold(n, 8)
    mv = O.maxvwgt(n, T)
    O.srand(1)
    match, cmap, nc, perm = O.cpu_match(x, g.adjncy, w, None, mv)
    cx, ca, cw, cv = O.cpu_contract(x, g.adjncy, w, None, match, cmap, nc)
    np.savez_compressed(HERE / "3elt_k8.npz", match=match, cmap=cmap, perm=perm, cxadj=cx,
                        cadjncy=ca, cadjwgt=cw, cvwgt=cv, T=T, maxvwgt=mv)
    print(f"3elt: n={n} edges={g.m // 2} T={T} maxvwgt={mv} nc={nc} mc={cx[-1]}")


if __name__ == "__main__":
    main(*sys.argv[1:])
