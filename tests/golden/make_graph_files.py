"""Writes the expected CSR of every graph-file fixture in tests/golden/graphs/
as the reference reads it: GAPBS's reader and builder, compiled from the
reference sources into oracle/_ref/ (``make -C oracle ref``), called the way
load_graph calls them (runtime/infra_gpu/graph.h:273-276). Also writes two
serialized fixtures (.sg, .wsg) with GAPBS's own writer. Run from the
repository root: ``python tests/golden/make_graph_files.py``.
"""
import glob
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

HERE = os.path.join(ROOT, "tests", "golden", "graphs")


def main():
    if O.ref() is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    O.ref_write_serialized(os.path.join(HERE, "dup_loops.el"), False, False,
                           os.path.join(HERE, "dup_loops.sg"))
    O.ref_write_serialized(os.path.join(HERE, "weighted.wel"), True, True,
                           os.path.join(HERE, "weighted_sym.wsg"))
    out = {}
    for path in sorted(glob.glob(os.path.join(HERE, "*"))):
        name = os.path.basename(path)
        if name.endswith(".npz"):
            continue
        serialized = name.endswith((".sg", ".wsg"))
        for sym in ((False,) if serialized else (False, True)):
            off, nb, w, directed = O.ref_read_graph(path, sym)
            key = f"{name}|{int(sym)}"
            out[key + "|offsets"] = off
            out[key + "|nbrs"] = nb
            out[key + "|weights"] = w if w is not None else np.zeros(0, np.int32)
            out[key + "|meta"] = np.array([int(directed), int(w is not None)], np.int32)
    np.savez(os.path.join(HERE, "expected.npz"), **out)
    print(f"{len(out) // 4} expected graphs written")


if __name__ == "__main__":
    main()
