"""Extracts the hard-coded golden vectors of the reference's own unit tests
into tests/golden/reference_vectors.json (data only: the integer/real
tables, not test code).

Run in a container where the reference checkout is present:
    python tests/golden/extract_reference_vectors.py /root/reference

Vectors taken (uofuseismo/eikonalxx):
  testing/graph.cpp:46-162     Graph3D node lists of all 8 sweeps on a
                               4 x 3 x 5 grid (nodeInLevelToGridPointRef1..8)
                               and levelStartPtrRef.
  testing/solver2d.cpp:573-600 2D level-set travel-time index tables on a
                               5 x 3 grid (tRef0..tRef3).
  testing/solver2d.cpp:724-759 2D sweep slowness tables on a 5 x 4 grid
                               (sweepSlowRef1..4).
"""
import json
import os
import re
import sys

NAMES = {
    "graph.cpp": ["nodeInLevelToGridPointRef%d" % i for i in range(1, 9)]
    + ["levelStartPtrRef"],
    "solver2d.cpp": ["tRef0", "tRef1", "tRef2", "tRef3", "tTimes", "slow"]
    + ["sweepSlowRef%d" % i for i in range(1, 5)],
}


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def extract(reference):
    out = {}
    for fname, names in NAMES.items():
        path = os.path.join(reference, "testing", fname)
        text = _strip_comments(open(path).read())
        for name in names:
            m = re.search(r"std::vector<(int|double)>\s+%s\s*\{([^}]*)\}"
                          % re.escape(name), text)
            if m is None:
                raise RuntimeError("%s not found in %s" % (name, fname))
            kind = m.group(1)
            vals = [v.strip() for v in m.group(2).split(",") if v.strip()]
            conv = int if kind == "int" else float
            out[name] = [conv(v) for v in vals]
    return out


def main():
    reference = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = extract(reference)
    here = os.path.dirname(os.path.abspath(__file__))
    dst = os.path.join(here, "reference_vectors.json")
    with open(dst, "w") as f:
        json.dump(data, f, separators=(",", ":"), sort_keys=True)
    print("wrote", dst, {k: len(v) for k, v in data.items()})


if __name__ == "__main__":
    main()
