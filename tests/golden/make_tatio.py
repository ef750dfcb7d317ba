"""Writes the TAT binary-format fixtures (tatio_*.bin, tatio_golden.json)
with the reference TAT driver (oracle/_ref/tat_interp, built from the
reference sources by oracle/ref/Makefile): each tensor is made by TAT,
filled by TAT's range_, written by TAT's ``dump`` and printed back.  The
shapes follow TAT/tests/test_binary_io.cpp (no symmetry 2x3x4, U(1) and
fermionic U(1) rank-3 tensors) plus Z(2) ones.

Run from the repository root: python tests/golden/make_tatio.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ref_runner, tat  # noqa: E402

CASES = {
    "No": [("Left", [(0, 2)], False), ("Right", [(0, 3)], False), ("Up", [(0, 4)], False)],
    "BoseU1": [("Left", [(-1, 3), (0, 1), (1, 2)], False), ("Right", [(-1, 1), (0, 2), (1, 3)], False),
               ("Up", [(-1, 2), (0, 3), (1, 1)], False)],
    "FermiU1": [("Left", [(-2, 3), (0, 1), (-1, 2)], True), ("Right", [(0, 2), (1, 3)], False),
                ("Up", [(0, 3), (1, 1)], False)],
    "BoseZ2": [("A", [(0, 2), (1, 3)], False), ("B", [(0, 1), (1, 2)], False), ("C", [(1, 2), (0, 2)], False)],
    "FermiZ2": [("A", [(0, 2), (1, 3)], True), ("B", [(1, 1), (0, 2)], False), ("C", [(0, 3), (1, 2)], True)],
}


def main():
    if not ref_runner.available():
        ref_runner.build()
    golden = {}
    for kind_name, spec in CASES.items():
        kind = tat.KINDS[kind_name]
        path = os.path.join(HERE, f"tatio_{kind_name}.bin")
        lines = [f"tensor t {len(spec)}"]
        for name, segs, arrow in spec:
            lines.append(f"edge {name} {len(segs)} {int(arrow)} " + " ".join(f"{q} {d}" for q, d in segs))
        lines += ["range t 2 1", f"dump t {path}", "print t"]
        t = ref_runner.run(kind, lines)["t"]
        golden[kind_name] = {"file": os.path.basename(path), "names": t.names,
                             "edges": [{"segments": [[int(q[0]) if kind.length else 0, int(d)] for q, d in e.segments],
                                        "arrow": bool(e.arrow)} for e in t.edges],
                             "storage": [float(v) for v in t.data]}
    with open(os.path.join(HERE, "tatio_golden.json"), "w") as f:
        json.dump(golden, f, indent=1)


if __name__ == "__main__":
    main()
