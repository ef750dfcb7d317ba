"""Driver for ``oracle/_ref/tat_interp`` (TEST INFRASTRUCTURE ONLY).

Serialises oracle tensors into the interpreter's script grammar
(``oracle/ref/tat_interp.cpp``), runs the binary built from the reference TAT
headers, and parses ``print`` output back into oracle tensors.  Used to pin
the oracle against real TAT and to write golden fixtures.
"""
import os
import subprocess

import numpy as np

from . import tat

HERE = os.path.dirname(os.path.abspath(__file__))
BINARY = os.path.join(HERE, "_ref", "tat_interp")


def available():
    return os.path.exists(BINARY) and os.access(BINARY, os.X_OK)


def build(reference="/root/reference"):
    """Compile the interpreter when the reference checkout is present."""
    if not os.path.isdir(os.path.join(reference, "TAT", "include")):
        return False
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "ref"), f"REF={reference}"], check=True)
    return available()


def _q(kind, q):
    return int(q[0]) if kind.length else 0


def tensor_lines(tid, t):
    out = [f"tensor {tid} {t.rank}"]
    for n, e in zip(t.names, t.edges):
        segs = " ".join(f"{_q(t.kind, q)} {d}" for q, d in e.segments)
        out.append(f"edge {n} {len(e.segments)} {int(e.arrow)} {segs}")
    vals = " ".join(repr(float(v)) for v in t.data)
    out.append(f"data {tid} {t.size} {vals}")
    return out


def run(kind, lines):
    """Run a script; returns {id: oracle Tensor} for every ``print``."""
    script = "\n".join([f"sym {kind.name}"] + lines) + "\n"
    res = subprocess.run([BINARY], input=script, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"tat_interp failed: {res.stderr.strip()}")
    return parse(kind, res.stdout)


def parse(kind, text):
    tensors = {}
    toks = text.split("\n")
    i = 0
    while i < len(toks):
        line = toks[i].split()
        if not line:
            i += 1
            continue
        if line[0] == "tensor":
            tid, rank = line[1], int(line[2])
            names, edges = [], []
            for r in range(rank):
                el = toks[i + 1 + r].split()
                name, nseg, arrow = el[1], int(el[2]), bool(int(el[3]))
                segs = [(int(el[4 + 2 * s]) if kind.length else None, int(el[5 + 2 * s])) for s in range(nseg)]
                names.append(name)
                edges.append(tat.Edge(kind, segs, arrow))
            dl = toks[i + 1 + rank].split()
            n = int(dl[1])
            vals = np.array([float(v) for v in dl[2:2 + n]])
            t = tat.Tensor(kind, names, edges)
            t.data[:] = vals
            tensors[tid] = t
            i += rank + 2
        elif line[0] == "time":
            tensors["__time__"] = (float(line[1]), float(line[2]))
            i += 1
        else:
            i += 1
    return tensors
