"""Writes the golden fixtures under tests/golden/ from the reference TAT.

Runs only where the reference checkout and its interpreter build
(``oracle/_ref/tat_interp``, built by ``oracle/ref/Makefile`` from the
header-only TAT) are available.  Inputs come from seeded numpy generators;
outputs are TAT's own results on those inputs:

* ``c1_ops.json``  : config-1 tensors (rank-4, every edge dimension 8 split
  {-1:2, 0:4, +1:2}) for FermiU1 / BoseU1 / FermiZ2: contract over (k,l),
  all 24 transposes of A, SVD (i,j)|(m,n) cut 16 (singular values, cut
  segments, U S V), QR / LQ (Q R product), conjugate, merge and split.
* ``c2_amplitudes.json`` : psi(s) of the 4x4 D=4 Dc=16 FermiU1 PEPS and of a
  3x3 FermiZ2 PEPS for a set of configurations, computed by TAT in
  tetragono's contraction order.
* ``peps8x8_amplitudes.json`` : the same for the 8x8 FermiZ2 D=6 Dc=36 and
  the 8x8 FermiU1 D=8 Dc=64 PEPS (a few configurations each).
* ``zero_amplitudes.json`` : configurations of those two 8x8 PEPS whose
  amplitude TAT evaluates to exactly zero, with a few nonzero ones.
* ``c1_ops.json`` also holds relative cuts (0.3 and 0.6 of the largest
  singular value) of the same SVDs.

    python tests/golden/make_golden.py
"""
import hashlib
import itertools
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import helpers  # noqa: E402
from oracle import peps, ref_runner as rr, tat  # noqa: E402


def digest(data):
    return hashlib.sha256(np.ascontiguousarray(data, dtype=np.float64).tobytes()).hexdigest()


def tensor_json(t, arrays=None, key=None):
    """Structure + sha256 of the float64 storage; with ``arrays`` the storage
    itself also goes into the npz under ``key``."""
    out = {"names": t.names, "edges": [[[list(q), d] for q, d in e.segments] for e in t.edges],
           "arrows": [e.arrow for e in t.edges], "sha256": digest(t.data), "size": int(t.size)}
    if arrays is not None:
        arrays[key] = np.asarray(t.data, dtype=np.float64)
    return out


def c1_pair(kind_name, seed):
    return helpers.c1_pair(tat, kind_name, seed)


def make_c1():
    out = {}
    arrays = {}
    for kind_name, seed in (("FermiU1", 11), ("BoseU1", 12), ("FermiZ2", 13)):
        A, B = c1_pair(kind_name, seed)
        lines = rr.tensor_lines("A", A) + rr.tensor_lines("B", B)
        lines += ["contract C A B 2 k k l l", "print C"]
        perms = ["".join(p) for p in itertools.permutations("ijkl")]
        for p in perms:
            lines += [f"transpose T{p} A 4 {' '.join(p)}", f"print T{p}"]
        lines += ["svd U S V C 2 i j cu cv su sv 16", "print U", "print S", "print V",
                  "contract US U S 1 cu su", "contract USV US V 1 sv cv", "print USV",
                  "svd U2 S2 V2 C 1 j cu cv su sv 5", "print S2",
                  "svd U3 S3 V3 C 2 i j cu cv su sv 0.3", "print S3",
                  "contract US3 U3 S3 1 cu su", "contract USV3 US3 V3 1 sv cv", "print USV3",
                  "svd U4 S4 V4 C 1 m cu cv su sv 0.6", "print S4",
                  "qr Q R C r 2 m n cq cr", "contract QR Q R 1 cq cr", "print QR", "print Q", "print R",
                  "qr Q2 R2 C q 2 m n cq cr", "contract QR2 Q2 R2 1 cq cr", "print QR2",
                  "print U3", "print V3",
                  "conjugate CJ A", "print CJ",
                  "merge M A mm 2 j i 1", "print M",
                  "merge M0 A mm 2 k l 0", "print M0",
                  "reverse RV A 2 i k 1", "print RV"]
        res = rr.run(A.kind, lines)
        entry = {"seed": seed, "A": tensor_json(A), "B": tensor_json(B)}
        # floating-point results: full storage in the npz
        for key in ["C", "S", "USV", "S2", "S3", "USV3", "S4", "QR", "QR2"]:
            entry[key] = tensor_json(res[key], arrays, f"{kind_name}_{key}")
        # structure-only (gauge-dependent factors) and exact copies (digest)
        for key in ["U", "V", "Q", "R", "U3", "V3", "CJ", "M", "M0", "RV"] + ["T" + p for p in perms]:
            entry[key] = tensor_json(res[key])
        out[kind_name] = entry
    with open(os.path.join(HERE, "c1_ops.json"), "w") as f:
        json.dump(out, f)
    np.savez_compressed(os.path.join(HERE, "c1_ops.npz"), **arrays)


def storage_digest(storages):
    h = hashlib.sha256()
    for s in storages:
        h.update(np.ascontiguousarray(s, dtype=np.float64).tobytes())
    return h.hexdigest()


def tat_psis(spec, storages, configs):
    """psi(s) of each configuration from TAT in tetragono's order: value (0
    when the one-element result has no block) and the result's edge names."""
    sd = {"kind": spec.kind, "L1": spec.L1, "L2": spec.L2, "N": spec.N, "segments": list(spec.segments)}
    sites = peps.build_sites(sd, storages)
    lines = []
    for c, cfg in enumerate(configs):
        lat = peps.project(sites, cfg)
        for l1 in range(spec.L1):
            for l2 in range(spec.L2):
                lines += rr.tensor_lines(f"s{c}_{l1}_{l2}", lat[l1][l2])
        lines += [f"amplitude psi{c} {spec.L1} {spec.L2} {spec.cut} s{c}", f"print psi{c}"]
    res = rr.run(tat.KINDS[spec.kind], lines)
    psis = []
    for c in range(len(configs)):
        t = res[f"psi{c}"]
        psis.append({"value": float(t.data[0]) if t.size else 0.0, "names": t.names, "size": int(t.size)})
    return sd, psis


def make_amplitudes(out, specs):
    from tnsp_amd import peps as ppeps  # host-only parts: layout, seeded storages, configurations

    cases = {}
    for name, spec, nconf, particles in specs(ppeps):
        storages = ppeps.random_storages(spec)
        configs = ppeps.sample_configurations(spec, nconf, seed=spec.seed + 100, particles=particles)
        sd, psis = tat_psis(spec, storages, configs)
        for p in psis:
            del p["size"]
        cases[name] = {"spec": {**sd, "cut": spec.cut, "seed": spec.seed}, "storage_sha256": storage_digest(storages),
                       "configs": configs.astype(int).tolist(), "psi": psis}
    with open(os.path.join(HERE, out), "w") as f:
        json.dump(cases, f)


def make_zero_amplitudes(out="zero_amplitudes.json", candidates=48, zeros=5, nonzeros=3):
    """Configurations of the two 8x8 workloads for which TAT's amplitude is
    exactly zero (some contraction or decomposition of the sweep keeps no
    block, or every kept singular value vanishes), next to a few nonzero
    ones, picked in draw order from ``candidates`` seeded draws."""
    from tnsp_amd import peps as ppeps

    cases = {}
    for name, spec, particles in (("C3_8x8_Z2_D6_Dc36_zero", ppeps.C3, (24, 40)),
                                  ("C4_8x8_U1_D8_Dc64_zero", ppeps.C4, None)):
        storages = ppeps.random_storages(spec)
        configs = ppeps.sample_configurations(spec, candidates, seed=spec.seed + 200, particles=particles)
        sd, psis = tat_psis(spec, storages, configs)
        zi = [i for i, p in enumerate(psis) if p["value"] == 0.0][:zeros]
        ni = [i for i, p in enumerate(psis) if p["value"] != 0.0][:nonzeros]
        keep = sorted(zi + ni)
        cases[name] = {"spec": {**sd, "cut": spec.cut, "seed": spec.seed}, "storage_sha256": storage_digest(storages),
                       "candidates": candidates, "zero_count": sum(p["value"] == 0.0 for p in psis),
                       "configs": configs[keep].astype(int).tolist(), "psi": [psis[i] for i in keep]}
    with open(os.path.join(HERE, out), "w") as f:
        json.dump(cases, f)


def small_specs(ppeps):
    return (
        ("C2_4x4_D4_Dc16", ppeps.C2, 24, None),
        ("Z2_3x3_D4_Dc8", ppeps.LatticeSpec("FermiZ2", 3, 3, (2, 2), 4, 8, seed=21), 12, (2, 6)),
        ("U1_3x3_D4_exact", ppeps.LatticeSpec("FermiU1", 3, 3, (2, 1, 1), 4, 1000, seed=22), 8, None),
    )


def large_specs(ppeps):
    """The two 8x8 workloads at their full bond and cut dimensions."""
    return (
        ("C3_8x8_Z2_D6_Dc36", ppeps.C3, 6, (24, 40)),
        ("C4_8x8_U1_D8_Dc64", ppeps.C4, 6, None),
    )


if __name__ == "__main__":
    if not rr.available():
        rr.build()
    make_c1()
    make_amplitudes("c2_amplitudes.json", small_specs)
    make_amplitudes("peps8x8_amplitudes.json", large_specs)
    make_zero_amplitudes()
    print("golden fixtures written")
