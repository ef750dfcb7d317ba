"""Golden vectors written by the reference TAT (tests/golden/make_golden.py):
the oracle (CPU) and the GPU package must reproduce them.

Bar: names / segments / arrows bit-exact; pure data movement (transpose,
merge, reverse, conjugate) bit-exact (sha256 of the storage); contraction and
decompositions within 1e-12 (oracle) / 1e-10 (GPU) relative on the
gauge-invariant results (C, S, U S V, Q R); psi(s) within 1e-10 relative
with the same sign.
"""
import itertools

import numpy as np
import pytest

from helpers import assert_structure, c1_pair, digest, load_golden, load_golden_8x8, load_golden_zero, storage, to_gpu

OPS, ARRAYS, AMPS = load_golden()
AMPS8 = load_golden_8x8()
ZERO = load_golden_zero()
KINDS = ["FermiU1", "BoseU1", "FermiZ2"]


@pytest.fixture(params=["oracle", pytest.param("gpu", marks=pytest.mark.gpu)])
def backend(request):
    return request.param


def rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)) if b.size else 0.0


def inputs(kind, backend):
    from oracle import tat

    ref = OPS[kind]
    A, B = c1_pair(tat, kind, ref["seed"])
    assert digest(A.data) == ref["A"]["sha256"] and digest(B.data) == ref["B"]["sha256"]
    if backend == "gpu":
        return to_gpu(A), to_gpu(B)
    return A, B


@pytest.mark.parametrize("kind", KINDS)
def test_c1_exact_data_movement(kind, backend):
    A, _ = inputs(kind, backend)
    ref = OPS[kind]
    for p in itertools.permutations("ijkl"):
        t = A.transpose(list(p))
        assert_structure(t, ref["T" + "".join(p)])
        assert digest(storage(t)) == ref["T" + "".join(p)]["sha256"]
    for key, t in (("CJ", A.conjugate()), ("M", A.merge_edge({"mm": ["j", "i"]}, True)),
                   ("M0", A.merge_edge({"mm": ["k", "l"]}, False)), ("RV", A.reverse_edge({"i", "k"}, True))):
        assert_structure(t, ref[key])
        assert digest(storage(t)) == ref[key]["sha256"], key


@pytest.mark.parametrize("kind", KINDS)
def test_c1_contract_svd_qr(kind, backend):
    A, B = inputs(kind, backend)
    ref = OPS[kind]
    tol = 1e-12 if backend == "oracle" else 1e-10
    C = A.contract(B, {("k", "k"), ("l", "l")})
    assert_structure(C, ref["C"])
    assert rel(storage(C), ARRAYS[f"{kind}_C"]) < tol
    u, s, v = C.svd({"i", "j"}, "cu", "cv", "su", "sv", 16)
    assert_structure(u, ref["U"])
    assert_structure(s, ref["S"])
    assert_structure(v, ref["V"])
    assert rel(storage(s), ARRAYS[f"{kind}_S"]) < tol
    usv = u.contract(s, {("cu", "su")}).contract(v, {("sv", "cv")})
    assert_structure(usv, ref["USV"])
    assert rel(storage(usv), ARRAYS[f"{kind}_USV"]) < tol
    _, s2, _ = C.svd({"j"}, "cu", "cv", "su", "sv", 5)
    assert_structure(s2, ref["S2"])
    assert rel(storage(s2), ARRAYS[f"{kind}_S2"]) < tol
    # relative cut (PyTAT.hpp:1040-1047: 0 < cut < 1 keeps the values above
    # cut * the largest one, svd.hpp:430-481)
    u3, s3, v3 = C.svd({"i", "j"}, "cu", "cv", "su", "sv", 0.3)
    assert_structure(u3, ref["U3"])
    assert_structure(s3, ref["S3"])
    assert_structure(v3, ref["V3"])
    assert rel(storage(s3), ARRAYS[f"{kind}_S3"]) < tol
    usv3 = u3.contract(s3, {("cu", "su")}).contract(v3, {("sv", "cv")})
    assert_structure(usv3, ref["USV3"])
    assert rel(storage(usv3), ARRAYS[f"{kind}_USV3"]) < tol
    _, s4, _ = C.svd({"m"}, "cu", "cv", "su", "sv", 0.6)
    assert_structure(s4, ref["S4"])
    assert rel(storage(s4), ARRAYS[f"{kind}_S4"]) < tol
    for key, direction in (("QR", "r"), ("QR2", "q")):
        q, r = C.qr(direction, {"m", "n"}, "cq", "cr")
        if key == "QR":
            assert_structure(q, ref["Q"])
            assert_structure(r, ref["R"])
        qr = q.contract(r, {("cq", "cr")})
        assert_structure(qr, ref[key])
        assert rel(storage(qr), ARRAYS[f"{kind}_{key}"]) < tol


def lattice_case(name):
    from tnsp_amd import peps as ppeps

    case = AMPS[name] if name in AMPS else (AMPS8[name] if name in AMPS8 else ZERO[name])
    sd = case["spec"]
    spec = ppeps.LatticeSpec(sd["kind"], sd["L1"], sd["L2"], tuple(sd["segments"]), sd["N"], sd["cut"], sd["seed"])
    storages = ppeps.random_storages(spec)
    h = __import__("hashlib").sha256()
    for st in storages:
        h.update(np.ascontiguousarray(st, dtype=np.float64).tobytes())
    assert h.hexdigest() == case["storage_sha256"], "site generator drifted"
    return spec, storages, np.array(case["configs"]), case["psi"]


@pytest.mark.parametrize("name", list(AMPS) + list(AMPS8))
def test_amplitude_oracle_vs_tat(name):
    from oracle import peps

    spec, storages, configs, psis = lattice_case(name)
    sd = {"kind": spec.kind, "L1": spec.L1, "L2": spec.L2, "N": spec.N, "segments": list(spec.segments)}
    sites = peps.build_sites(sd, storages)
    for cfg, ref in list(zip(configs, psis))[:8 if name in AMPS else 1]:
        v, t = peps.psi(sites, cfg, spec.cut)
        assert t.names == ref["names"]
        assert abs(v - ref["value"]) <= 1e-10 * abs(ref["value"]) + 1e-300


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(AMPS) + list(AMPS8))
def test_amplitude_gpu_plain_and_engine_vs_tat(name):
    from tnsp_amd import peps as ppeps
    from tnsp_amd.engine import AmplitudeEngine

    spec, storages, configs, psis = lattice_case(name)
    lat = ppeps.Lattice(spec, storages)
    ref = np.array([p["value"] for p in psis])
    for c in range(min(6 if name in AMPS else 1, len(configs))):
        v, t = ppeps.amplitude_plain(lat, configs[c])
        assert t.names == psis[c]["names"]
        assert abs(v - ref[c]) <= 1e-10 * abs(ref[c]) + 1e-300
    for chunk, lanes in ((7, 1), (5, 3)):
        eng = AmplitudeEngine(lat, chunk=chunk, lanes=lanes)
        got = eng.evaluate(configs)
        np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-300)
        assert np.array_equal(np.sign(got), np.sign(ref))


def test_zero_fixture_has_zero_and_nonzero_amplitudes():
    """The C4 fixture holds configurations whose TAT amplitude is exactly 0
    and others whose amplitude is not; the C3 draws had none."""
    c4 = ZERO["C4_8x8_U1_D8_Dc64_zero"]
    vals = [p["value"] for p in c4["psi"]]
    assert c4["zero_count"] > 0 and 0.0 in vals and any(v != 0.0 for v in vals)


def test_zero_amplitude_oracle_vs_tat():
    """The oracle gives exactly 0 where TAT does (one zero and one nonzero
    C4 configuration: each takes seconds in numpy)."""
    from oracle import peps

    spec, storages, configs, psis = lattice_case("C4_8x8_U1_D8_Dc64_zero")
    sd = {"kind": spec.kind, "L1": spec.L1, "L2": spec.L2, "N": spec.N, "segments": list(spec.segments)}
    sites = peps.build_sites(sd, storages)
    zi = next(i for i, p in enumerate(psis) if p["value"] == 0.0)
    ni = next(i for i, p in enumerate(psis) if p["value"] != 0.0)
    for i in (zi, ni):
        v, t = peps.psi(sites, configs[i], spec.cut)
        assert t.names == psis[i]["names"]
        if psis[i]["value"] == 0.0:
            assert v == 0.0
        else:
            assert abs(v - psis[i]["value"]) <= 1e-10 * abs(psis[i]["value"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(ZERO))
def test_zero_amplitudes_gpu_plain_and_engine_vs_tat(name):
    """Exactly zero where TAT's amplitude is exactly zero, TAT's value to
    1e-10 with its sign elsewhere, on the plain path and on the engine."""
    from tnsp_amd import peps as ppeps
    from tnsp_amd.engine import AmplitudeEngine

    spec, storages, configs, psis = lattice_case(name)
    lat = ppeps.Lattice(spec, storages)
    ref = np.array([p["value"] for p in psis])
    zero = ref == 0.0
    for c in range(len(configs)):
        if c > 0 and zero[c] == zero[c - 1]:
            continue  # the plain path is slow: one zero and one nonzero case suffice
        v, t = ppeps.amplitude_plain(lat, configs[c])
        assert t.names == psis[c]["names"]
        assert (v == 0.0) if zero[c] else abs(v - ref[c]) <= 1e-10 * abs(ref[c])
    for chunk, lanes in ((8, 1), (3, 2)):
        got = AmplitudeEngine(lat, chunk=chunk, lanes=lanes).evaluate(configs)
        assert np.array_equal(got == 0.0, zero), (got, ref)
        np.testing.assert_allclose(got[~zero], ref[~zero], rtol=1e-10, atol=0)
        assert np.array_equal(np.sign(got), np.sign(ref))
