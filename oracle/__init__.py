"""CPU oracle for the symmetric block-sparse tensor path and the PEPS amplitude.

TEST INFRASTRUCTURE ONLY.  This package is a plain CPU restatement of the
reference algorithms (TAT's tensor core and tetragono's single-layer
boundary-MPS amplitude).  It exists so that the GPU product path can be
checked against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import, call or link anything under
``oracle/`` -- and only as the checker, never as the thing measured or
shipped.  The product package ``tnsp_amd`` never imports it.

Modules
-------
``oracle.tat``   Symmetry / Edge / block-sparse Tensor with TAT semantics:
                 edge_operator (split, reverse, transpose, merge, cut) with the
                 four fermion sign sources, contract, svd (with TAT's global
                 truncation rule), qr, conjugate, norms.
``oracle.peps``  tetragono-style PEPS construction and the single-layer
                 boundary-MPS amplitude psi(s), plus the Metropolis ratio.
``oracle/ref``   Build recipe for ``oracle/_ref``: small C++ drivers compiled
                 directly against the reference TAT headers, used to produce
                 golden vectors and as the timed CPU baseline.

Parity pinning (what checks each function, see ``tests/test_oracle_*.py``):

* block layout / storage order: PyTAT known-answer tests restated
  (test_create_symmetry_tensor, test_create_fermi_tensor) + TAT golden vectors.
* edge_operator signs: closed-form transpose sign, merge/split ``(n_odd & 2)``
  sign, round trips, and TAT golden vectors for all 24 permutations.
* contract: exact-integer known answers, symmetric == dense (brute force),
  fermion invariances, TAT golden vectors.
* svd / qr: reconstruction + orthonormality, singular values == numpy on the
  dense blocks, truncation == brute-force global top-k, TAT golden vectors of
  the gauge-invariant products (U S V, Q R, singular values, cut segments).
* amplitude: exact dense contraction (no truncation) on small lattices, gauge
  invariance, and the oracle/_ref TAT-based amplitude driver.
"""
