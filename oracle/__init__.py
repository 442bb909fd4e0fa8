"""CPU oracle for the two-site DMRG / TDVP hot path -- TEST INFRASTRUCTURE ONLY.

This package is a NumPy/SciPy FP64 (complex128) restatement of the
reference pure-Python backend (dmrgpy's ``pyitensor``), written with fixed
array layouts instead of an Index engine.  It exists to check the HIP
product path; nothing in ``dmrgpy_amd`` may import it.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg use it,
and only as the checker / the timed CPU baseline.

Layouts (row-major, complex128, dimension-1 links at the open ends, which
is arithmetically identical to the reference's link-free boundary tensors):

* MPS site      ``A[Dl, d, Dr]``
* MPO site      ``W[wl, s_in, s_out, wr]``  (the reference's (in,out) storage
  convention, ``src/dmrgpy/pyitensor/sites/base.py:1-16``)
* environments  ``L[a_bra, w, a_ket]`` and ``R[b_bra, w, b_ket]``

Modules and what pins them (each function cites the reference line it
follows):

* ``sites``      operator tables        -- pinned by commutation relations and
  the reference's own table entries in tests/test_oracle_sites.py
* ``mpo``        terms -> MPO, dense H  -- pinned by dense Kronecker H vs ED
* ``linalg``     truncation rule, Gram SVD, Lanczos, Krylov expm -- pinned by
  brute force / numpy.linalg / scipy.linalg.expm on small inputs
* ``mps``        random MPS, canonical moves, inner -- pinned by dense vectors
* ``dmrg``       sweeps                 -- pinned by ED, closed forms and the
  golden energies written from the reference (tests/golden/)
* ``tdvp``       TDVP step              -- pinned by dense expm at N<=8 and the
  golden correlators from the reference
* ``ed``         exact diagonalisation  -- pinned by closed forms (dimer, N=3,
  free fermions)
* ``free_fermion`` XX-chain closed forms
"""
