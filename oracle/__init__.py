"""CPU oracle for the block-sparse backend hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker or as the timed
CPU baseline.  The product path (``yastn_amd``) never imports it and has no CPU
fallback.

``oracle.np_backend`` restates, in numpy, the reference numpy backend's hot-path
functions and the reference torch backend's hand-written adjoints; each function
cites the reference file:line it follows.  ``oracle/Makefile`` builds the
reference's own C merge worker (``experimental/tm_worker.c``) into
``oracle/_ref/`` when the reference tree is present; tests use it to pin the
numpy restatement of ``transpose_and_merge``.

How each oracle function is pinned (tests/test_oracle.py):
  transpose_and_merge  golden vectors of the reference numpy backend; the
                       reference C worker (oracle/_ref); dense closed form
                       (np.transpose of the densified tensor)
  unmerge              golden vectors; round trip unmerge(merge(x)) == x
  dot                  golden vectors; dense np.tensordot closed form
  svd                  golden vectors (singular values); U S Vh == A,
                       orthonormality; LAPACK through numpy.linalg.svd
  eigh                 golden vectors (eigenvalues); U diag(w) U^H == A
  fix_svd_signs        known-answer test (reference tests/tensor/test_svd.py:159-186)
                       and golden vectors
  apply_mask           golden vectors of the truncation scenarios
  *_adjoint (copies)   adjoint identity <M x, y> == <x, M^T y> against the forward
  dot_adjoint          finite differences; reference torch gradients (golden)
  svd_adjoint          reference torch gradients (golden, per-block fixtures);
                       central finite differences
  eigh_adjoint         central finite differences (torch's documented formula)
"""
