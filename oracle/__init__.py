"""FP64 CPU oracle for the BP-Layers hot path (TEST INFRASTRUCTURE ONLY).

This package restates, in plain numpy at float64, the algorithms of the
reference's CUDA ops (VLOGroup/bp-layers):

* ``sad``     -- stereo SAD cost volume (ops/sad) and the fused flow
                 SAD + min-projection op (ops/flow_mp_sad), forward and backward;
* ``lbp``     -- min-sum loopy belief propagation on the 4-connected grid for
                 the three pairwise variants (ops/lbp_stereo,
                 ops/lbp_semantic_pw, ops/lbp_semantic_pw_pixel): forward
                 sweeps, the reference's backward ("mode R") and the exact
                 adjoint ("mode E");
* ``packing`` -- the Python-side parameter packing around those ops.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker. The product
path (``bp-layers_amd/``) never imports it and has no CPU fallback.

Every function cites the reference file:line it follows and names the
independent check that pins it (tests/test_oracle_*.py).
"""
