"""CPU oracle for the ptychography forward/adjoint path -- TEST INFRASTRUCTURE ONLY.

Nothing in the shipped package imports this directory.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may use
it, and only as the checker (or, for ``bench.py``, as the timed CPU baseline),
never as the thing that is measured or shipped.

Contents
--------
``ptycho_oracle``   FP64 NumPy restatement of the four operators (conventional,
                    Fourier, coded, near-field), their hand-written adjoints and
                    the amplitude (MSE) loss.  Every function cites the
                    reference file:line it follows.
``torch_literal``   FP32 torch-CPU restatement of the reference's torch
                    fallback, including its per-view Python loops; this is the
                    "reference CPU path" timed as the CPU baseline.

How the oracle is pinned (``tests/test_oracle.py``): brute-force DFT and
closed-form fftshift on small frames, brute-force crop/wrap loops, central
finite differences of every gradient (Re/Im separately), FP64 dot-product
adjoint identities, closed-form special cases, Parseval, and the reference's
own ``test_ops.py`` cases (same seeds and shapes, torch autograd of the
reference's native formulas).  No function here is "parity unpinned".
"""
