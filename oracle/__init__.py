"""CPU oracle for the redshift chi2 scan — TEST INFRASTRUCTURE, NOT PRODUCT.

This package is a float64 restatement of redrock's CPU algorithm for the chi2
scan path (rebin, IGM, normal equations, solve, chi2, [OII] penalty, minima
refinement).  It exists only to check the HIP implementation in
``redrock_amd``: the tests, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py`` may import it, and nothing else may.  The product path
never routes through it.

Each function cites the reference file:line it follows.  Parity status of
each function (what pins it besides the restatement itself) is recorded in
the module docstring of ``oracle.rr_oracle``.
"""
