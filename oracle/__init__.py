"""CPU FP64 restatement of the curved-sky MASTER path -- TEST INFRASTRUCTURE ONLY.

This package is the oracle: a plain numpy/scipy (and, under ``oracle/c``, C/OpenMP)
restatement of the reference algorithm, written from the mathematics.  It exists to
check the HIP product path.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import or run it, and only as the checker:
the product package ``gmaster_amd`` never imports it and has no CPU fallback.

Every function cites the reference file:line whose behaviour it restates and says
what pins it (closed forms, invariants, brute force, library routines).  Functions
with no independent check say "parity unpinned" in their docstring.

Modules
    geometry  HEALPix RING geometry and healpy alm packing.
    legendre  Normalised Wigner-d rows lambda_lm^(s)(theta) by FP64 recurrence with an
              extended exponent; Wigner's explicit sum for pinning.
    sht       Ring DFT, spin-0 / spin-s analysis and synthesis, Jacobi refinement, E/B.
    wigner3j  Exact (integer) Racah 3j symbols and the FP64 recursions in l3.
    coupling  Mode-coupling matrices 0-0, 0-2, 2-2 (+/-) and the block layout.
    master    Coupled C_ell, binning, binned MCM and the FP64 decoupling solve.
    apodize   C1/C2 mask apodization.
    toeplitz  NaMaster's Toeplitz approximation of the scalar coupling matrix.
"""
