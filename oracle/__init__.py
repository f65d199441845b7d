"""CPU FP64 oracle for the periodic HEX8 effective-stiffness path.

TEST INFRASTRUCTURE ONLY.  This package restates, in NumPy/SciPy, the
reference's CPU algorithm (and the mathematics of its GPU chain) so that the
HIP implementation can be checked against it.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or as the timed CPU baseline.  The product
package never imports it; a missing HIP library is an error there, never a
reason to fall back to this code.

Modules
-------
materials          material law, phase table, orientation frames, rotations
fem                8-point Q1 assembly + sparse direct solve, matrix-free
                   element action, RHS, stress recovery, uniform stencil
spectral           FI Fourier symbol of an isotropic reference medium and the
                   FFT Green preconditioner
reference_medium   Zeman sampling and the closed-form soft-phase design
pcg                preconditioned CG with the reference's safeguards
homogenize         the CPU backend (sparse direct, <=4096 voxels) and an FP64
                   matrix-free FFT-PCG C_eff solver for larger cells

Pinning (what fixes each function independently of its own formula) is listed
in each module docstring.  Every function listed there has at least one check
in ``tests/test_oracle_*.py``.
"""
