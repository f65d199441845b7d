"""MI355X-native periodic HEX8 homogenization: the effective-stiffness solve path.

Drop-in for the GPU backend of Voxel-LMFEM's ``homogenize`` (same signature,
inputs and result keys); the arithmetic runs in libvxl.so, hand-written gfx950
HIP kernels reached through a C ABI.
"""

__version__ = "0.1.3+mi355x"


def homogenize(phase_id, materials, domain_size, *, orientation=None, backend="cuda", warmup=False,
               precision="mixed", tol=1e-8, fields=False, preconditioner="fft", element="fi",
               hourglass_coefficient=0.1, max_iter=10000, stencil="modal", solver="cg"):
    """Solve six unit-strain loads; returns the reference's result dictionary
    (see ``voxel_lmfem.solver.homogenize``)."""
    from .solver import homogenize as solve
    return solve(phase_id, materials, domain_size, orientation=orientation, backend=backend,
                 warmup=warmup, precision=precision, tol=tol, fields=fields,
                 preconditioner=preconditioner, element=element, hourglass_coefficient=hourglass_coefficient,
                 max_iter=max_iter, stencil=stencil, solver=solver)


__all__ = ["homogenize", "__version__"]
