"""CPU oracle for the Fraunhofer/MFT path (test infrastructure only).

This package is a complex128 NumPy restatement of the reference algorithm
(lentil v0.8.8).  It is the checker the GPU path is compared against and is
never part of the product: only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it.

Modules
-------
``mft``      dft2 / idft2 and the analytic adjoint (fourier.py:79-121,193-198)
``optics``   extents, Field bookkeeping, Wavefront, Plane phasors, tilts and
             propagate_dft (extent.py, field.py, wavefront.py, plane.py,
             propagate.py)
``inputs``   input generators used to build seeded configurations
             (shape.py, segmented.py, zernike.py, util.py, helper.py)
``configs``  the five named workloads as seeded input builders and the
             reference user loop over wavelengths
"""
