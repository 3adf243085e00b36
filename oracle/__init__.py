"""CPU oracle for the Fraunhofer/MFT path (test infrastructure only).

This package is a complex128 NumPy restatement of the reference algorithm
(lentil v0.8.8).  It is the checker the GPU path is compared against and is
never part of the product: only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it.

Modules
-------
``mft``        dft2 / idft2 and the analytic adjoint (fourier.py:79-121,193-198)
``optics``     extents, Field bookkeeping, Wavefront, Plane phasors, tilts,
               propagate_dft / propagate_fft and ``broadband_image``, the
               reference user loop over wavelengths (extent.py, field.py,
               wavefront.py, plane.py, propagate.py,
               docs/user/diffraction.rst:148-160)
``inputs``     input generators used to build seeded configurations
               (shape.py, segmented.py, zernike.py, util.py, helper.py)
``detector``   image operators on the accumulated PSF: pixel, pixelate,
               jitter, smear, rebin, rescale (detector.py, convolvable.py,
               util.py), with numpy FFTs and scipy map_coordinates as the
               reference calls them
``retrieval``  the complex128 phase-retrieval forward model and its exact
               gradient (no reference counterpart; pinned by finite
               differences)

The five named workloads are built by ``lentil_amd.workloads`` (seeded host
arrays only); the oracle evaluates the reference loop on them.
"""
