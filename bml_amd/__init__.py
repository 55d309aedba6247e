"""bml_amd: MI355X (gfx950) backend for bml's thresholded sparse-multiply path.

* :mod:`bml_amd.hip` -- ctypes binding of ``libbml_hip.so`` (the C ABI in
  ``include/bml_hip.h``) with a thin mirror of bml's C API.
* :mod:`bml_amd.ell` -- host ELLPACK-R / ELLBLOCK containers and file format.
* :mod:`bml_amd.generators` -- seeded synthetic inputs.

Importing the package does not touch the GPU; :func:`bml_amd.hip.lib` loads
the shared library (and raises if it is missing).
"""
__all__ = ["ell", "generators", "hip"]
