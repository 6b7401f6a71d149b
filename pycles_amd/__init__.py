"""MI355X (gfx950) dynamical core of PyCLES.

Python operator classes mirroring the reference's ``update(...)`` surface
(``Grid``, ``ReferenceState``, ``PrognosticVariables``, ``ScalarAdvection``,
``MomentumAdvection``, ``ThermodynamicsSA``/``ThermodynamicsDry``,
``PressureSolver``, ``TimeStepping``, ``Simulation3d``) over the HIP kernels
of ``libpycles_amd.so`` (C ABI: ``include/pycles_amd.h``).  Importing the
package does not touch the GPU; the operators need the built library and a
visible GPU and raise otherwise.
"""
__all__ = ["build"]
