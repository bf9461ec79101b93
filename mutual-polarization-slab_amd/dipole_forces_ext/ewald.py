"""dipole_forces_ext.ewald: the GPU spectral-Ewald module under the package
name the reference uses ("CPU - HOOMD simulation test/dipole_forces_ext/ewald.py",
identical to "CPU - Ewald test/ewald.py")."""
import ewald as _impl

globals().update({k: v for k, v in vars(_impl).items() if not k.startswith("__")})
