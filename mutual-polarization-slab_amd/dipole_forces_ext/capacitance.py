"""dipole_forces_ext.capacitance: the GPU mutual-polarization solver under the
package name the reference uses
("CPU - HOOMD simulation test/dipole_forces_ext/capacitance.py")."""
import capacitance as _impl

globals().update({k: v for k, v in vars(_impl).items() if not k.startswith("__")})
