"""Physical constants of the dynamical core.

Values restated from reference ``generate_parameters.py:20-39``; the derived
constants use the same double-precision operations, so they carry the same
bits as the reference's generated ``parameters.h``.  The HIP sources carry
the same definitions in ``csrc/params.h``.
"""
import math

pi = math.pi
g = 9.80665
Rd = 287.1
Rv = 461.5
eps_v = Rd / Rv
eps_vi = 1.0 / eps_v
cpd = 1004.0
cpv = 1859.0
cl = 4218.0
ci = 2106.0
kappa = Rd / cpd
Tf = 273.15
Tt = 273.16
T_tilde = 298.15
p_tilde = 10.0 ** 5
pv_star_t = 611.7
sd_tilde = 6864.8
sv_tilde = 10513.6
