'''Length conversion used for geometries (CODATA 2014, as orbkit/units.py:9-28).'''
from math import pi

_h = 6.626070040e-34        # J s
_hbar = _h / (2 * pi)
_e = 1.6021766208e-19       # C
_me = 9.10938356e-31        # kg
_e0 = 8.854187817e-12       # F/m
_a0 = 4 * pi * _e0 * _hbar ** 2 / (_me * _e ** 2)   # Bohr radius, m

aa_to_a0 = 1e-10 / _a0
a0_to_aa = _a0 / 1e-10
