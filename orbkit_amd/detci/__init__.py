'''Grid reductions of orbkit's detCI module on the device path.'''
from .ci_core import rho, jab, a_nabla_b, get_rho_full, get_j_full, get_jab_full  # noqa: F401
