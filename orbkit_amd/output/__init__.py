'''Output writers for grid data (Gaussian cube).'''
from .cube import cube_creator  # noqa: F401
