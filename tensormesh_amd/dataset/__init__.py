from .grf import grf_lattice, lognormal_kappa_cube, lognormal_kappa_points
from .batched import batched_square_meshes
