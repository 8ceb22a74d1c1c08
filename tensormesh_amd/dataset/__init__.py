from .grf import (grf_lattice, grf_lattice_device, lognormal_kappa_cube, lognormal_kappa_cube_device,
                  lognormal_kappa_points)
from .batched import batched_square_meshes
