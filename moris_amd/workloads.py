"""Seeded level-set inputs of the assembly workloads (sphere arrays, random sphere sets, the 2D
circle and the 3D sphere diffusion cases).

Random numbers come from splitmix64 so that every host produces the same spheres.
"""
import numpy as np

from . import IWG_STRUC_LINEAR_BULK, make_spec

BASE_SEED = 20261202
_MASK = (1 << 64) - 1


def splitmix64_uniform(seed, count):
    """count uniforms in [0, 1) from the splitmix64 stream started at seed."""
    out = np.empty(count)
    x = seed & _MASK
    for i in range(count):
        x = (x + 0x9E3779B97F4A7C15) & _MASK
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        z ^= z >> 31
        out[i] = (z >> 11) * (1.0 / (1 << 53))
    return out


def sphere_array(variant, per_axis=4, radius=0.1, spread=0.1, seed=BASE_SEED + 5):
    """Spheres on a per_axis^3 grid of the unit cube, radii radius * U[1 - spread, 1 + spread]
    drawn per sphere and per design variant."""
    n = per_axis ** 3
    u = splitmix64_uniform(seed * 1000003 + variant, n)
    g = (np.arange(per_axis) + 0.5) / per_axis
    cz, cy, cx = np.meshgrid(g, g, g, indexing="ij")
    r = radius * (1.0 - spread + 2.0 * spread * u)
    return np.stack([cx.ravel(), cy.ravel(), cz.ravel(), r], axis=1)


def random_spheres(n_spheres=64, seed=BASE_SEED + 3, rmin=0.05, rmax=0.12):
    """n_spheres spheres with centres uniform in the unit cube and radii uniform in [rmin, rmax]."""
    u = splitmix64_uniform(seed, 4 * n_spheres).reshape(n_spheres, 4)
    return np.hstack([u[:, :3], (rmin + (rmax - rmin) * u[:, 3]).reshape(-1, 1)])


def elasticity_spec(n, spheres, material_phase=1):
    """3D linear elasticity (E = 1, nu = 0.3) on an n^3 hex8 lattice of the unit cube; the material
    is the region outside the spheres (material_phase 1, level set positive)."""
    return make_spec(3, (n, n, n), iwg=IWG_STRUC_LINEAR_BULK, spheres=spheres, material_phase=material_phase)


def circle_2d_spec(n=64):
    """Unit square, n x n QUAD4, circle hole r = 0.25 at the centre (material outside), plane stress
    E = 1, nu = 0.3, p = 1 (TRI_7 on the cut sub-triangles)."""
    return make_spec(2, (n, n), iwg=IWG_STRUC_LINEAR_BULK, spheres=[[0.5, 0.5, 0.0, 0.25]], seed=BASE_SEED)


def sphere_diffusion_spec(n=128):
    """Unit cube, n^3 HEX8 diffusion, sphere |x - c| - 0.3 (c the centre; material outside),
    k_e = exp(N(0, 0.5^2)) per element (seeded), p = 1 (TET_11 on the cut sub-tets)."""
    return make_spec(3, (n, n, n), spheres=[[0.5, 0.5, 0.5, 0.3]], k_sigma=0.5, seed=BASE_SEED + 2)

