"""Seeded synthetic workloads shared by the tests and the benchmark.

`multi_particle_sweep` is the many-particle extinction sweep: particles of
uneven face counts, each with its own Drude gold-like material, solved over
a common wavelength grid in vacuum.
"""
import numpy as np

from . import shapes

LEVEL_FACES = {2: 320, 3: 1280, 4: 5120}


def multi_particle_sweep(seed=2026, nparticles=12):
    """Particles of the sweep: semi-axes U(5, 40) nm each, icosphere level in
    {2, 3, 4} (320, 1,280 or 5,120 faces), Drude eps_inf in U[5, 12], wp in
    U[8, 9.5] eV and gamma in U[0.05, 0.2] eV; eps outside 1, wavelengths
    linspace(400, 900, 128).  Returns (particles, wavelengths); each particle
    is a dict with verts, faces, semi_axes, level and drude (eps_inf, wp,
    gamma)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nparticles):
        axes = tuple(float(a) for a in rng.uniform(5.0, 40.0, 3))
        level = int(rng.choice([2, 3, 4]))
        drude = (float(rng.uniform(5.0, 12.0)), float(rng.uniform(8.0, 9.5)), float(rng.uniform(0.05, 0.2)))
        v, f = shapes.ellipsoid(level, axes)
        out.append({"verts": v, "faces": f, "semi_axes": axes, "level": level, "drude": drude})
    return out, np.linspace(400.0, 900.0, 128)
