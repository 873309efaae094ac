"""Case descriptions used by the tests and the benchmark (parameters.h fields by name).

The benchmark cases run with T_HOT=+0.5, T_COLD=-0.5, T_REF=0 and an initial temperature that
is linear between the walls plus seeded noise.  At Ra=1e8 the reference's default VISC would give
a free-fall velocity of 0.45 lattice units (globals.cu:26-30), so those cases set
VISC = 0.1414*LX*sqrt(Pr/Ra), the free-fall velocity of the reference's own case (0.1414).
"""
import math

RB_BASE = dict(t_hot=0.5, t_cold=-0.5, t_ref=0.0, activate_particles=0, particle_fraction=0.0)


def visc_for(lx, ra, pr):
    return 0.1414 * lx * math.sqrt(pr / ra)


def single_phase(lx, ly, lz, ra=1e5, pr=0.71, visc=None, **kw):
    d = dict(RB_BASE, lx=lx, ly=ly, lz=lz, rayleigh=ra, prandtl=pr)
    if visc is not None:
        d["visc"] = visc
    d.update(kw)
    return d


def config1():
    """64x64x32, Ra=1e5, Pr=0.71, default BETA and VISC (g0=1.2503, u0=0.0707)."""
    return single_phase(64, 64, 32, ra=1e5)


def config2(precision=0):
    """320x320x160, Ra=1e8, Pr=0.71."""
    return single_phase(320, 320, 160, ra=1e8, visc=visc_for(320, 1e8, 0.71), precision=precision)


def particle_case(lx, ly, lz, npart, diam, ra=1e5, pr=0.71, beta=None, **kw):
    d = single_phase(lx, ly, lz, ra=ra, pr=pr, **kw)
    d.update(activate_particles=1, npart_override=npart, particle_diam_frac=diam / lx,
             rho_particle_ratio=2.0, cp_particle_ratio=1.0, particle_fraction=0.0)
    if beta is not None:
        d["beta"] = beta
    return d


def config3(precision=0):
    """320x320x160, Ra=1e8, Pr=0.71, 300 spheres of diameter 10 (PARTICLE_DIAM_FRAC = 1/32,
    fraction 0.0096 -> floor(N*0.0096/V) = 300), density ratio 2, heat-capacity ratio 1.
    BETA is raised to 0.625 so that g0 = 1e-4 (fluid buoyancy g0*BETA is unchanged) and the
    spheres' buoyancy (1 - 1/rho_r) g0 stays a small acceleration."""
    d = single_phase(320, 320, 160, ra=1e8, visc=visc_for(320, 1e8, 0.71), precision=precision,
                     beta=0.625)
    d.update(activate_particles=1, particle_diam_frac=1.0 / 32.0, particle_fraction=0.0096,
             rho_particle_ratio=2.0, cp_particle_ratio=1.0, npart_override=-1)
    return d


def local_particle_temperature(temperature, ppos, lx, ly, lz):
    """Initial sphere temperature = fluid temperature at the cell holding the centre."""
    import numpy as np
    np_ = len(ppos) // 3
    x, y, z = ppos[:np_], ppos[np_:2 * np_], ppos[2 * np_:]
    ix = np.clip(np.floor(x).astype(int), 0, lx - 1)
    iy = np.clip(np.floor(y).astype(int), 0, ly - 1)
    iz = np.clip(np.floor(z).astype(int), 0, lz - 1)
    return temperature[(iz * ly + iy) * lx + ix]


def config5(precision=0):
    """512x512x256, Ra=1e9, Pr=0.71, 5000 spheres of diameter 8 (PARTICLE_DIAM_FRAC = 1/64,
    fraction 0.019975 -> floor(N*0.019975/V) = 5000), density ratio 2, heat-capacity ratio 1.
    BETA = 0.39 puts g0 at 1.0e-4, as for the 320x320x160 particle case."""
    d = single_phase(512, 512, 256, ra=1e9, visc=visc_for(512, 1e9, 0.71), precision=precision,
                     beta=0.39)
    d.update(activate_particles=1, particle_diam_frac=1.0 / 64.0, particle_fraction=0.019975,
             rho_particle_ratio=2.0, cp_particle_ratio=1.0, npart_override=-1)
    return d
