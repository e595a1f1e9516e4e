"""Named SimParams used across the tests.

The reference's configs/*.cfg (solid_angle, absorption, default,
wrapped_lambertian) and the benchmark workloads: C1 plain small (50 mm
polished cube, 10^6 photons) and C2 long histories (100 x 100 x 50 mm
Lambertian box, 10^8 photons). Values restated from configs/*.cfg and
params.hpp:42-62.
"""
import math

from gpu_optical_photons_amd.params import SimParams, Source, Surface

INF = math.inf


def black_box(**kw):
    """tests/test_transport_cpu.cpp:41-49: index-matched readout, black faces,
    no bulk processes, prompt emission."""
    d = dict(surface=Surface.BLACK, n_det=1.63, abs_length=INF, scat_length=INF, tau=0.0)
    d.update(kw)
    return SimParams(**d)


def solid_angle(**kw):  # configs/solid_angle.cfg
    return black_box(src_x=10.0, src_y=-5.0, src_z=3.0, **kw)


def absorption(**kw):  # configs/absorption.cfg
    return black_box(abs_length=100.0, src_x=6.0, src_y=-8.0, src_z=-5.0, **kw)


def default(**kw):  # configs/default.cfg
    return SimParams(**kw)


def wrapped_lambertian(**kw):  # configs/wrapped_lambertian.cfg
    d = dict(surface=Surface.LAMBERTIAN, reflectivity=0.95, n_det=1.5, abs_length=1000.0,
             scat_length=2000.0, source=Source.VOLUME, tau=4.0)
    d.update(kw)
    return SimParams(**d)


def c1_plain(**kw):
    """Plain method, small: 50 mm cube, polished faces to n_out 1.0, readout
    n_det 1.5, abs 1000 mm, no Rayleigh, point source at the centre, tau 4 ns,
    max_steps 10000, seed 1."""
    d = dict(n_scint=1.63, n_det=1.5, n_out=1.0, abs_length=1000.0, scat_length=INF,
             surface=Surface.POLISHED, source=Source.POINT, tau=4.0, max_steps=10000, seed=1)
    d.update(kw)
    return SimParams(**d)


def c2_long(**kw):
    """Long histories: 100 x 100 x 50 mm, Rayleigh 300 mm, Lambertian R 0.97,
    readout n_det 1.5, volume source, abs 2000 mm, n_scint 1.63, tau 4, seed 1."""
    d = dict(hx=50.0, hy=50.0, hz=25.0, n_scint=1.63, n_det=1.5, n_out=1.0, abs_length=2000.0,
             scat_length=300.0, surface=Surface.LAMBERTIAN, reflectivity=0.97,
             source=Source.VOLUME, tau=4.0, max_steps=10000, seed=1)
    d.update(kw)
    return SimParams(**d)


def every_process(**kw):
    """tests/test_gpu.cpp:81-83: polished walls (Fresnel, TIR), Rayleigh,
    absorption, volume source, emission time."""
    d = dict(scat_length=300.0, source=Source.VOLUME)
    d.update(kw)
    return SimParams(**d)


def specular(**kw):
    d = dict(surface=Surface.SPECULAR, reflectivity=0.97, n_det=1.45, scat_length=150.0,
             source=Source.VOLUME, hx=10.0, hy=30.0, hz=20.0)
    d.update(kw)
    return SimParams(**d)


# Configurations the parity tests sweep: every surface model, both sources,
# Rayleigh on and off, a max_steps cap, prompt emission, off-centre sources.
PARITY = {
    "solid_angle": solid_angle(),
    "absorption": absorption(),
    "default": default(),
    "wrapped_lambertian": wrapped_lambertian(),
    "c1_plain": c1_plain(),
    "c2_long": c2_long(),
    "every_process": every_process(),
    "specular": specular(),
    "capped": every_process(max_steps=5),
    "black_scatter": black_box(scat_length=15.0, seed=31),
    "mirror_trap": SimParams(surface=Surface.SPECULAR, reflectivity=1.0, n_det=1.0,
                             abs_length=INF, scat_length=INF, max_steps=1000, src_x=3.0,
                             src_y=-11.0, src_z=7.0, seed=5),
}
