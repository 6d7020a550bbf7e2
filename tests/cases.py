"""Shared parity cases: a mesh + equations configuration and a seeded state, evaluated by the
oracle (CPU) and by the device library on the same inputs."""
import numpy as np

from oracle_lib import Case, random_state, tgv_state

TWO_PI = 2 * np.pi

# (name, case, state kind)
RESIDUAL_CASES = [
    ("euler_tgv_hllc_bj", Case((8, 7, 6), lo=(0, 0, 0), hi=(TWO_PI,) * 3), "tgv"),
    ("ns_tgv_hllc_bj", Case((8, 7, 6), lo=(0, 0, 0), hi=(TWO_PI,) * 3, kind="navierstokes",
                            mu=1 / 1600), "tgv"),
    ("ns_tgv_hllc_vk", Case((6, 6, 6), lo=(0, 0, 0), hi=(TWO_PI,) * 3, kind="navierstokes",
                            mu=1 / 1600, limiter="venkatakrishnan", coefficient=1.0), "tgv"),
    ("euler_roe_nolim", Case((6, 5, 4), limiter=None, flux="roe"), "random"),
    ("euler_ausmd", Case((6, 5, 4), flux="ausmd"), "random"),
    ("euler_ausmdv", Case((6, 5, 4), flux="ausmdv"), "random"),
    ("euler_hll", Case((6, 5, 4), flux="hll"), "random"),
    ("euler_hlle", Case((6, 5, 4), flux="hlle"), "random"),
    ("euler_lxf", Case((6, 5, 4), flux="lxf"), "random"),
    ("euler_central", Case((6, 5, 4), flux="central"), "random"),
    ("euler_godunov", Case((6, 5, 4), flux="godunov"), "random"),
    ("euler_o1_hllc", Case((6, 5, 4), space_order=1), "random"),
    ("euler_walls", Case((6, 5, 4), periodic=(0, 0, 0), flux="hllc", bcs=[
        ("left", "supersonic inflow", (1.0, 2.0, 0.1, 0.0, 0.7)),
        ("right", "supersonic outflow", None),
        ("bottom", "symmetry", None), ("top", "slipwall", None),
        ("back", "subsonic inflow", (1.1, 0.1, 0.0, 0.2, 1.0)),
        ("front", "subsonic outflow", (1.0, 0.0, 0.0, 0.0, 0.9))]), "random"),
    ("euler_farfield", Case((5, 4, 4), periodic=(1, 0, 0), flux="roe", bcs=[
        ("bottom", "farfield", (1.0, 0.3, 0.1, 0.0, 0.8)),
        ("top", "pressure outflow", (1.0, 0.0, 0.0, 0.0, 0.9)),
        ("back", "farfield", (1.0, -0.2, 0.0, 0.1, 1.1)),
        ("front", "farfield", (1.0, 0.0, 0.0, -0.3, 1.0))]), "random"),
    ("ns_walls", Case((6, 5, 4), periodic=(1, 0, 1), kind="navierstokes", mu=0.01, bcs=[
        ("bottom", "adiabatic wall", None), ("top", "wall", None)]), "random"),
    ("ns_2d", Case((9, 7, 1), periodic=(1, 0, 0), kind="navierstokes", mu=0.02,
                   limiter="venkatakrishnan", coefficient=1.0, bcs=[
                       ("bottom", "wall", None), ("top", "symmetry", None)]), "random"),
    ("euler_1d", Case((12, 1, 1), periodic=(0, 0, 0), bcs=[
        ("left", "supersonic outflow", None), ("right", "supersonic outflow", None)]), "random"),
]

# meshes of many 256-cell device blocks (full and partial blocks, block faces shared between
# blocks, walls cutting through blocks)
LARGE_CASES = [
    ("ns_tgv_24", Case((24, 24, 24), lo=(0, 0, 0), hi=(TWO_PI,) * 3, kind="navierstokes",
                       mu=1 / 1600), "tgv"),
    ("euler_walls_17", Case((17, 13, 11), periodic=(0, 0, 0), flux="hllc", bcs=[
        ("left", "supersonic inflow", (1.0, 2.0, 0.1, 0.0, 0.7)),
        ("right", "supersonic outflow", None),
        ("bottom", "symmetry", None), ("top", "slipwall", None),
        ("back", "subsonic inflow", (1.1, 0.1, 0.0, 0.2, 1.0)),
        ("front", "subsonic outflow", (1.0, 0.0, 0.0, 0.0, 0.9))]), "random"),
]


def initial_state(case, arrays, kind, seed=7):
    nc, ni = arrays["num_cells"], arrays["num_inner"]
    if kind == "tgv":
        P = tgv_state(arrays["center"][:nc])
    else:
        P = random_state(nc, seed)
    P[ni:] = 0.0
    return np.ascontiguousarray(P)


def gpu_bcs(case):
    return {ent: (name, ctx) for ent, name, ctx in case.bcs}
