"""Closed-form scatter cases that pin roughness squaring and the degenerate-
direction rule of src/renderer/cpu.rs.

Every case hits the top of a unit sphere at the origin with the axial ray
(0,0,5) -> (0,0,-1): t = 4, P = (0,0,1), n = (0,0,1), front face.  The draw
words are hand-picked so that random_in_unit_sphere (utils/mod.rs:47-55,
normalize(U[-1,1]^3)) returns an axis vector exactly:
  U_X    = (0.75, 0.5, 0.5) as U[0,1) -> (0.5, 0, 0) -> (1, 0, 0)
  U_NEGZ = (0.5, 0.5, 0.0)            -> (0, 0, -1) -> (0, 0, -1)
so each expected direction is a closed form of the roughness:
  metal, roughness r:   normalize(reflect + r^2 * u)        cpu.rs:262-263, 276-277, 305-306
  roughness test:       diffuse iff u2 < r^2                cpu.rs:308-309
  glass, roughness r:   normalize(refract + r^2 * u)        cpu.rs:300-302
  |dir|^2 < 1e-10:      dir = n                             cpu.rs:326-328
A slip to plain `roughness` changes cases metal_r05, spec_r05_u030,
spec_r055_u031 and glass_r05; dropping the degenerate rule changes
degenerate_diffuse (the mutant-oracle test checks exactly that).
"""
import numpy as np

from raydar_amd import abi
from raydar_amd.scene import Scene, make_camera, make_object, sky_world


def U(x):
    """Draw word whose U[0,1) value ((w >> 8) * 2^-24) is x (x a multiple of 2^-24)."""
    return np.uint32(int(round(x * 2 ** 24)) << 8)


U_X = (U(0.75), U(0.5), U(0.5))
U_NEGZ = (U(0.5), U(0.5), U(0.0))
RAY = np.array([0, 0, 5, 0, 0, -1], np.float32)
P = np.array([0.0, 0.0, 1.0])
N = np.array([0.0, 0.0, 1.0])


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


# name: (material, draw words by slot, expected direction, origin offset along "n" or "dir")
CASES = {
    "metal_r05": (dict(roughness=0.5, metallic=1.0), {0: U(0.5), 1: U(0.5), 4: U_X},
                  _unit([0.25, 0, 1]), "n"),
    "diffuse_r06_u030": (dict(roughness=0.6), {0: U(0.5), 1: U(0.5), 2: U(0.3), 4: U_X},
                         np.array([1.0, 0, 1]), "n"),
    "spec_r05_u030": (dict(roughness=0.5), {0: U(0.5), 1: U(0.5), 2: U(0.3), 4: U_X},
                      _unit([0.25, 0, 1]), "n"),
    "spec_r055_u031": (dict(roughness=0.55), {0: U(0.5), 1: U(0.5), 2: U(0.31), 4: U_X},
                       _unit([np.float32(0.55) * np.float32(0.55), 0, 1]), "n"),
    "glass_r05": (dict(roughness=0.5, transmission=1.0, ior=1.5), {0: U(0.5), 1: U(0.999), 4: U_X},
                  _unit([0.25, 0, -1]), "dir"),
    "degenerate_diffuse": (dict(roughness=1.0), {0: U(0.5), 1: U(0.5), 2: U(0.5), 4: U_NEGZ},
                           N.copy(), "n"),
}

# cases a mutant oracle must get wrong
CAUGHT_BY = {
    "ORACLE_MUTANT_PLAIN_ROUGHNESS": ("metal_r05", "spec_r05_u030", "spec_r055_u031", "glass_r05"),
    "ORACLE_MUTANT_NO_DEGENERATE_RULE": ("degenerate_diffuse",),
}


def case(name):
    """(scene, packed n x 16 scatter input, bits n x 8, expected direction,
    expected origin) for one case, in the layout of oracle.scatter /
    raydar_amd.debug_scatter."""
    mat, slots, d_exp, along = CASES[name]
    m = dict(roughness=0.0, metallic=0.0, transmission=0.0, ior=1.5, albedo=(0.9, 0.8, 0.7))
    m.update(mat)
    sc = Scene(make_camera((0, 0, 10), (0, 0, 0), (0, 1, 0), 8, 8, fov_deg=40), sky_world(),
               np.array([make_object(abi.RD_SPHERE, (0, 0, 0), 1.0, **m)]))
    packed = np.zeros((1, 16), np.float32)
    packed[0, :6] = RAY
    packed[0, 6] = 4.0
    packed[0, 7] = np.uint32(0).view(np.float32)
    packed[0, 8:11] = 1.0
    bits = np.zeros((1, 8), np.uint32)
    for s, w in slots.items():
        if isinstance(w, tuple):
            bits[0, s:s + 3] = w
        else:
            bits[0, s] = w
    off = N if along == "n" else d_exp
    return sc, packed, bits, d_exp, P + off * 1e-4


def failures(scatter):
    """Names of the cases where scatter(scene, packed, bits) -> n x 16 misses
    the closed form (direction and origin within 1e-6)."""
    bad = []
    for name in CASES:
        sc, packed, bits, d_exp, o_exp = case(name)
        out = scatter(sc, packed, bits)
        if not (np.allclose(out[0, 3:6], d_exp, atol=1e-6, rtol=0)
                and np.allclose(out[0, 0:3], o_exp, atol=1e-6, rtol=0)):
            bad.append(name)
    return bad
