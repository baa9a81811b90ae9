"""Seeded ray sets for the closest-hit parity tests.

Each generator returns n x 6 float32 rays (origin, direction) chosen to hit
the cases where a BVH can disagree with raydar's linear scan
(src/renderer/cpu.rs:344-352): camera rays, secondary rays leaving surfaces
(the reference's 1e-4 offset, cpu.rs:316-325), rays grazing primitive
silhouettes (where the FP32 sphere test of cpu.rs:39-52 is ill-conditioned),
axis-parallel rays (zero direction components), and far origins.
"""
import numpy as np

import oracle


def camera(scene, n, rng):
    pix = rng.integers(0, scene.width * scene.height, n).astype(np.uint32)
    return oracle.camera_rays(scene, pix)


def secondary(scene, n, rng, unnormalized=True):
    rays = camera(scene, n, rng)
    ids, ts, nrm, _ = oracle.trace(scene, rays)
    h = ids >= 0
    p = rays[h, :3] + rays[h, 3:] * ts[h, None]
    o = p + nrm[h] * np.float32(1e-4)
    d = rng.normal(size=(int(h.sum()), 3))
    if unnormalized:  # diffuse directions are n + unit vector, never normalised (cpu.rs:267-271)
        d = nrm[h] + d / np.linalg.norm(d, axis=1)[:, None]
    return np.concatenate([o, d], 1).astype(np.float32)


def grazing(scene, n, rng):
    objs = scene.objects
    k = rng.integers(0, len(objs), n)
    c = objs["center"][k].astype(np.float64)
    s = objs["size"][k].astype(np.float64)
    o = c + rng.normal(size=(n, 3)) * np.maximum(s, 1.0)[:, None] * rng.choice([3, 30, 300], n)[:, None]
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    target = c + u * s[:, None] * (1.0 + rng.normal(size=n)[:, None] * 1e-3)
    d = target - o
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.concatenate([o, d], 1).astype(np.float32)


def axis_parallel(scene, n, rng, extent=12.0):
    o = rng.uniform(-extent, extent, (n, 3))
    o[:, 1] = rng.uniform(0, 3, n)
    d = np.zeros((n, 3))
    ax = rng.integers(0, 3, n)
    d[np.arange(n), ax] = rng.choice([-1.0, 1.0], n)
    return np.concatenate([o, d], 1).astype(np.float32)


def far(scene, n, rng, dist=3000.0):
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    o = u * dist
    objs = scene.objects
    k = rng.integers(0, len(objs), n)
    d = objs["center"][k].astype(np.float64) - o + rng.normal(size=(n, 3)) * 0.5
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.concatenate([o, d], 1).astype(np.float32)


ALL = (camera, secondary, grazing, axis_parallel, far)


def mixed(scene, n_each, seed=0):
    rng = np.random.default_rng(seed)
    return np.concatenate([g(scene, n_each, rng) for g in ALL], 0)
