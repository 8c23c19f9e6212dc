"""FP64 CPU restatement of the bounce stage (shading, scatter, sky, resolve).

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of femrt_frame_scatter
and femrt_frame_resolve.

Follows FEM-AABB/raytracer.cu:21-98 (scatter: Lambertian, metal, dielectric),
362-406 (ray_color: attenuation product, emission, sky, depth cap) and
460-475 (mean, gamma sqrt, clamp, 8 bits), with the library's documented
choices: counter-based random numbers keyed by (seed, path, depth, draw),
direct unit-vector sampling, the hit face's tangent cross product for the
normal (outward by the element's orientation), material 1 on local face 0.
"""
from __future__ import annotations

import numpy as np

M64 = (1 << 64) - 1


def _mix(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def path_rand(seed, path, depth, k):
    h = _mix(seed ^ _mix((path * 0x9E3779B97F4A7C15 + ((depth << 8) | k)) & M64))
    return float(np.float32(h >> 40) * np.float32(1.0 / 16777216.0))


def unit_vec(a, b):
    z = 1.0 - 2.0 * a
    phi = 2.0 * np.pi * b
    r = np.sqrt(max(0.0, 1.0 - z * z))
    return np.array([r * np.cos(phi), r * np.sin(phi), z])


FACE_FIX = [2, 1, 0, 1, 0, 2]
FACE_VAL = [-1, -1, 1, 1, -1, 1]


def sky(d):
    t = 0.5 * (d[1] + 1.0)
    return np.array([(1 - t) + t * 0.5, (1 - t) + t * 0.7, 1.0])


def scatter(orc, rays, hits, mats, seed, depth, last, tp, acc, alive, ties=None, ids=None):
    """One bounce of every live path; rays (6 x n) are the rays that were
    traced, hits the oracle's trace of them.  Updates tp, acc, alive in place
    and returns the next rays (6 x n, float64).  ties (bool n, optional) is
    set where a dielectric's reflect/refract decision lies within 1e-5 of its
    threshold (Fresnel draw or total internal reflection), which FP32 may
    decide the other way.  ids (int n, optional) are the frame's path
    indices of the columns when they are a subset of a frame (the random
    draws are keyed by path index); by default column p is path p."""
    n = rays.shape[1]
    nxt = np.zeros((6, n))
    nxt[:3] = 1e30
    nxt[3] = 1.0
    for p in range(n):
        if not alive[p]:
            continue
        pid = int(ids[p]) if ids is not None else p
        o = rays[:3, p].astype(np.float64)
        d = rays[3:, p].astype(np.float64)
        e = int(hits["elem"][p])
        if e < 0:
            acc[:, p] += tp[:, p] * sky(d)
            alive[p] = False
            continue
        f = int(hits["face"][p])
        t = float(hits["t"][p])
        _, J = orc.map(e, hits["rst"][:, p])
        fix = FACE_FIX[f]
        a, b = (fix + 1) % 3, (fix + 2) % 3
        nrm = np.cross(J[:, a], J[:, b])
        nrm /= np.linalg.norm(nrm)
        if FACE_VAL[f] < 0:
            nrm = -nrm
        front = np.dot(d, nrm) < 0
        if not front:
            nrm = -nrm
        po = o + t * d
        mi = (1 if f == 0 else 0) if len(mats) > 1 else 0
        ty, col, em, rough, ior = mats[mi]
        att = np.array(col, float)
        em = np.array(em, float)
        scattered = True
        if ty == 1:
            rv = d - 2 * np.dot(d, nrm) * nrm
            us = unit_vec(path_rand(seed, pid, depth + 1, 0), path_rand(seed, pid, depth + 1, 1))
            rad = np.cbrt(path_rand(seed, pid, depth + 1, 2))
            sd = rv + rough * rad * us
            scattered = np.dot(sd, nrm) > 0
        elif ty == 2:
            att = np.ones(3)
            eta = 1.0 / ior if front else ior
            cosv = min(-np.dot(d, nrm), 1.0)
            sinv = np.sqrt(max(0.0, 1 - cosv * cosv))
            r0 = ((1 - eta) / (1 + eta)) ** 2
            sch = r0 + (1 - r0) * (1 - cosv) ** 5
            xi = path_rand(seed, pid, depth + 1, 0)
            if ties is not None and (abs(eta * sinv - 1) < 1e-5 or abs(sch - xi) < 1e-5):
                ties[p] = True
            if eta * sinv > 1 or sch > xi:
                sd = d + 2 * cosv * nrm
            else:
                k = 1 - eta * eta * (1 - cosv * cosv)
                sd = eta * (d + cosv * nrm) - np.sqrt(max(k, 0.0)) * nrm
        else:
            us = unit_vec(path_rand(seed, pid, depth + 1, 0), path_rand(seed, pid, depth + 1, 1))
            sd = nrm + us
            if np.dot(sd, sd) < 1e-12:
                sd = nrm
        if not scattered:
            acc[:, p] += tp[:, p] * em
            alive[p] = False
            continue
        tp[:, p] *= att
        if np.any(em != 0) or last:
            if np.any(em != 0):
                acc[:, p] += tp[:, p] * em
            alive[p] = False
            continue
        nxt[:3, p] = po
        nxt[3:, p] = sd / np.linalg.norm(sd)
    return nxt


def resolve(acc, W, H, spp):
    c = acc.reshape(3, W * H, spp).mean(axis=2)
    g = np.minimum(1.0, np.sqrt(np.maximum(c, 0.0)))
    px = (255.99 * g).astype(np.uint32)
    return px[0] | (px[1] << 8) | (px[2] << 16) | np.uint32(0xFF000000)
