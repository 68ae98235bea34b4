"""Closed, outward-oriented triangle meshes for rigid boundaries, and an
OBJ reader for mesh files such as the reference's applications/*/meshes."""
from __future__ import annotations

from pathlib import Path

import numpy as np


def _tube_faces(nu: int, nv: int) -> np.ndarray:
    """Faces of an nu x nv periodic quad grid (u major), two triangles per
    quad, oriented so that the normal points along d(v) x d(u) (outward for
    the tube parametrisation below)."""
    u = np.arange(nu)[:, None]
    v = np.arange(nv)[None, :]
    a = (u * nv + v).ravel()
    b = (((u + 1) % nu) * nv + v).ravel()
    c = (((u + 1) % nu) * nv + (v + 1) % nv).ravel()
    d = (u * nv + (v + 1) % nv).ravel()
    return np.concatenate([np.stack([a, c, b], 1), np.stack([a, d, c], 1)]).astype(np.uint32)


def torus_knot(p: int = 2, q: int = 3, nu: int = 1024, nv: int = 32, radius: float = 0.2,
               scale: float = 0.3, center=(0.0, 0.0, 0.0)):
    """Tube of the given radius around the (p, q) torus knot
    ((2 + cos(q t)) cos(p t), (2 + cos(q t)) sin(p t), sin(q t)) * scale,
    with parallel-transport frames; 2*nu*nv triangles."""
    t = np.linspace(0.0, 2.0 * np.pi, nu, endpoint=False)
    c = np.stack([(2 + np.cos(q * t)) * np.cos(p * t), (2 + np.cos(q * t)) * np.sin(p * t),
                  np.sin(q * t)], 1) * scale
    tang = np.roll(c, -1, 0) - np.roll(c, 1, 0)
    tang /= np.linalg.norm(tang, axis=1, keepdims=True)
    # parallel transport of a normal around the closed curve, then spread the
    # holonomy twist evenly so the frame closes
    n = np.zeros_like(c)
    ref = np.array([0.0, 0.0, 1.0]) if abs(tang[0, 2]) < 0.9 else np.array([1.0, 0.0, 0.0])
    n[0] = np.cross(tang[0], ref)
    n[0] /= np.linalg.norm(n[0])
    for i in range(1, nu):
        v = n[i - 1] - np.dot(n[i - 1], tang[i]) * tang[i]
        n[i] = v / np.linalg.norm(v)
    v = n[-1] - np.dot(n[-1], tang[0]) * tang[0]
    v /= np.linalg.norm(v)
    b0 = np.cross(tang[0], n[0])
    twist = np.arctan2(np.dot(v, b0), np.dot(v, n[0]))
    ang = -twist * np.arange(nu) / nu
    b = np.cross(tang, n)
    n2 = np.cos(ang)[:, None] * n + np.sin(ang)[:, None] * b
    b2 = np.cross(tang, n2)
    phi = np.linspace(0.0, 2.0 * np.pi, nv, endpoint=False)
    verts = (c[:, None, :] + radius * (np.cos(phi)[None, :, None] * n2[:, None, :] +
                                       np.sin(phi)[None, :, None] * b2[:, None, :]))
    verts = verts.reshape(-1, 3) + np.asarray(center, np.float64)
    return verts, _tube_faces(nu, nv)


def uv_sphere(radius: float = 0.5, nu: int = 64, nv: int = 32, center=(0.0, 0.0, 0.0)):
    """UV sphere with poles, outward oriented."""
    verts = [np.array([0.0, 0.0, radius])]
    for j in range(1, nv):
        th = np.pi * j / nv
        for i in range(nu):
            ph = 2 * np.pi * i / nu
            verts.append(radius * np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph),
                                            np.cos(th)]))
    verts.append(np.array([0.0, 0.0, -radius]))
    verts = np.array(verts) + np.asarray(center, np.float64)
    faces = []
    for i in range(nu):
        faces.append([0, 1 + i, 1 + (i + 1) % nu])
    for j in range(nv - 2):
        r0, r1 = 1 + j * nu, 1 + (j + 1) * nu
        for i in range(nu):
            a, b = r0 + i, r0 + (i + 1) % nu
            c, d = r1 + i, r1 + (i + 1) % nu
            faces.append([a, c, d])
            faces.append([a, d, b])
    last = len(verts) - 1
    r0 = 1 + (nv - 2) * nu
    for i in range(nu):
        faces.append([last, r0 + (i + 1) % nu, r0 + i])
    return verts, np.array(faces, np.uint32)


def superellipsoid(a=0.4, b=0.3, c=0.25, e1=0.5, e2=0.5, nu=64, nv=32, center=(0.0, 0.0, 0.0)):
    """Superellipsoid |x/a|^(2/e2)+|y/b|^(2/e2))^(e2/e1) + |z/c|^(2/e1) = 1 by
    the signed-power map of a UV sphere (same connectivity)."""
    verts, faces = uv_sphere(1.0, nu, nv)
    th = np.arccos(np.clip(verts[:, 2], -1, 1))
    ph = np.arctan2(verts[:, 1], verts[:, 0])

    def spow(x, e):
        return np.sign(x) * np.abs(x) ** e

    x = a * spow(np.sin(th), e1) * spow(np.cos(ph), e2)
    y = b * spow(np.sin(th), e1) * spow(np.sin(ph), e2)
    z = c * spow(np.cos(th), e1)
    return np.stack([x, y, z], 1) + np.asarray(center, np.float64), faces


def load_obj(path) -> tuple[np.ndarray, np.ndarray]:
    """Vertices and triangles of a Wavefront OBJ file (polygons fanned)."""
    verts, faces = [], []
    for line in Path(path).read_text().splitlines():
        parts = line.split()
        if not parts:
            continue
        if parts[0] == "v":
            verts.append([float(x) for x in parts[1:4]])
        elif parts[0] == "f":
            idx = [int(tok.split("/")[0]) for tok in parts[1:]]
            idx = [i - 1 if i > 0 else len(verts) + i for i in idx]
            for k in range(1, len(idx) - 1):
                faces.append([idx[0], idx[k], idx[k + 1]])
    return np.array(verts, np.float64), np.array(faces, np.uint32)


def signed_volume(verts, faces) -> float:
    """Enclosed volume (positive for an outward-oriented closed mesh)."""
    a, b, c = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)
