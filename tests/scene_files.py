"""Writers of small scene files (scene.cfg, OBJ/MTL key frames, camera.cfg,
light.cfg) in the formats the reference reads, for the scene-loading tests."""
import os

import numpy as np


def write_obj(path, verts, faces, normals=None, mtllib=None, groups=None, extra=(), objects=()):
    """verts (nv, 3); faces: list of corner lists, each corner an int (1-based
    vertex) or (v, n) pair written as v//n; groups: list of (first_face, material);
    objects: face indices before which an "o <k>" line starts a new object."""
    lines = []
    if mtllib:
        lines.append(f"mtllib {mtllib}")
    for v in verts:
        lines.append("v %.9g %.9g %.9g" % tuple(v))
    if normals is not None:
        for n in normals:
            lines.append("vn %.9g %.9g %.9g" % tuple(n))
    lines.extend(extra)
    starts = dict(groups or [])
    obj_at = {}
    for k, f0 in enumerate(objects):
        obj_at.setdefault(f0, []).append(k)
    for i, f in enumerate(faces):
        for k in obj_at.get(i, ()):
            lines.append(f"o {k}")
        if i in starts:
            lines.append(f"usemtl {starts[i]}")
        corners = []
        for c in f:
            corners.append(f"{c[0]}//{c[1]}" if isinstance(c, tuple) else str(c))
        lines.append("f " + " ".join(corners))
    for k in obj_at.get(len(faces), ()):
        lines.append(f"o {k}")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_mtl(path, materials):
    """materials: dict name -> dict of MTL keys (Kd, Ks, Ke, Ns, Ni, illum)."""
    lines = []
    for name, m in materials.items():
        lines.append(f"newmtl {name}")
        for k, v in m.items():
            if isinstance(v, (list, tuple)):
                lines.append(f"{k} " + " ".join("%.9g" % x for x in v))
            else:
                lines.append(f"{k} {v}")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_camera(path, pos, orient, up, fov=None, res=None, rotation=None):
    lines = ["# camera", "position %.9g %.9g %.9g" % tuple(pos),
             "orientation %.9g %.9g %.9g" % tuple(orient), "up %.9g %.9g %.9g" % tuple(up)]
    if rotation is not None:
        lines.append("rotation %.9g %.9g %.9g" % tuple(rotation))
    if res is not None:
        lines += [f"resX {res[0]}", f"resY {res[1]}"]
    if fov is not None:
        lines.append(f"FOV {fov}")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_lights(path, lights, blank_lines=False):
    """lights: list of dicts with position, intensity, edge1, edge2 and optional normal."""
    lines = []
    for L in lights:
        for k in ("position", "normal", "intensity", "edge1", "edge2"):
            if k in L:
                lines.append(f"{k} " + " ".join("%.9g" % x for x in L[k]))
        if blank_lines:
            lines.append("")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_cfg(path, entries):
    """entries: list of (key, value) lines, or raw strings."""
    lines = [e if isinstance(e, str) else f"{e[0]} = {e[1]}" for e in entries]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def heightfield(q, phase):
    """(q+1)^2 vertices over [-1,1]^2 in x/z with a smooth height field, its
    analytic unit normals, and 2*q*q triangles as corner lists (v, n)."""
    xs = np.linspace(-1.0, 1.0, q + 1)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    H = 0.25 * np.sin(2.0 * X + phase) * np.cos(3.0 * Z) + 0.1 * X * Z
    dHdx = 0.5 * np.cos(2.0 * X + phase) * np.cos(3.0 * Z) + 0.1 * Z
    dHdz = -0.75 * np.sin(2.0 * X + phase) * np.sin(3.0 * Z) + 0.1 * X
    verts = np.stack([X.ravel(), H.ravel(), Z.ravel()], 1).astype(np.float32)
    n = np.stack([-dHdx.ravel(), np.ones(X.size), -dHdz.ravel()], 1)
    n = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
    faces = []
    for i in range(q):
        for j in range(q):
            a = i * (q + 1) + j + 1
            b = a + q + 1
            faces.append([(a, a), (b, b), (b + 1, b + 1)])
            faces.append([(a, a), (b + 1, b + 1), (a + 1, a + 1)])
    return verts, n, faces


def animated_heightfield_scene(root, q=24, frames=3, step=0.4, res=(96, 64)):
    """An animated scene: key frames anim/hf<i>.obj with per-vertex normals and
    two materials, a camera.cfg and a light.cfg.  Returns the scene.cfg path."""
    os.makedirs(os.path.join(root, "anim"), exist_ok=True)
    write_mtl(os.path.join(root, "anim", "hf.mtl"),
              {"matte": {"Kd": (0.7, 0.6, 0.5)},
               "gloss": {"Kd": (0.2, 0.3, 0.6), "Ks": (0.5, 0.5, 0.5), "Ns": 30}})
    digits = len(str(frames - 1)) if frames > 1 else 0
    for f in range(frames):
        v, n, faces = heightfield(q, 0.7 * f)
        name = "hf" + str(f).zfill(digits) + ".obj"
        write_obj(os.path.join(root, "anim", name), v, faces, normals=n, mtllib="hf.mtl",
                  groups=[(0, "matte"), (len(faces) // 2, "gloss")])
    write_camera(os.path.join(root, "camera.cfg"), (0.0, 1.6, 2.6), (0.0, -0.5, -1.0),
                 (0.0, 1.0, 0.0), fov=55, res=res)
    write_lights(os.path.join(root, "light.cfg"),
                 [{"position": (-0.4, 1.5, -0.4), "intensity": (6.0, 6.0, 5.5),
                   "edge1": (0.8, 0.0, 0.0), "edge2": (0.0, 0.0, 0.8)}])
    cfg = os.path.join(root, "scene.cfg")
    write_cfg(cfg, ["# animated height field", ("frame prefix", "anim/hf"), ("frame suffix", ".obj"),
                    ("number of frames", frames), ("frame step", step),
                    ("camera", "camera.cfg"), ("light", "light.cfg")])
    return cfg


def write_instances(path, objects, matrices, aabb_min=None, aabb_max=None, comments=True):
    """The instances file of WFObject::loadInstances: per instance an obj_id
    line, m_row_0..2 (world-to-object rows with the translation last) and the
    AABB_min / AABB_max lines (omitted where aabb_min / aabb_max is None)."""
    lines = []
    for i, k in enumerate(objects):
        if comments:
            lines.append("#new instance#")
        lines.append(f"obj_id {int(k)}")
        for r in range(3):
            lines.append(f"m_row_{r} " + " ".join("%.9g" % x for x in matrices[i][r]))
        if aabb_min is not None:
            lines.append("AABB_min " + " ".join("%.9g" % x for x in aabb_min[i]))
        if aabb_max is not None:
            lines.append("AABB_max " + " ".join("%.9g" % x for x in aabb_max[i]))
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def instanced_scene_files(root, sc, ins):
    """scene.cfg + one OBJ whose objects are the Instancing's triangle ranges
    ("o" lines), the instances file and a camera.cfg.  Returns the scene.cfg path."""
    faces = [[int(a) + 1, int(b) + 1, int(c) + 1] for a, b, c in sc.idx]
    write_obj(os.path.join(root, "inst.obj"), sc.verts.T, faces, objects=list(ins.obj_first[:-1]))
    write_instances(os.path.join(root, "instances.txt"), ins.objects, ins.matrices, ins.aabb_min,
                    ins.aabb_max)
    c = sc.camera
    write_camera(os.path.join(root, "camera.cfg"), c["pos"], c["dir"], c["up"], fov=c["fov"], res=c["res"])
    cfg = os.path.join(root, "scene.cfg")
    write_cfg(cfg, [("obj", "inst.obj"), ("camera", "camera.cfg"), ("instances", "instances.txt")])
    return cfg
