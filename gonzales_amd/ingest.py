"""Triangle-mesh ingest from pbrt scene files: the shapes the hot path renders, as flat buffers.

Follows the reference's scene description for what reaches the accelerator
(Api/SceneDescription.swift:317-399, Shape/Triangle.swift:560-647, Shape/PlyMesh.swift):
  - `Shape "trianglemesh"` ("point3 P", "integer indices", "normal N", "point2 uv",
    "integer faceIndices") and `Shape "plymesh"` ("string filename", relative to the scene file,
    `.gz` decompressed; read by gz_ply_parse) become meshes whose points are transformed to world
    space by the current transform at the Shape (Transform * Point, Transform.swift:52-66:
    row sums left to right, then divided by the homogeneous weight, in float32);
  - each non-light Shape appends its own material (the current named or anonymous material) to
    the material table, so a primitive's material index is its shape's position among the
    non-light shapes (SceneDescription.swift:331-340);
  - inside an `AreaLightSource "diffuse"` attribute, a shape's triangles become area lights of
    radiance "rgb L", numbered in order (one light per triangle);
  - transforms: Identity, Translate, Scale, Rotate, LookAt, Transform, ConcatTransform,
    AttributeBegin/End and TransformBegin/End; WorldBegin resets the transform; Camera records
    the world-to-camera transform;
  - Film x/y resolution, Sampler pixelsamples, PixelFilter gaussian/box radius and sigma,
    Integrator maxdepth, Camera perspective fov are read for the render settings.
Other shapes (spheres, disks, curves, ...) and directives are skipped: the hot path renders
triangles. Transforms are composed in float64 and rounded to float32 once (the reference
composes in float32, so composed transforms can differ in the last bit; identity, the case of
every scene in the reference's repository, is exact).
Materials: diffuse, conductor, dielectric and coatedconductor with the parameters and defaults
of Material/{Diffuse,Conductor,Dielectric,CoatedConductor}.swift; rgb values only (a named
spectrum other than the copper defaults is rejected).
"""
from __future__ import annotations

import ctypes
import gzip
import os
import re
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .accelerator import Primitives, Scene, TriangleMesh

_TOKEN = re.compile(r'"[^"]*"|\[|\]|[^\s\[\]"#]+|#[^\n]*')


def tokenize(text: str) -> list[str]:
    return [t for t in _TOKEN.findall(text) if not t.startswith("#")]


@dataclass
class Param:
    type: str
    name: str
    values: list


def _parse_value(tok: str):
    if tok.startswith('"'):
        return tok[1:-1]
    if tok in ("true", "false"):
        return tok == "true"
    return float(tok)


class _Tokens:
    def __init__(self, toks):
        self.toks = toks
        self.i = 0

    def peek(self):
        return self.toks[self.i] if self.i < len(self.toks) else None

    def next(self):
        t = self.toks[self.i]
        self.i += 1
        return t

    def numbers(self, n=None):
        """n numbers, bracketed or not (for n None: a bracketed list)."""
        out = []
        if self.peek() == "[":
            self.next()
            while self.peek() != "]":
                out.append(float(self.next()))
            self.next()
        else:
            for _ in range(n or 0):
                out.append(float(self.next()))
        if n is not None and len(out) != n:
            raise ValueError(f"expected {n} numbers, got {len(out)}")
        return out

    def params(self) -> dict:
        """The parameter list after a directive: "type name" value-or-[values] pairs."""
        out = {}
        while self.peek() is not None and self.peek().startswith('"'):
            decl = self.next()[1:-1].split()
            if len(decl) != 2:
                raise ValueError(f"bad parameter declaration {decl}")
            if self.peek() == "[":
                self.next()
                vals = []
                while self.peek() != "]":
                    vals.append(_parse_value(self.next()))
                self.next()
            else:
                vals = [_parse_value(self.next())]
            out[decl[1]] = Param(decl[0], decl[1], vals)
        return out


def _translate(x, y, z):
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m


def _scale(x, y, z):
    return np.diag([x, y, z, 1.0])


def _rotate(angle, x, y, z):
    a = np.asarray([x, y, z], np.float64)
    a /= np.linalg.norm(a)
    s, c = np.sin(np.radians(angle)), np.cos(np.radians(angle))
    m = np.eye(4)
    m[0, :3] = (a[0] * a[0] + (1 - a[0] * a[0]) * c, a[0] * a[1] * (1 - c) - a[2] * s, a[0] * a[2] * (1 - c) + a[1] * s)
    m[1, :3] = (a[0] * a[1] * (1 - c) + a[2] * s, a[1] * a[1] + (1 - a[1] * a[1]) * c, a[1] * a[2] * (1 - c) - a[0] * s)
    m[2, :3] = (a[0] * a[2] * (1 - c) - a[1] * s, a[1] * a[2] * (1 - c) + a[0] * s, a[2] * a[2] + (1 - a[2] * a[2]) * c)
    return m


def transform_points(m: np.ndarray, p: np.ndarray) -> np.ndarray:
    """Transform * Point (Transform.swift:52-66) in float32: each row summed left to right,
    then divided by the homogeneous weight."""
    m = np.asarray(m, np.float32)
    p = np.asarray(p, np.float32).reshape(-1, 3)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    rows = [((m[r, 0] * x + m[r, 1] * y) + m[r, 2] * z) + m[r, 3] for r in range(4)]
    w = rows[3]
    return np.stack([rows[0] / w, rows[1] / w, rows[2] / w], 1).astype(np.float32)


def read_ply(path: str) -> dict:
    """A binary PLY file (gzip-compressed when it ends in .gz) through gz_ply_parse."""
    with open(path, "rb") as f:
        data = f.read()
    if path.endswith(".gz"):
        data = gzip.decompress(data)
    buf = np.frombuffer(data, np.uint8)
    lib = abi.load()
    info = abi.GzPlyInfo()
    abi.check(lib.gz_ply_parse(buf.ctypes.data, len(buf), ctypes.byref(info), None, None, None, None, None),
              "gz_ply_parse")
    v = info.vertexCount
    out = {"P": np.zeros((v, 3), np.float32), "N": np.zeros((v, 3), np.float32), "uv": np.zeros((v, 2), np.float32),
           "indices": np.zeros(info.indexCount, np.int64), "faceIndices": np.zeros(info.faceCount, np.int64)}
    abi.check(lib.gz_ply_parse(buf.ctypes.data, len(buf), ctypes.byref(info), out["P"].ctypes.data,
                               out["N"].ctypes.data, out["uv"].ctypes.data, out["indices"].ctypes.data,
                               out["faceIndices"].ctypes.data), "gz_ply_parse")
    if not info.hasNormals:
        out["N"] = np.zeros((0, 3), np.float32)
    if not info.hasUvs:
        out["uv"] = np.zeros((0, 2), np.float32)
    if not info.hasFaceIndices:
        out["faceIndices"] = np.zeros(0, np.int64)
    return out


def _rgb(params, name, default):
    p = params.get(name)
    if p is None:
        return tuple(default)
    if p.type in ("rgb", "color") and len(p.values) == 3:
        return tuple(float(v) for v in p.values)
    if p.type == "spectrum" and len(p.values) == 1 and isinstance(p.values[0], str):
        raise ValueError(f"named spectrum {p.values[0]} is not supported (rgb values only)")
    raise ValueError(f"parameter {name}: expected rgb, got {p.type} {p.values}")


def _real(params, name, default=None):
    p = params.get(name)
    return float(p.values[0]) if p is not None else default


def make_material(kind: str, params: dict) -> np.ndarray:
    """GzMaterial of a pbrt material (Material/*.swift create functions and their defaults)."""
    from .integrator import (CU_ETA, CU_K, coated_conductor_material, conductor_material, dielectric_material,
                             diffuse_material)
    if kind == "diffuse":
        return diffuse_material(_rgb(params, "reflectance", (0.5, 0.5, 0.5)))
    if kind == "conductor":
        r = _real(params, "roughness")
        if r is None:
            r = _real(params, "uroughness", 0.5)  # the kernel's conductor is isotropic
        return conductor_material(_rgb(params, "eta", CU_ETA), _rgb(params, "k", CU_K), roughness=r)
    if kind == "dielectric":
        r = _real(params, "roughness")
        if r is None:
            r = _real(params, "uroughness", 0.0)
        remap = params["remaproughness"].values[0] if "remaproughness" in params else True
        return dielectric_material(_real(params, "eta", 1.5), roughness=r, remap=bool(remap))
    if kind == "coatedconductor":
        ri = _real(params, "interface.roughness", _real(params, "roughness"))
        rc = _real(params, "conductor.roughness", _real(params, "roughness"))
        ri = _real(params, "uroughness", 0.5) if ri is None else ri
        rc = _real(params, "uroughness", 0.5) if rc is None else rc
        remap = params["remaproughness"].values[0] if "remaproughness" in params else True
        return coated_conductor_material(
            reflectance=_rgb(params, "reflectance", (1.0, 1.0, 1.0)), interface_roughness=ri, conductor_roughness=rc,
            eta=_real(params, "eta", 1.5), thickness=_real(params, "thickness", 0.01),
            albedo=_rgb(params, "albedo", (0.0, 0.0, 0.0)), g=_real(params, "g", 0.0),
            conductor_eta=_rgb(params, "conductor.eta", CU_ETA), conductor_k=_rgb(params, "conductor.k", CU_K),
            max_depth=int(_real(params, "maxdepth", 10)), n_samples=int(_real(params, "nsamples", 1)),
            remap=bool(remap))
    raise ValueError(f"material type {kind!r} is not supported by the MIS kernel")


def material_c_table(gz: np.ndarray) -> np.ndarray:
    """The Material_C table the wavefront diffuse shade reads, as prepareMaterials builds it
    (BoundingHierarchy.swift:91-150): diffuse -> its reflectance, conductor and coated conductor
    -> 0.7 grey, dielectric -> 0.9 grey (all shaded as diffuse there)."""
    out = np.zeros(len(gz), abi.MATERIAL)
    out["type"] = 1
    out["albedo"] = 0.5
    t = gz["type"]
    out["albedo"][t == abi.GZ_MAT_DIFFUSE] = gz["reflectance"][t == abi.GZ_MAT_DIFFUSE]
    out["albedo"][(t == abi.GZ_MAT_CONDUCTOR) | (t == abi.GZ_MAT_COATED_CONDUCTOR)] = 0.7
    out["albedo"][t == abi.GZ_MAT_DIELECTRIC] = 0.9
    return out


@dataclass
class IngestedScene:
    scene: Scene
    prims: Primitives
    materials: np.ndarray            # GzMaterial table (one entry per non-light shape)
    world_to_camera: np.ndarray | None
    fov: float = 90.0
    resolution: tuple = (1280, 720)
    spp: int = 16
    max_depth: int = 5
    filter: dict = field(default_factory=lambda: {"type": "gaussian", "radius": (1.5, 1.5), "sigma": 0.5})
    skipped_shapes: list = field(default_factory=list)

    def camera(self, **kw):
        """The scene's PerspectiveCamera (film, spp and filter as the file sets them)."""
        from .camera import Film, PerspectiveCamera
        f = self.filter
        ftype = abi.GZ_FILTER_GAUSSIAN if f["type"] == "gaussian" else abi.GZ_FILTER_BOX
        return PerspectiveCamera(self.world_to_camera, self.fov, Film(*self.resolution), self.spp,
                                 filter_type=ftype, filter_radius=f["radius"], filter_sigma=f["sigma"], **kw)


def ingest(path: str) -> IngestedScene:
    """Reads a pbrt scene file's triangle shapes, materials, area lights and render settings."""
    with open(path) as f:
        tk = _Tokens(tokenize(f.read()))
    base = os.path.dirname(os.path.abspath(path))
    ctm = np.eye(4)
    stack = []
    named: dict[str, np.ndarray] = {}
    current = make_material("diffuse", {})
    area_L = None
    meshes, parts, mats, emission = [], [], [], []
    out = IngestedScene(None, None, None, None)
    n_lights = 0
    while tk.peek() is not None:
        d = tk.next()
        if d == "Identity":
            ctm = np.eye(4)
        elif d == "Translate":
            ctm = ctm @ _translate(*tk.numbers(3))
        elif d == "Scale":
            ctm = ctm @ _scale(*tk.numbers(3))
        elif d == "Rotate":
            ctm = ctm @ _rotate(*tk.numbers(4))
        elif d == "LookAt":
            from .camera import look_at
            v = tk.numbers(9)
            ctm = ctm @ look_at(v[0:3], v[3:6], v[6:9])
        elif d == "Transform":
            ctm = np.asarray(tk.numbers(16), np.float64).reshape(4, 4).T
        elif d == "ConcatTransform":
            ctm = ctm @ np.asarray(tk.numbers(16), np.float64).reshape(4, 4).T
        elif d in ("AttributeBegin", "TransformBegin"):
            stack.append((ctm.copy(), current, area_L))
        elif d in ("AttributeEnd", "TransformEnd"):
            ctm, cur, al = stack.pop()
            if d == "AttributeEnd":
                current, area_L = cur, al
        elif d == "WorldBegin":
            ctm = np.eye(4)
        elif d == "WorldEnd":
            pass
        elif d in ("ObjectBegin", "ObjectEnd", "ObjectInstance", "Include", "Import"):
            raise ValueError(f"{d}: object instancing and included files are not supported")
        elif d == "Camera":
            kind = tk.next()[1:-1]
            p = tk.params()
            if kind != "perspective":
                raise ValueError(f"camera {kind!r}: only the perspective camera is supported")
            out.world_to_camera = ctm.copy()
            out.fov = _real(p, "fov", 90.0)
        elif d == "Film":
            tk.next()
            p = tk.params()
            out.resolution = (int(_real(p, "xresolution", 1280)), int(_real(p, "yresolution", 720)))
        elif d == "Sampler":
            tk.next()
            out.spp = int(_real(tk.params(), "pixelsamples", 16))
        elif d == "Integrator":
            tk.next()
            out.max_depth = int(_real(tk.params(), "maxdepth", 5))
        elif d == "PixelFilter":
            kind = tk.next()[1:-1]
            p = tk.params()
            if kind == "gaussian":
                out.filter = {"type": "gaussian", "radius": (_real(p, "xradius", 1.5), _real(p, "yradius", 1.5)),
                              "sigma": _real(p, "sigma", 0.5)}
            elif kind == "box":
                out.filter = {"type": "box", "radius": (_real(p, "xradius", 0.5), _real(p, "yradius", 0.5)),
                              "sigma": 0.5}
            else:
                raise ValueError(f"pixel filter {kind!r}: gaussian and box are supported")
        elif d == "MakeNamedMaterial":
            name = tk.next()[1:-1]
            p = tk.params()
            named[name] = make_material(p["type"].values[0] if "type" in p else "diffuse", p)
        elif d == "NamedMaterial":
            name = tk.next()[1:-1]
            if name not in named:
                raise ValueError(f"unknown named material {name!r}")
            current = named[name]
        elif d == "Material":
            kind = tk.next()[1:-1]
            current = make_material(kind, tk.params())
        elif d == "AreaLightSource":
            kind = tk.next()[1:-1]
            p = tk.params()
            if kind != "diffuse":
                raise ValueError(f"area light {kind!r}: only diffuse area lights are supported")
            area_L = _rgb(p, "L", (1.0, 1.0, 1.0))
        elif d == "Shape":
            kind = tk.next()[1:-1]
            p = tk.params()
            if kind == "trianglemesh":
                data = {"P": np.asarray(p["P"].values, np.float32).reshape(-1, 3),
                        "indices": np.asarray(p["indices"].values, np.int64),
                        "N": np.asarray(p["N"].values, np.float32).reshape(-1, 3) if "N" in p else None,
                        "uv": np.asarray(p["uv"].values, np.float32).reshape(-1, 2) if "uv" in p else None,
                        "faceIndices": np.asarray(p["faceIndices"].values, np.int64) if "faceIndices" in p else None}
            elif kind == "plymesh":
                data = read_ply(os.path.join(base, p["filename"].values[0]))
            else:
                out.skipped_shapes.append(kind)
                continue
            pts = transform_points(ctm, data["P"])
            nrm = data.get("N")
            if nrm is not None and len(nrm):
                inv_t = np.linalg.inv(ctm).T[:3, :3].astype(np.float32)
                nrm = (nrm @ inv_t.T).astype(np.float32)
            mesh = TriangleMesh.from_xyz(pts, data["indices"],
                                         face_indices=data["faceIndices"] if data.get("faceIndices") is not None
                                         else np.zeros(0, np.int64),
                                         normals=nrm if nrm is not None and len(nrm) else None,
                                         uvs=data["uv"] if data.get("uv") is not None and len(data["uv"]) else None)
            mi = len(meshes)
            meshes.append(mesh)
            nt = mesh.number_triangles
            if area_L is not None:
                parts.append(Primitives.area_lights(mi, nt, n_lights))
                emission += [area_L] * nt
                n_lights += nt
            else:
                parts.append(Primitives.geometric(mi, nt, len(mats)))
                mats.append(current)
        else:
            # other directives (LightSource, Texture, Accelerator, MakeNamedMedium, ...): skip the
            # directive's string argument and parameter list
            while tk.peek() is not None and (tk.peek().startswith('"') or tk.peek() == "["):
                t = tk.next()
                if t == "[":
                    while tk.next() != "]":
                        pass
    gz = abi.concat(mats, abi.GZ_MATERIAL) if mats else make_material("diffuse", {})
    out.scene = Scene(meshes, material_c_table(gz), light_emission=np.asarray(emission, np.float32).reshape(-1, 3))
    out.prims = Primitives.concat(parts)
    out.materials = gz
    return out


__all__ = ["IngestedScene", "ingest", "read_ply", "tokenize", "transform_points", "make_material", "material_c_table"]
