"""spica scene files: the XML scene description, Radiance RGBE images and OBJ meshes.

`load_xml(path)` reads the Mitsuba-style XML spica's SceneParser reads
(sources/spica/sceneparser.cc:72-335) into this package's Scene and Camera and the
render settings, following the parser's rules:

* elements are visited depth first, children before their parent (`parseChildren`),
  except that a <transform> is read as a whole;
* <boolean>/<integer>/<float>/<string>/<rgb>/<spectrum>/<transform> with a name store a
  parameter in one flat table (`storeToParam`); a <string name="filename"> is made
  absolute relative to the XML file;
* every other element creates the plugin named by its type from the parameters stored
  so far, which the plugin consumes; the object is stored under its name, else under
  its id (if any) and its element name (`bsdf`, `film`, `sampler`, `sensor`, ...);
* <ref id> makes the referenced material the current `bsdf`; an <emitter type="area">
  inside a <shape> turns that shape's triangles into area lights with the current
  "radiance" (one light per triangle, sceneparser.cc:168-188);
* <transform> composes its children in order, each applied after the previous
  (`trans = sub * trans`); <rotate angle> is passed to Transform::rotate as radians,
  as spica does (no degree conversion, sceneparser.cc:141-146);
* vectors and colours parse as "x y z", "x, y, z" or one value for all three
  (vector3d_detail.h:30-44).

Supported plugins are the ones on the path tracer's path: integrators "path" and
"directlighting"; sensor
"perspective"; film "hdrfilm"; samplers "independent" (and "sobol", "halton",
"ldsampler", which map to the counter-based Sobol stream); bsdfs "diffuse",
"conductor", "dielectric", "roughconductor", "plastic", "roughdielectric" and
"roughplastic" with distribution "beckmann" (the default) or "ggx"; shapes "ply" and
"obj"; emitters "area" and "envmap" (RGBE .hdr).  Anything else raises ValueError.
"""
import math
import os
import xml.etree.ElementTree as ET

import numpy as np

from . import _abi
from .meshio import load_ply
from .scene import (Camera, Scene, conductor, dielectric, diffuse, look_at, plastic, roughconductor, roughdielectric,
                    roughplastic)

# ---------------------------------------------------------------- RGBE images
def read_rgbe(path):
    """Radiance .hdr with run-length encoded scanlines, as spica's Image::loadHdr
    (image.cc:252-388): returns float32 [h][w][3], value = m * 2^(e-128) / 256."""
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    valid = False
    while True:
        end = data.index(b"\n", pos)
        line = data[pos:end].decode("ascii", errors="replace")
        pos = end + 1
        if line.startswith("#?RADIANCE") or line.startswith("#?RGBE"):
            valid = True
        if line == "":
            break
    if not valid:
        raise ValueError(f"{path}: not a Radiance HDR file")
    end = data.index(b"\n", pos)
    tok = data[pos:end].decode("ascii").split()
    pos = end + 1
    if len(tok) != 4 or tok[0] != "-Y" or tok[2] != "+X":
        raise ValueError(f"{path}: unsupported HDR orientation {' '.join(tok)}")
    h, w = int(tok[1]), int(tok[3])
    px = np.zeros((h, w, 4), dtype=np.uint8)
    for y in range(h):
        if pos + 4 > len(data) or data[pos] != 2 or data[pos + 1] != 2:
            raise ValueError(f"{path}: scanline {y} is not run-length encoded")
        sw = (data[pos + 2] << 8) | data[pos + 3]
        if sw != w:
            raise ValueError(f"{path}: scanline width {sw} != {w}")
        pos += 4
        for c in range(4):
            x = 0
            while x < w:
                n = data[pos]
                pos += 1
                if n > 128:
                    n -= 128
                    px[y, x:x + n, c] = data[pos]
                    pos += 1
                else:
                    px[y, x:x + n, c] = np.frombuffer(data, dtype=np.uint8, count=n, offset=pos)
                    pos += n
                x += n
    e = px[..., 3].astype(np.float64)
    scale = np.power(2.0, e - 128.0) / 256.0
    return (px[..., :3].astype(np.float64) * scale[..., None]).astype(np.float32)


def write_rgbe(path, rgb):
    """Radiance .hdr (RLE scanlines, literal runs) of float [h][w][3], as spica's
    Image::saveHdr (image.cc:390-435) lays it out."""
    rgb = np.asarray(rgb, dtype=np.float64)
    h, w, _ = rgb.shape
    v = rgb.max(axis=2)
    m, e = np.frexp(v)
    ok = v > 1e-32
    sc = np.where(ok, m * 256.0 / np.where(ok, v, 1.0), 0.0)
    enc = np.zeros((h, w, 4), dtype=np.uint8)
    enc[..., :3] = np.clip(rgb * sc[..., None], 0, 255).astype(np.uint8)
    enc[..., 3] = np.where(ok, e + 128, 0).astype(np.uint8)
    out = bytearray(b"#?RADIANCE\n# Made with 100% pure HDR Shop\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0000000000000\n\n")
    out += f"-Y {h} +X {w}\n".encode("ascii")
    for y in range(h):
        out += bytes([2, 2, (w >> 8) & 0xFF, w & 0xFF])
        for c in range(4):
            x = 0
            while x < w:
                n = min(127, w - x)
                out.append(n)
                out += enc[y, x:x + n, c].tobytes()
                x += n
    with open(path, "wb") as f:
        f.write(out)


# ---------------------------------------------------------------- OBJ meshes
def load_obj(path, to_world=None):
    """Wavefront OBJ: v / vn / f (polygons fanned, negative indices allowed).  Returns the
    world-space triangles [n,3,3] and per-corner normals [n,3,3] or None, as load_ply."""
    pos, nrm, faces = [], [], []
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "v":
                pos.append([float(a) for a in t[1:4]])
            elif t[0] == "vn":
                nrm.append([float(a) for a in t[1:4]])
            elif t[0] == "f":
                corners = []
                for c in t[1:]:
                    parts = c.split("/")
                    vi = int(parts[0])
                    ni = int(parts[2]) if len(parts) > 2 and parts[2] else 0
                    corners.append((vi - 1 if vi > 0 else len(pos) + vi, (ni - 1 if ni > 0 else len(nrm) + ni) if ni else -1))
                for j in range(1, len(corners) - 1):
                    faces.append((corners[0], corners[j], corners[j + 1]))
    P = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    if to_world is not None:
        m = np.asarray(to_world, dtype=np.float64)
        P = P @ m[:3, :3].T + m[:3, 3]
    tris = np.array([[P[c[0]] for c in f] for f in faces], dtype=np.float32).reshape(-1, 3, 3)
    normals = None
    if nrm and all(c[1] >= 0 for f in faces for c in f):
        N = np.asarray(nrm, dtype=np.float64)
        if to_world is not None:
            N = N @ np.linalg.inv(np.asarray(to_world, dtype=np.float64)[:3, :3])
        normals = np.array([[N[c[1]] for c in f] for f in faces], dtype=np.float32).reshape(-1, 3, 3)
    return tris, normals


# ---------------------------------------------------------------- transforms
def parse_vec3(s):
    s = s.strip()
    for sep in (",", None):
        parts = [p for p in (s.split(sep) if sep else s.split()) if p.strip()]
        if len(parts) == 3:
            try:
                return np.array([float(p) for p in parts])
            except ValueError:
                pass
    try:
        v = float(s.split(",")[0].split()[0])
        return np.array([v, v, v])
    except (ValueError, IndexError):
        raise ValueError(f"cannot parse '{s}' as a vector")


def rotate(theta, axis):
    """Transform::rotate (transform.cc:136-160), theta in radians."""
    a = np.asarray(axis, dtype=np.float64)
    a = a * (1.0 / np.linalg.norm(a))
    c, s = math.cos(theta), math.sin(theta)
    x, y, z = a
    m = np.eye(4)
    m[0, :3] = [x * x + (1 - x * x) * c, x * y * (1 - c) - z * s, x * z * (1 - c) + y * s]
    m[1, :3] = [x * y * (1 - c) + z * s, y * y + (1 - y * y) * c, y * z * (1 - c) - x * s]
    m[2, :3] = [x * z * (1 - c) - y * s, y * z * (1 - c) + x * s, z * z + (1 - z * z) * c]
    return m


def parse_transform(elem):
    """SceneParser::parseTransform: compose the children in order (trans = sub * trans)."""
    trans = np.eye(4)
    for ch in elem:
        sub = np.eye(4)
        if ch.tag == "matrix":
            vals = [float(v) for v in ch.get("value").split()]
            if len(vals) != 16:
                raise ValueError("matrix needs 16 values")
            sub = np.array(vals).reshape(4, 4)
        elif ch.tag == "scale":
            sub = np.diag([float(ch.get("x")), float(ch.get("y")), float(ch.get("z")), 1.0])
        elif ch.tag == "rotate":
            sub = rotate(float(ch.get("angle")), [float(ch.get("x")), float(ch.get("y")), float(ch.get("z"))])
        elif ch.tag == "translate":
            sub[:3, 3] = [float(ch.get("x")), float(ch.get("y")), float(ch.get("z"))]
        elif ch.tag == "lookAt":
            sub = look_at(parse_vec3(ch.get("origin")), parse_vec3(ch.get("target")), parse_vec3(ch.get("up")))
        trans = sub @ trans
    return trans


# ---------------------------------------------------------------- the parser
class _Params:
    """spica's RenderParams: one flat table; plugins take (and usually remove) entries."""

    def __init__(self):
        self.d = {}

    def add(self, name, value):
        self.d[name] = value

    def take(self, name, default=None, required=False, remove=True):
        if name not in self.d:
            if required:
                raise ValueError(f"parameter '{name}' is missing")
            return default
        return self.d.pop(name) if remove else self.d[name]


class XmlScene:
    """A parsed scene file: geometry, camera and the render request."""

    def __init__(self):
        self.scene = Scene()
        self.camera = None
        self.spp = None
        self.max_depth = 16
        self.sampler = _abi.SPK_SAMPLER_INDEPENDENT
        self.seed = 0
        self.integrator = None
        self.output = ""

    def workload(self):
        from .configs import Workload
        kind = _abi.SPK_INTEGRATOR_DIRECT if self.integrator == "directlighting" else _abi.SPK_INTEGRATOR_PATH
        return Workload(os.path.basename(self.output or "scene"), self.scene, self.camera, self.spp, self.max_depth,
                        self.sampler, self.seed, integrator=kind)


def _gray(v):
    """A texture's gray() (spectrum.cc: 0.2126 r + 0.7152 g + 0.0722 b) of a float or RGB value."""
    if np.ndim(v) == 0:
        return float(v)
    v = np.asarray(v, dtype=np.float64)
    return float(0.2126 * v[0] + 0.7152 * v[1] + 0.0722 * v[2])


def _dist(p):
    dist = p.take("distribution", "beckmann")
    if dist not in ("beckmann", "ggx"):
        raise ValueError(f"microfacet distribution '{dist}' is not supported (beckmann, ggx)")
    return dist


def _material(kind, p):
    """bsdfs/*.cc (RenderParams constructors): the keys, defaults and gray() reductions spica
    applies; a texture spica dereferences without a default is required here."""
    if kind == "diffuse":
        return diffuse(p.take("reflectance", required=True))
    if kind == "conductor":
        return conductor(eta=p.take("eta", required=True), k=p.take("k", required=True))
    if kind == "dielectric":
        return dielectric(_gray(p.take("intIOR", 1.333)), reflectance=p.take("specularReflectance", (1.0, 1.0, 1.0)),
                          transmittance=p.take("specularTransmittance", (1.0, 1.0, 1.0)))
    if kind == "roughconductor":
        return roughconductor(alpha=_gray(p.take("alpha", 0.1)), eta=p.take("eta", required=True),
                              k=p.take("k", required=True), distribution=_dist(p))
    if kind == "plastic":
        return plastic(p.take("diffuseReflectance", required=True), p.take("specularReflectance", required=True),
                       _gray(p.take("intIOR", 1.5)))
    if kind == "roughdielectric":
        return roughdielectric(alpha=_gray(p.take("alpha", 0.1)), int_ior=_gray(p.take("intIOR", 1.3333)),
                               reflectance=p.take("specularReflectance", required=True),
                               transmittance=p.take("specularTransmittance", required=True), distribution=_dist(p))
    if kind == "roughplastic":
        return roughplastic(alpha=_gray(p.take("alpha", 0.1)), diffuse_reflectance=p.take("diffuseReflectance", required=True),
                            specular_reflectance=p.take("specularReflectance", required=True),
                            int_ior=_gray(p.take("intIOR", 1.5)), distribution=_dist(p))
    raise ValueError(f"bsdf type '{kind}' is not supported")


def load_xml(path):
    """Parse a spica scene file into an XmlScene."""
    path = os.path.abspath(path)
    base = os.path.dirname(path)
    root = ET.parse(path).getroot()
    if root.tag != "scene":
        raise ValueError("XML root node should be \"scene\"")
    out = XmlScene()
    p = _Params()
    p.add("maxDepth", 16)
    state = {"area": False}

    def store(elem):
        tag = elem.tag
        name = elem.get("name", "")
        if tag == "boolean":
            if name:
                p.add(name, elem.get("value") == "true")
        elif tag == "integer":
            if name:
                p.add(name, int(elem.get("value")))
        elif tag == "float":
            if name:
                p.add(name, float(elem.get("value")))
        elif tag == "string":
            v = elem.get("value")
            if name == "filename":
                v = os.path.normpath(os.path.join(base, v))
            if name:
                p.add(name, v)
        elif tag == "rgb":
            if name:
                p.add(name, tuple(parse_vec3(elem.get("value"))))
        elif tag == "spectrum":
            raise ValueError("sampled <spectrum> values are not supported (use <rgb>)")
        elif tag == "transform":
            if name:
                p.add(name, parse_transform(elem))
        elif tag == "shape":
            kind = elem.get("type")
            mat = p.take("bsdf", None)
            if mat is None:
                raise ValueError("shape without a bsdf")
            to_world = p.take("toWorld", np.eye(4))
            if kind == "ply":
                tris, nrm = load_ply(p.take("filename", required=True), to_world)
            elif kind == "obj":
                tris, nrm = load_obj(p.take("filename", required=True), to_world)
            else:
                raise ValueError(f"shape type '{kind}' is not supported (ply, obj)")
            radiance = None
            if state["area"]:
                radiance = p.take("radiance", required=True, remove=False)
            out.scene.add_triangles(tris, mat, radiance=radiance, normals=nrm)
            state["area"] = False
        elif tag == "ref":
            obj = p.take(elem.get("id"), required=True, remove=False)
            if isinstance(obj, dict):  # a material
                p.add("bsdf", obj)
        elif tag == "integrator":
            kind = elem.get("type")
            if kind not in ("path", "directlighting"):
                raise ValueError(f"integrator '{kind}' is not supported (path, directlighting)")
            p.add("integrator", kind)
        else:
            kind = elem.get("type")
            if tag == "emitter" and kind == "area":
                state["area"] = True
                return
            if not kind:
                raise ValueError(f"type is not specified for <{tag}>")
            value = create(tag, kind)
            if name:
                p.add(name, value)
            else:
                if elem.get("id"):
                    p.add(elem.get("id"), value)
                p.add(tag, value)
            if tag == "sensor":
                if out.camera is not None:
                    raise ValueError("multiple cameras are specified")
                out.camera = value

    def create(tag, kind):
        if tag == "bsdf":
            return _material(kind, p)
        if tag == "film":
            if kind != "hdrfilm":
                raise ValueError(f"film '{kind}' is not supported (hdrfilm)")
            p.take("rfilter", None)
            return dict(width=p.take("width", required=True), height=p.take("height", required=True),
                        output=p.take("outputFile", p.take("filename", "")))
        if tag == "rfilter":
            if kind != "box":
                raise ValueError(f"reconstruction filter '{kind}' is not supported (box)")
            return kind
        if tag == "sampler":
            if kind == "independent":
                out.sampler = _abi.SPK_SAMPLER_INDEPENDENT
            elif kind in ("sobol", "halton", "ldsampler"):
                out.sampler = _abi.SPK_SAMPLER_SOBOL
            else:
                raise ValueError(f"sampler '{kind}' is not supported")
            out.seed = int(p.take("seed", 0))
            p.take("ns", None)
            p.take("isPermite", None)
            return kind
        if tag == "sensor":
            if kind != "perspective":
                raise ValueError(f"sensor '{kind}' is not supported (perspective)")
            film = p.take("film", required=True)
            cam = Camera(film["width"], film["height"], p.take("toWorld", required=True), p.take("fov", required=True),
                         p.take("apertureRadius", 0.0), p.take("focusDistance", 50.0))
            out.output = film["output"]
            return cam
        if tag == "emitter":
            if kind != "envmap":
                raise ValueError(f"emitter '{kind}' is not supported (area, envmap)")
            img = read_rgbe(p.take("filename", required=True))
            center = p.take("worldCenter", (0.0, 0.0, 0.0))
            if isinstance(center, str):
                center = tuple(parse_vec3(center))
            out.scene.set_envmap(img, to_world=p.take("toWorld", np.eye(4)), scale=float(p.take("scale", 1.0)),
                                 center=tuple(center), radius=float(p.take("worldRadius", 2.0)))
            return kind
        raise ValueError(f"<{tag} type=\"{kind}\"> is not supported")

    def walk(parent):
        for elem in parent:
            if len(elem) and elem.tag != "transform":
                walk(elem)
            store(elem)

    walk(root)
    if out.camera is None:
        raise ValueError("sensor is not specified")
    out.spp = p.take("sampleCount", required=True, remove=False)
    out.max_depth = p.take("maxDepth", 16, remove=False)
    out.integrator = p.take("integrator", "path", remove=False)
    return out


def spica_filename(fmt, ident):
    """sprintf(fmt, id) as Film::save names its file (film.cc:31-33), for names holding at
    most one integer conversion (%d, %i, %u with flags and width) and %%."""
    out, used, i = [], False, 0
    while i < len(fmt):
        ch = fmt[i]
        if ch != "%":
            out.append(ch)
            i += 1
            continue
        if fmt[i + 1:i + 2] == "%":
            out.append("%")
            i += 2
            continue
        j = i + 1
        while j < len(fmt) and fmt[j] in "-+ 0#":
            j += 1
        while j < len(fmt) and fmt[j].isdigit():
            j += 1
        if used or j >= len(fmt) or fmt[j] not in "diu" or j - i > 8:
            raise ValueError(f"output name {fmt!r} needs at most one %d-style conversion")
        out.append(("%" + fmt[i + 1:j] + "d") % ident)
        used = True
        i = j + 1
    return "".join(out)


def render_xml(path, output=None, save_every=None, queue_size=1 << 24, block_size=64, guiding=False):
    """Render a scene file on the GPU and save the film as spica's HDR film does: the
    file spica_filename(output, samples so far) + ".hdr" (hdrfilm.cc:21-25), every
    save_every samples (spica saves after every pass, integrator.cc:100; default here:
    once, after the last).  output defaults to the scene path without its extension, as
    spica's command line sets "outputFile" (main.cc:37-42).  Returns (image, files)."""
    from . import render
    xs = load_xml(path)
    w = xs.workload()
    if output is None:
        output = xs.output or os.path.splitext(os.path.abspath(path))[0]
    every = max(1, min(w.spp, save_every or w.spp))
    spica_filename(output, 0)   # a bad name fails before the render
    gs = render.GpuScene(w.scene)
    gr = render.GpuRender(gs, w.camera, w.render_desc(), queue_size=queue_size, block_size=block_size)
    g = None
    if guiding:
        g = render.Guide()
        gr.set_guide(g)
    files = []
    done = 0
    img = None
    try:
        while done < w.spp:
            k = min(every, w.spp - done)
            gr.run(k)
            done += k
            img = gr.image()
            name = spica_filename(output, done) + ".hdr"
            write_rgbe(name, img)
            files.append(name)
    finally:
        gr.close()
        if g is not None:
            g.close()
        gs.close()
    return img, files


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="render a spica scene file on the GPU")
    ap.add_argument("scene")
    ap.add_argument("-o", "--output", default=None, help="output name (may hold a %%d for the sample count)")
    ap.add_argument("--save-every", type=int, default=None, help="save every K samples (default: once)")
    ap.add_argument("--guiding", action="store_true")
    a = ap.parse_args()
    for f in render_xml(a.scene, a.output, a.save_every, guiding=a.guiding)[1]:
        print(f)
