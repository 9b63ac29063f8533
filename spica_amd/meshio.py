"""Minimal PLY reader (binary little-endian or ASCII) for triangle meshes.

spica loads PLY meshes into one Triangle primitive per face with the mesh's
transform applied (sources/core/meshio.cc:76-169).  This reader returns the
world-space triangle soup [n, 3, 3] (float32) and, when the file carries vertex
normals, per-corner normals [n, 3, 3]; polygons are fanned into triangles.
"""
import numpy as np

_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2",
    "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
    "float": "f4", "float32": "f4", "double": "f8", "float64": "f8",
}


def load_ply(path, to_world=None):
    with open(path, "rb") as f:
        data = f.read()
    end = data.index(b"end_header")
    end = data.index(b"\n", end) + 1
    header = data[:end].decode("ascii", errors="replace").splitlines()
    if not header or header[0].strip() != "ply":
        raise ValueError(f"{path}: not a PLY file")
    fmt = None
    elements = []
    for line in header[1:]:
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            elements.append([tok[1], int(tok[2]), []])
        elif tok[0] == "property":
            if tok[1] == "list":
                elements[-1][2].append((tok[4], "list", tok[2], tok[3]))
            else:
                elements[-1][2].append((tok[2], tok[1]))
    if fmt not in ("binary_little_endian", "ascii"):
        raise ValueError(f"{path}: unsupported PLY format {fmt}")
    body = data[end:]
    verts = normals = faces = None
    if fmt == "ascii":
        words = body.split()
        pos = 0
        for name, count, props in elements:
            if name == "vertex":
                nprop = len(props)
                arr = np.array(words[pos:pos + count * nprop], dtype=np.float64).reshape(count, nprop)
                pos += count * nprop
                names = [p[0] for p in props]
                verts = arr[:, [names.index(a) for a in ("x", "y", "z")]]
                if all(a in names for a in ("nx", "ny", "nz")):
                    normals = arr[:, [names.index(a) for a in ("nx", "ny", "nz")]]
            elif name == "face":
                polys = []
                for _ in range(count):
                    k = int(words[pos])
                    polys.append([int(w) for w in words[pos + 1:pos + 1 + k]])
                    pos += 1 + k
                faces = polys
            else:
                pos += count * len(props)
    else:
        off = 0
        for name, count, props in elements:
            if all(p[1] != "list" for p in props):
                dt = np.dtype([(p[0], "<" + _PLY_TYPES[p[1]]) for p in props])
                arr = np.frombuffer(body, dtype=dt, count=count, offset=off)
                off += dt.itemsize * count
                if name == "vertex":
                    verts = np.stack([arr["x"], arr["y"], arr["z"]], axis=1).astype(np.float64)
                    if all(a in arr.dtype.names for a in ("nx", "ny", "nz")):
                        normals = np.stack([arr["nx"], arr["ny"], arr["nz"]], axis=1).astype(np.float64)
            else:
                if len(props) != 1:
                    raise ValueError(f"{path}: element {name} mixes lists and scalars")
                _, _, ct, it = props[0]
                cdt, idt = np.dtype("<" + _PLY_TYPES[ct]), np.dtype("<" + _PLY_TYPES[it])
                k0 = int(np.frombuffer(body, dtype=cdt, count=1, offset=off)[0])
                rec = np.dtype([("n", cdt), ("i", idt, (k0,))])
                arr = np.frombuffer(body, dtype=rec, count=count, offset=off)
                if np.all(arr["n"] == k0):
                    polys = arr["i"]
                    off += rec.itemsize * count
                else:
                    polys = []
                    for _ in range(count):
                        k = int(np.frombuffer(body, dtype=cdt, count=1, offset=off)[0])
                        off += cdt.itemsize
                        polys.append(np.frombuffer(body, dtype=idt, count=k, offset=off).tolist())
                        off += idt.itemsize * k
                if name == "face":
                    faces = polys
    if verts is None or faces is None:
        raise ValueError(f"{path}: no vertex/face elements")
    tri_idx = []
    for p in faces:
        p = list(p)
        for j in range(1, len(p) - 1):
            tri_idx.append((p[0], p[j], p[j + 1]))
    tri_idx = np.asarray(tri_idx, dtype=np.int64)
    if to_world is not None:
        m = np.asarray(to_world, dtype=np.float64)
        verts = verts @ m[:3, :3].T + m[:3, 3]
        if normals is not None:
            normals = normals @ np.linalg.inv(m[:3, :3])
    tris = verts[tri_idx].astype(np.float32)
    nrm = None if normals is None else normals[tri_idx].astype(np.float32)
    return tris, nrm
