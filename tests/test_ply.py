"""gz_ply_parse against PLY files written here with known contents (the writer is the oracle:
every value read must be the value written), in both byte orders, every vertex property set the
reference's PlyMesh accepts, uchar and uint face counts, face indices, and the header and body
errors it raises (PlyMesh.swift:24-344)."""
import ctypes

import numpy as np
import pytest

from gonzales_amd import abi

LAYOUTS = {
    "xyz": ["x", "y", "z"],
    "xyz_uv": ["x", "y", "z", "u", "v"],
    "xyz_n": ["x", "y", "z", "nx", "ny", "nz"],
    "xyz_n_uv": ["x", "y", "z", "nx", "ny", "nz", "u", "v"],
    "xyz_n_st": ["x", "y", "z", "nx", "ny", "nz", "s", "t"],
    "xyz_uv_n": ["x", "y", "z", "u", "v", "nx", "ny", "nz"],
}


def write_ply(vertex_props, verts, faces, big=False, list_type="uchar", face_indices=None, comments=("made here",)):
    order = ">" if big else "<"
    head = ["ply", f"format binary_{'big' if big else 'little'}_endian 1.0"]
    head += [f"comment {c}" for c in comments]
    head.append(f"element vertex {len(verts)}")
    head += [f"property float {p}" for p in vertex_props]
    head.append(f"element face {len(faces)}")
    head.append(f"property list {list_type} int vertex_indices")
    if face_indices is not None:
        head.append("property int face_indices")
    head.append("end_header")
    body = bytearray("\n".join(head).encode() + b"\n")
    body += np.asarray(verts, np.float32).astype(order + "f4").tobytes()
    cnt = order + ("u1" if list_type in ("uchar", "uint8") else "u4")
    for k, f in enumerate(faces):
        body += np.array([len(f)], cnt).tobytes()
        body += np.asarray(f, order + "i4").tobytes()
        if face_indices is not None:
            body += np.array([face_indices[k]], order + "i4").tobytes()
    return bytes(body)


def parse(data: bytes):
    lib = abi.load()
    buf = np.frombuffer(data, np.uint8)
    info = abi.GzPlyInfo()
    rc = lib.gz_ply_parse(buf.ctypes.data, len(buf), ctypes.byref(info), None, None, None, None, None)
    if rc != abi.GZ_OK:
        raise ValueError(abi.last_error())
    v = info.vertexCount
    pts = np.zeros((v, 3), np.float32)
    nrm = np.zeros((v, 3), np.float32)
    uvs = np.zeros((v, 2), np.float32)
    idx = np.zeros(info.indexCount, np.int64)
    fidx = np.zeros(info.faceCount, np.int64)
    rc = lib.gz_ply_parse(buf.ctypes.data, len(buf), ctypes.byref(info), pts.ctypes.data, nrm.ctypes.data,
                          uvs.ctypes.data, idx.ctypes.data, fidx.ctypes.data)
    if rc != abi.GZ_OK:
        raise ValueError(abi.last_error())
    return info, pts, nrm, uvs, idx, fidx


@pytest.mark.parametrize("layout", list(LAYOUTS))
@pytest.mark.parametrize("big", [False, True])
@pytest.mark.parametrize("list_type", ["uchar", "uint8", "uint"])
def test_layouts_byte_orders_and_list_types(layout, big, list_type):
    props = LAYOUTS[layout]
    g = np.random.default_rng(len(props) + 10 * big)
    nv, nf = 37, 50
    verts = g.standard_normal((nv, len(props))).astype(np.float32)
    faces = g.integers(0, nv, (nf, 3))
    fi = g.integers(0, 1000, nf)
    data = write_ply(props, verts, faces, big=big, list_type=list_type, face_indices=fi)
    info, pts, nrm, uvs, idx, fidx = parse(data)
    assert (info.vertexCount, info.faceCount, info.indexCount) == (nv, nf, 3 * nf)
    assert info.bigEndian == int(big) and info.hasFaceIndices == 1
    col = {p: k for k, p in enumerate(props)}
    assert np.array_equal(pts, verts[:, [col["x"], col["y"], col["z"]]])
    if "nx" in col:
        assert info.hasNormals == 1
        assert np.array_equal(nrm, verts[:, [col["nx"], col["ny"], col["nz"]]])
    else:
        assert info.hasNormals == 0
    u = col.get("u", col.get("s"))
    if u is not None:
        assert info.hasUvs == 1
        assert np.array_equal(uvs, verts[:, [u, u + 1]])
    assert np.array_equal(idx, faces.reshape(-1)) and np.array_equal(fidx, fi)


def test_non_triangle_faces_are_kept_as_they_are():
    verts = np.arange(12, dtype=np.float32).reshape(4, 3)
    data = write_ply(LAYOUTS["xyz"], verts, [[0, 1, 2, 3], [0, 2, 3]])
    info, _, _, _, idx, _ = parse(data)
    assert info.indexCount == 7 and idx.tolist() == [0, 1, 2, 3, 0, 2, 3]


@pytest.mark.parametrize("header,msg", [
    ("plx\n", "first line"),
    ("ply\nformat ascii 1.0\nend_header\n", "unknown format"),
    ("ply\nformat binary_little_endian 1.1\nend_header\n", "1.0"),
    ("ply\nformat binary_little_endian 1.0\nelement edge 3\nend_header\n", "unknown element"),
    ("ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty double x\nend_header\n", "unknown property"),
    ("ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty float w\nend_header\n", "unknown float"),
    ("ply\nformat binary_little_endian 1.0\nelement face 1\nproperty float x\nend_header\n", "outside the vertex"),
    ("ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty float y\nproperty float x\n"
     "property float z\nend_header\n", "unsupported vertex"),
    ("ply\r\nformat binary_little_endian 1.0\r\nend_header\r\n", "unknown header word"),
    ("ply\nformat binary_little_endian 1.0\n", "end_header"),
])
def test_header_errors(header, msg):
    with pytest.raises(ValueError, match=msg):
        parse(header.encode())


def test_truncated_bodies_are_errors():
    verts = np.zeros((5, 3), np.float32)
    data = write_ply(LAYOUTS["xyz"], verts, [[0, 1, 2], [1, 2, 3]])
    parse(data)
    for cut in (1, 5, 13, 40):
        with pytest.raises(ValueError, match="truncated"):
            parse(data[:-cut])
