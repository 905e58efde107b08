"""CPU oracle of STL ingestion (TEST INFRASTRUCTURE ONLY): parseSTL and normalizeAndCenterGeometry of
src/geometry/stlLoader.ts:9-127, restated in Python.

The ASCII path runs the reference's own regular expression (stlLoader.ts:80) with Python's `re`, whose
backtracking order for this pattern is the ECMAScript one; `\\s` and `\\d` are spelled out as ECMAScript
defines them (WhiteSpace + LineTerminator, and ASCII digits), because Python's own classes differ (they
include U+001C..U+001F and U+0085 and exclude U+FEFF). The text is decoded as TextDecoder decodes it (UTF-8,
invalid sequences replaced). Numbers go through float() (nearest double, as parseFloat) and then float32
(as the Float32BufferAttribute store).

The normalisation restates three.js r170 (package.json pins three ^0.170.0, not vendored here):
Box3.setFromBufferAttribute (Math.min/Math.max per point), Box3.getCenter/getSize (zero for an empty box),
BufferGeometry.translate/scale -> applyMatrix4 -> Vector3.applyMatrix4 (row dot product plus e[12..14],
times w = 1 / (e[3]x + e[7]y + e[11]z + e[15]), in double, stored to float32), and applyMatrix4 recomputing
the geometry's existing bounding box in place, which is why the size tested against 10 is the size of the
translated positions.

Pinned by tests/test_stl.py: the reference's ASCII tetrahedron (src/validation/phase1.test.ts:52-100),
hand-derived outcomes of the pattern on edge tokens, a decimal whose double-then-float rounding differs
from direct float rounding, binary files built field by field, and closed-form normalisations.
"""
import re
import struct

import numpy as np

_WS = "\t\n\v\f\r \u00a0\u1680\u2000-\u200a\u2028\u2029\u202f\u205f\u3000\ufeff"
_S = f"[{_WS}]"
_NUM = r"([+-]?[0-9]*\.?[0-9]+(?:[eE][+-]?[0-9]+)?)"
# stlLoader.ts:80, with \s and \d spelled out
VERTEX_RE = re.compile(f"vertex{_S}+{_NUM}{_S}+{_NUM}{_S}+{_NUM}")

INVALID = "Invalid or corrupted STL file."
NO_TRIANGLES = "ASCII STL file contains no valid triangles."


class STLError(ValueError):
    pass


def _decode(b):
    return bytes(b).decode("utf-8", errors="replace")


def parse_ascii(text):
    """parseAsciiSTL (stlLoader.ts:78-95)."""
    pos = [np.float32(float(v)) for m in VERTEX_RE.finditer(text) for v in m.groups()]
    if len(pos) < 9 or len(pos) % 9 != 0:
        raise STLError(NO_TRIANGLES)
    return np.array(pos, dtype=np.float32).reshape(-1, 9)


def parse_binary(b, faces):
    """parseBinarySTL (stlLoader.ts:40-76): the three vertices of each 50-byte record."""
    out = np.zeros((faces, 9), dtype=np.float32)
    for f in range(faces):
        o = 84 + 50 * f + 12
        out[f] = struct.unpack_from("<9f", b, o)
    return out


def parse_stl(b):
    """parseSTL (stlLoader.ts:9-38): (n, 9) float32 positions, or STLError with the reference's message."""
    b = bytes(b)
    if len(b) < 84:
        return parse_ascii(_decode(b))
    faces = struct.unpack_from("<I", b, 80)[0]
    expected = 84 + faces * 50
    if expected == len(b):
        return parse_binary(b, faces)
    header = _decode(b[:min(256, len(b))])
    if re.sub(f"^{_S}+", "", header).lower().startswith("solid"):
        full = _decode(b)
        if "facet" in full and "vertex" in full:
            return parse_ascii(full)
    if expected <= len(b) and faces > 0:
        return parse_binary(b, faces)
    raise STLError(INVALID)


def _js_min(a, b):
    if a != a or b != b:
        return float("nan")
    if a == b:
        return a if np.signbit(a) else b
    return a if a < b else b


def _js_max(a, b):
    if a != a or b != b:
        return float("nan")
    if a == b:
        return b if np.signbit(a) else a
    return a if a > b else b


def _box(t):
    mn = [float("inf")] * 3
    mx = [float("-inf")] * 3
    for v in t.reshape(-1, 3):
        for k in range(3):
            mn[k] = _js_min(mn[k], float(v[k]))
            mx[k] = _js_max(mx[k], float(v[k]))
    empty = mx[0] < mn[0] or mx[1] < mn[1] or mx[2] < mn[2]
    return mn, mx, empty


def _apply(t, e):
    v = t.reshape(-1, 3)
    for i in range(v.shape[0]):
        x, y, z = float(v[i, 0]), float(v[i, 1]), float(v[i, 2])
        w = 1.0 / (e[3] * x + e[7] * y + e[11] * z + e[15])
        v[i, 0] = np.float32((e[0] * x + e[4] * y + e[8] * z + e[12]) * w)
        v[i, 1] = np.float32((e[1] * x + e[5] * y + e[9] * z + e[13]) * w)
        v[i, 2] = np.float32((e[2] * x + e[6] * y + e[10] * z + e[14]) * w)


def normalize_and_center(tris):
    """normalizeAndCenterGeometry (stlLoader.ts:101-127): (centred copy, autoScaledFromMm)."""
    t = np.array(tris, dtype=np.float32).reshape(-1, 9).copy()
    mn, mx, empty = _box(t)
    c = [0.0, 0.0, 0.0] if empty else [(mn[k] + mx[k]) * 0.5 for k in range(3)]
    _apply(t, [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, -c[0], -c[1], -c[2], 1])
    mn, mx, empty = _box(t)
    size = [0.0, 0.0, 0.0] if empty else [mx[k] - mn[k] for k in range(3)]
    max_dim = _js_max(_js_max(size[0], size[1]), size[2])
    scaled = False
    if max_dim > 10.0:
        s = 0.001 if max_dim > 50 else 1.0 / max_dim
        _apply(t, [s, 0, 0, 0, 0, s, 0, 0, 0, 0, s, 0, 0, 0, 0, 1])
        scaled = True
    return t, scaled


def binary_stl(tris, header=b"", normals=None, attr=0, trailing=b"", count=None):
    """Binary STL bytes of (n, 9) positions: 80-byte header, u32 count, 50-byte records."""
    t = np.asarray(tris, dtype=np.float32).reshape(-1, 9)
    n = t.shape[0] if count is None else count
    out = bytearray(header[:80].ljust(80, b" "))
    out += struct.pack("<I", n)
    for i in range(t.shape[0]):
        nv = (0.0, 0.0, 0.0) if normals is None else tuple(normals[i])
        out += struct.pack("<3f", *nv) + t[i].tobytes() + struct.pack("<H", attr)
    return bytes(out) + trailing
