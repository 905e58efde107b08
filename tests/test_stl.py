"""STL ingestion: the oracle (oracle/stl_oracle.py) pinned by the reference's own ASCII tetrahedron, by
outcomes of the vertex pattern derived by hand, by binary files built field by field and by closed-form
normalisations; then the library's aero_stl_parse / aero_stl_normalize against the oracle, bit for bit,
on the same cases and on seeded random files. Host code only (no GPU)."""
import struct

import numpy as np
import pytest

import aerosim_amd as A
from oracle import stl_oracle as O

# src/validation/phase1.test.ts:52-84 (the reference's ASCII STL test input)
TETRA = """solid test_tetrahedron
      facet normal 0 0 -1
        outer loop
          vertex 0 0 0
          vertex 0 1 0
          vertex 1 0 0
        endloop
      endfacet
      facet normal 0 -1 0
        outer loop
          vertex 0 0 0
          vertex 1 0 0
          vertex 0 0 1
        endloop
      endfacet
      facet normal -1 0 0
        outer loop
          vertex 0 0 0
          vertex 0 0 1
          vertex 0 1 0
        endloop
      endfacet
      facet normal 0.577 0.577 0.577
        outer loop
          vertex 1 0 0
          vertex 0 1 0
          vertex 0 0 1
        endloop
      endfacet
    endsolid test_tetrahedron"""


def _bits(a):
    return np.asarray(a, dtype=np.float32).view(np.uint32)


def _same(a, b):
    return a.shape == b.shape and np.array_equal(_bits(a), _bits(b))


def _tri(*v):
    return np.array(v, dtype=np.float32).reshape(1, 9)


def _product(data):
    try:
        return A.parse_stl(data), None
    except A.AeroError as e:
        return None, str(e)


def _oracle(data):
    try:
        return O.parse_stl(data), None
    except O.STLError as e:
        return None, str(e)


def _agree(data):
    p, pe = _product(data)
    o, oe = _oracle(data)
    if oe is not None:
        assert p is None and oe in pe, (data[:120], pe, oe)
    else:
        assert pe is None and _same(p, o), (data[:120], pe)


# ---- oracle pins ---------------------------------------------------------------------------------

def test_oracle_reference_tetrahedron():
    t = O.parse_stl(TETRA.encode())
    assert t.shape == (4, 9)
    assert t[0].tolist() == [0, 0, 0, 0, 1, 0, 1, 0, 0]
    assert t[3].tolist() == [1, 0, 0, 0, 1, 0, 0, 0, 1]
    v = t.reshape(-1, 3, 3).astype(np.float64)
    vol6 = np.einsum("ij,ij->i", v[:, 0], np.cross(v[:, 1], v[:, 2])).sum()
    assert abs(abs(vol6) / 6 - 1 / 6) < 1e-15  # the reference expects volume 1/6


# (text after "solid s facet\n", expected vertices or None when no vertex matches there)
PATTERN_CASES = [
    ("vertex 1 2 3", [1, 2, 3]),
    ("vertex 1. 2. 3.", None),                 # "1." is followed by '.', not by whitespace
    ("vertex 1 2 3.", [1, 2, 3]),              # the last number needs nothing after it: "3"
    ("vertex\t-1.5e+2\n+.25  7E-1", [-150, 0.25, 0.7]),
    ("VERTEX 1 2 3", None),                    # the pattern is case-sensitive
    ("xvertex 4 5 6", [4, 5, 6]),              # no word boundary in the pattern
    ("vertex 1,2 3 4", None),
    ("vertex 1e5 2e 3", None),                 # "2e" is not a number
    ("vertex 1 2 3e", [1, 2, 3]),              # exponent needs a digit: the last number is "3"
    ("vertex 1 2 .", None),
    ("vertex 1 2 3vertex 4 5 6", [1, 2, 3, 4, 5, 6]),
    ("vertex vertex 7 8 9", [7, 8, 9]),        # the first "vertex" fails, the second matches
    ("vertex\u00a01\u20032\u30003", [1, 2, 3]),  # ECMAScript whitespace outside ASCII
    ("vertex 1\u00852 3", None),               # U+0085 is not ECMAScript whitespace
    ("vertex 1\u001c2 3", None),               # nor is U+001C
    ("vertex 1\ufeff2\ufeff3", [1, 2, 3]),     # U+FEFF is
    ("vertex 1 2 1.2.3", [1, 2, 1.2]),
    ("vertex -0 +0 00012", [-0.0, 0.0, 12]),
    ("vertex 1e400 -1e400 1e-400", [np.inf, -np.inf, 0.0]),
    ("vertex 1 2\n\n\n 3 vertex 4", [1, 2, 3]),
]


@pytest.mark.parametrize("text,want", PATTERN_CASES)
def test_oracle_vertex_pattern(text, want):
    body = ("solid s facet\n" + text + "\n").encode()
    # pad the vertex list to whole triangles with a known triangle so that the count rule passes
    pad = "".join(f" vertex {i} {i} {i}" for i in range(3))
    if want is None:
        got = O.parse_ascii((body.decode() + pad))
        assert got.reshape(-1).tolist() == [0, 0, 0, 1, 1, 1, 2, 2, 2]
    else:
        need = (-len(want)) % 9
        extra = "".join(f" vertex {i} {i} {i}" for i in range(need // 3))
        got = O.parse_ascii(body.decode() + extra).reshape(-1)
        w = np.array(want, dtype=np.float32)
        assert np.array_equal(_bits(got[:len(want)]), _bits(w)), (got[:len(want)], w)


def test_oracle_rounds_through_double():
    # 1 + 2^-24 + 1e-25: above the float halfway point 1 + 2^-24, so direct rounding to float gives
    # 1 + 2^-23, but its nearest double is 1 + 2^-24 exactly, which then rounds to even: 1.0
    s = "1.0000000596046447753906251"
    assert float(s) == 1 + 2 ** -24
    t = O.parse_ascii(f"vertex {s} 0 0 vertex 0 0 0 vertex 0 0 0")
    assert t[0, 0] == np.float32(1.0)


def _bin(tris, header=b"", count=None, trailing=b""):
    """Binary STL written field by field (little-endian)."""
    t = np.asarray(tris, dtype=np.float32).reshape(-1, 9)
    n = t.shape[0] if count is None else count
    b = header.ljust(80, b"\0")[:80] + struct.pack("<I", n)
    for row in t:
        b += struct.pack("<3f", 0.5, -0.5, 9.0) + struct.pack("<9f", *row.tolist()) + struct.pack("<H", 7)
    return b + trailing


def test_oracle_binary_layouts():
    rng = np.random.default_rng(1)
    t = rng.standard_normal((5, 9)).astype(np.float32)
    assert _same(O.parse_stl(_bin(t)), t)
    # many exporters write "solid" into a binary header: the exact size decides first
    assert _same(O.parse_stl(_bin(t, header=b"solid part facet vertex")), t)
    # trailing bytes: no "solid" header, the count fits -> binary
    assert _same(O.parse_stl(_bin(t, trailing=b"xyz")), t)
    # "solid" header and trailing bytes but no "facet": binary fallback
    assert _same(O.parse_stl(_bin(t, header=b"  SOLID x", trailing=b"\0\0")), t)
    # count 0 and 84 bytes: an empty geometry, not an error
    assert O.parse_stl(_bin(np.zeros((0, 9)))).shape == (0, 9)
    # count larger than the file: rejected
    with pytest.raises(O.STLError, match="Invalid or corrupted"):
        O.parse_stl(_bin(t, count=6))
    # "solid" header with the ASCII words: parsed as ASCII, and its failure is not a binary fallback
    fake = _bin(t, header=b"solid facet vertex", trailing=b"")
    fake = fake + b"tail"
    with pytest.raises(O.STLError, match="no valid triangles"):
        O.parse_stl(fake)
    # short input is ASCII whatever it holds
    with pytest.raises(O.STLError, match="no valid triangles"):
        O.parse_stl(b"\0" * 83)


def test_oracle_normalization_closed_forms():
    box = A.mesh_box((0, 0, 0), (100, 40, 20))              # 100 mm: mm -> m
    t, s = O.normalize_and_center(box)
    v = t.reshape(-1, 3)
    assert s and np.allclose(v.min(0), [-0.05, -0.02, -0.01]) and np.allclose(v.max(0), [0.05, 0.02, 0.01])
    t, s = O.normalize_and_center(A.mesh_box((-3, 1, 2), (17, 6, 4)))  # largest side 20: scaled to 1
    v = t.reshape(-1, 3)
    assert s and v[:, 0].min() == -0.5 and v[:, 0].max() == 0.5
    t, s = O.normalize_and_center(A.mesh_box((0, 0, 0), (10, 3, 3)))   # exactly 10: untouched size
    assert not s and t.reshape(-1, 3)[:, 0].max() == 5.0
    t, s = O.normalize_and_center(A.mesh_box((0, 0, 0), (50, 3, 3)))   # exactly 50: 1/50, not 0.001
    assert s and t.reshape(-1, 3)[:, 0].max() == np.float32(25.0 / 50.0)
    t, s = O.normalize_and_center(np.zeros((0, 9), dtype=np.float32))
    assert t.shape == (0, 9) and not s


def test_oracle_threshold_uses_the_translated_box():
    """A side just over 10 on the input floats that the translation to float rounds to exactly 10: the
    reference measures the translated box, so it does not rescale."""
    a = np.float32(7e-7)
    b = np.nextafter(np.float32(10.0), np.float32(np.inf))  # 10 + 2^-20
    assert float(b) - float(a) > 10.0                         # the input box is wider than 10
    c = (float(a) + float(b)) * 0.5
    assert float(np.float32(float(b) - c)) - float(np.float32(float(a) - c)) == 10.0   # the premise
    tri = _tri(a, 0, 0, b, 1, 0, a, 0, 1)
    t, s = O.normalize_and_center(tri)
    assert not s
    assert t[0, 3] - t[0, 0] == np.float32(10.0)


# ---- the library against the oracle -------------------------------------------------------------

def test_reference_tetrahedron_through_the_library():
    """phase1.test.ts:86-99: 4 triangles, watertight, volume 1/6 (4 digits), frontal area 0.5 (2 digits)."""
    t = A.parse_stl(TETRA.encode())
    assert _same(t, O.parse_stl(TETRA.encode()))
    st = A.mesh_stats(t)
    assert t.shape[0] == 4 and st["watertight"]
    assert abs(st["volume"] - 1 / 6) < 5e-5
    assert abs(st["frontal_area"] - 0.5) < 5e-3


@pytest.mark.parametrize("text,want", PATTERN_CASES)
def test_library_vertex_pattern(text, want):
    pad = "".join(f" vertex {i} {i} {i}" for i in range(9))
    for body in ("solid s facet\n" + text + pad, text + pad, "solid s facet\n" + text):
        _agree(body.encode())


def test_library_rounding_and_errors():
    _agree(b"vertex 1.0000000596046447753906251 0 0 vertex 0 0 0 vertex 0 0 0")
    _agree(b"solid nothing here")
    _agree(b"")
    _agree(b"\0" * 83)
    p, pe = _product(b"vertex 1 2 3 vertex 4 5 6")
    assert p is None and "ASCII STL file contains no valid triangles." in pe


def test_library_binary_layouts():
    rng = np.random.default_rng(2)
    t = rng.standard_normal((7, 9)).astype(np.float32)
    for data in (_bin(t), _bin(t, header=b"solid part facet vertex"), _bin(t, trailing=b"xyz"),
                 _bin(t, header=b"  SOLID x", trailing=b"\0\0"), _bin(np.zeros((0, 9))), _bin(t, count=8),
                 _bin(t, header=b"solid facet vertex") + b"tail", _bin(t, count=0, trailing=b"")[:84] + b"1"):
        _agree(data)


def _rand_text(rng):
    nums = ["0", "1", "-2", "+3", "4.5", "-.25", "6.", ".", "1e5", "1E-5", "2e", "1.e3", "1,5", "nan", "Infinity",
            "1e400", "-1e-400", "0.1", "123456789012345678901234567890", "3.4028235e38", "1.4e-45", "-0", "9e+09",
            "1.0000000596046447753906251", "+.5e-3"]
    words = ["vertex", "vertex", "vertex", "VERTEX", "facet", "normal", "outer", "loop", "endloop", "endfacet",
             "solid", "x"]
    seps = [" ", " ", " ", "  ", "\t", "\n", "\r\n", "\u00a0", "\u2003", "\u0085", "\u001c", "\ufeff", "\u3000", ""]
    parts = []
    if rng.random() < 0.7:
        parts.append(rng.choice(["solid", "SOLID", "  solid", "\ufeffsolid", "\u00a0solid", "sol"]))
    for _ in range(int(rng.integers(0, 60))):
        parts.append(str(rng.choice(words if rng.random() < 0.3 else nums)))
        parts.append(str(rng.choice(seps)))
    b = "".join(parts).encode("utf-8")
    if rng.random() < 0.2:  # stray bytes that are not UTF-8
        i = int(rng.integers(0, len(b) + 1))
        b = b[:i] + bytes(rng.choice([b"\xff", b"\xc2", b"\xe2\x80", b"\xa0"])) + b[i:]
    return b


def _rand_facets(rng):
    """Facet-structured ASCII with unusual number spellings and separators, and rare corruptions."""
    good = ["0", "1", "-2", "+3", "4.5", "-.25", "1e5", "1E-5", "0.1", "-0", "9e+09", "3.4028235e38", "1.4e-45",
            "1e400", "-1e-400", "1.0000000596046447753906251", "+.5e-3", "123456789012345678901234567890"]
    bad = ["6.", ".", "2e", "1,5", "nan", "1.e3", "VERTEX", "vertex"]
    seps = [" ", " ", " ", "\t", "  ", "\u00a0", "\u3000", "\ufeff", "\u0085"]
    lines = [str(rng.choice(["solid part", "  SOLID", "solid"]))]
    for _ in range(int(rng.integers(1, 9))):
        lines.append("facet normal 0 0 1\n outer loop")
        for _ in range(3):
            sp = [str(rng.choice(seps)) if rng.random() < 0.15 else " " for _ in range(3)]
            nums = [str(rng.choice(bad)) if rng.random() < 0.01 else str(rng.choice(good)) for _ in range(3)]
            lines.append("vertex" + "".join(a + b for a, b in zip(sp, nums)))
        lines.append(" endloop\nendfacet")
    lines.append("endsolid")
    return "\n".join(lines).encode("utf-8")


def test_library_matches_oracle_on_random_text():
    rng = np.random.default_rng(3)
    parsed = 0
    for i in range(3000):
        data = _rand_text(rng) if i % 2 else _rand_facets(rng)
        _agree(data)
        parsed += _oracle(data)[1] is None
    assert parsed > 500  # both the parse and the rejection paths are exercised


def test_library_matches_oracle_on_random_binary():
    rng = np.random.default_rng(4)
    for _ in range(300):
        n = int(rng.integers(0, 12))
        bits = rng.integers(0, 2 ** 32, size=(n, 9), dtype=np.uint64).astype(np.uint32)
        bits &= ~np.uint32(0x40000000)  # keep exponents below 0xFF: no NaN or infinity
        t = bits.view(np.float32)
        header = rng.choice([b"", b"solid x", b"solid facet vertex", b"binary"])
        trailing = bytes(rng.integers(0, 256, size=int(rng.integers(0, 3)) * int(rng.integers(0, 60)),
                                      dtype=np.uint8))
        count = None if rng.random() < 0.7 else int(rng.integers(0, 20))
        _agree(_bin(t, header=bytes(header), count=count, trailing=trailing))


def test_library_normalization_matches_oracle():
    rng = np.random.default_rng(5)
    cases = [A.mesh_box((0, 0, 0), (100, 40, 20)), A.mesh_box((-3, 1, 2), (17, 6, 4)),
             A.mesh_box((0, 0, 0), (10, 3, 3)), A.mesh_box((0, 0, 0), (50, 3, 3)), np.zeros((0, 9), np.float32),
             A.mesh_icosphere(2, 0.5, (3.0, -1.0, 2.0))]
    a = np.float32(7e-7)
    b = np.nextafter(np.float32(10.0), np.float32(np.inf))
    cases.append(_tri(a, 0, 0, b, 1, 0, a, 0, 1))
    for _ in range(40):
        scale = 10 ** rng.uniform(-2, 4)
        cases.append((rng.standard_normal((int(rng.integers(1, 30)), 9)) * scale + rng.standard_normal(9) * scale)
                     .astype(np.float32))
    for c in cases:
        p, ps = A.normalize_and_center_geometry(c)
        o, os_ = O.normalize_and_center(c)
        assert ps == os_ and _same(p, o)


def test_load_stl_file(tmp_path):
    path = tmp_path / "part.stl"
    path.write_bytes(_bin(A.mesh_box((0, 0, 0), (120, 60, 30))))
    t, scaled = A.load_stl(str(path))
    assert scaled and t.shape == (12, 9)
    assert np.allclose(t.reshape(-1, 3).max(0), [0.06, 0.03, 0.015])


# ---- mesh statistics details ---------------------------------------------------------------------

def test_mesh_stats_welds_with_math_round_ties():
    """Vertices are welded on Math.round(v * 1e5), whose halves go toward +infinity: -0.015625 * 1e5 =
    -1562.5 rounds to -1562, the key of -0.01562. A tetrahedron whose apex is written once as -0.015625 and
    twice as -0.01562 is therefore watertight."""
    a, b = -0.015625, -0.01562
    assert np.float32(a) * 1e5 == -1562.5
    P = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]
    tris = np.array([[*P[0], *P[2], *P[1]],
                     [*P[0], *P[1], 0, 0, a],
                     [*P[1], *P[2], 0, 0, b],
                     [*P[2], *P[0], 0, 0, b]], dtype=np.float32)
    assert A.mesh_stats(tris)["open_edges"] == 0


def test_mesh_stats_skips_the_edge_check_above_100k_triangles():
    """The reference checks watertightness only up to 100000 triangles (stlLoader.ts:190)."""
    rng = np.random.default_rng(6)
    soup = rng.standard_normal((100001, 9)).astype(np.float32)
    assert A.mesh_stats(soup)["open_edges"] == 0
    assert A.mesh_stats(soup[:1000])["open_edges"] == 3000
