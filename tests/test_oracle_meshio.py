"""Pins the CPU oracle of the STL/OBJ readers (oracle/meshio.c) and checks the host
(text) side of the product's readers, which runs without a GPU.

Checks: the reference's own reader tests (tests/io/test_stl.cpp, tests/io/test_obj.cpp),
the merge against an independent first-occurrence merge (a Python dict on coordinate
tuples, where -0.0 == 0.0 as in the reference's ordered map), the number grammar against
std::stod / std::stol compiled here, and the reference's example files (bunny STL,
antenna OBJ) when the reference tree is present."""
from __future__ import annotations

import os
import struct
import subprocess

import numpy as np
import pytest

import oracle
from tests.stlfiles import shared_surface, write_ascii_stl, write_binary_stl, write_obj

REF = "/root/reference"


def dict_weld(soup):
    seen = {}
    pts, idx = [], []
    for v in np.asarray(soup, np.float64).reshape(-1, 3):
        key = (float(v[0]), float(v[1]), float(v[2]))
        if key not in seen:
            seen[key] = len(pts)
            pts.append(v)
        idx.append(seen[key])
    return np.array(pts).reshape(-1, 3), np.array(idx, np.int32)


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(np.asarray(a).view(np.uint64),
                                                 np.asarray(b).view(np.uint64))


def test_weld_against_first_occurrence_dict():
    for seed, n in ((0, 5000), (1, 777)):
        soup = shared_surface(n, seed).astype(np.float64)
        pts, idx = oracle.weld(soup.reshape(-1, 3))
        wp, wi = dict_weld(soup)
        assert same_bits(pts, wp) and np.array_equal(idx, wi)
        assert len(pts) < 0.6 * soup.size / 3  # vertices really are shared


def test_weld_keeps_first_signed_zero():
    soup = np.array([[-0.0, 1, 2], [0.0, 1, 2], [3, 4, 5], [0.0, 1, 2.0]])
    pts, idx = oracle.weld(soup)
    assert np.array_equal(idx, [0, 0, 1, 0])
    assert np.signbit(pts[0, 0])


def tetra_soup():
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
    F = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
    return P, F, P[np.array(F)]


def test_reference_stl_cases(tmp_path):
    # tests/io/test_stl.cpp:35-55: ascii round trip merges back to the 4 corners
    P, F, soup = tetra_soup()
    p = str(tmp_path / "t.stl")
    write_ascii_stl(p, soup, fmt="{:g}")
    pts, tris, binary = oracle.read_stl(p)
    assert not binary and len(tris) == 4 and len(pts) == 4
    assert tris.min() >= 0 and tris.max() < 4
    # :57-79: one binary triangle, 3 points
    b = str(tmp_path / "b.stl")
    write_binary_stl(b, np.array([[[0, 0, 0], [2, 0, 0], [0, 3, 0]]], np.float32))
    pts, tris, binary = oracle.read_stl(b)
    assert binary and len(tris) == 1 and len(pts) == 3
    # :95-109: truncated ascii vertex
    bad = str(tmp_path / "bad.stl")
    open(bad, "w").write("solid bad\n  facet normal 0 0 1\n    outer loop\n      vertex 0 0\n"
                         "    endloop\n  endfacet\nendsolid bad\n")
    with pytest.raises(oracle.ReadError, match="short vertex record"):
        oracle.read_stl(bad)
    with pytest.raises(oracle.ReadError, match="cannot open file"):
        oracle.read_stl(str(tmp_path / "missing.stl"))
    empty = str(tmp_path / "e.stl")
    write_binary_stl(empty, np.zeros((0, 3, 3), np.float32))
    with pytest.raises(oracle.ReadError, match="positive multiple of 3"):
        oracle.read_stl(empty)


def test_reference_obj_cases(tmp_path):
    # tests/io/test_obj.cpp:31-56 round trip
    P, F, _ = tetra_soup()
    p = str(tmp_path / "t.obj")
    write_obj(p, P, F)
    pts, polys = oracle.read_obj(p)
    assert np.array_equal(pts, P) and [list(x) for x in polys] == [list(f) for f in F]
    # :58-77 v/vt/vn corners use the vertex index
    s = str(tmp_path / "s.obj")
    open(s, "w").write("# comment line\nv 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\n"
                       "v 1 1 0\nv 1 0 1\nv 0 1 1\nvt 0 0\nvn 0 0 1\nf 1/2/3 4/5/6 7/8/9\n")
    assert list(oracle.read_obj(s)[1][0]) == [0, 3, 6]
    # :79-92 v//vn
    open(s, "w").write("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1//7 2//8 3//9\n")
    assert list(oracle.read_obj(s)[1][0]) == [0, 1, 2]
    # :94-108 statements skipped
    open(s, "w").write("mtllib scene.mtl\no cube\ng group1\ns 1\nusemtl red\n"
                       "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    pts, polys = oracle.read_obj(s)
    assert len(pts) == 3 and len(polys) == 1
    with pytest.raises(oracle.ReadError, match="cannot open file"):
        oracle.read_obj(str(tmp_path / "missing.obj"))
    open(s, "w").write("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2\n")
    with pytest.raises(oracle.ReadError, match="at least 3 corners"):
        oracle.read_obj(s)
    open(s, "w").write("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 x/2 3\n")
    with pytest.raises(oracle.ReadError, match="malformed face corner: x/2"):
        oracle.read_obj(s)


NUMBERS = ["0", "-0", "1.5", "1e-5", "1E+5", "0x1p-3", "inf", "-nan", "1e400", "-1e400",
           "4.9e-324", "2.2250738585072014e-308", "1e-310", "12abc", "", "+3", "0.1,",
           ".5", "5.", "1e", "nan(123)", "  7"]


def _stod_truth(tmp_path):
    """std::stod and std::stol as the reference calls them (whole token consumed,
    exceptions are failures), compiled here."""
    src = tmp_path / "s.cpp"
    src.write_text(r'''
#include <cstdio>
#include <string>
int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        std::string s(argv[i]);
        try { size_t n; double v = std::stod(s, &n);
              if (n == s.size()) std::printf("ok %a\n", v); else std::printf("bad\n"); }
        catch (...) { std::printf("bad\n"); }
    }
}''')
    exe = tmp_path / "s"
    subprocess.run(["g++", "-O0", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), *NUMBERS], capture_output=True, text=True, check=True)
    return out.stdout.split("\n")[: len(NUMBERS)]


def test_number_grammar_matches_std_stod(tmp_path):
    truth = _stod_truth(tmp_path)
    p = str(tmp_path / "n.obj")
    for tok, want in zip(NUMBERS, truth):
        if "," in tok or tok.strip() != tok or tok == "":
            continue  # separators and blanks never reach the number parser
        open(p, "w").write(f"v {tok} 0 0\n")
        try:
            pts, _ = oracle.read_obj(p)
            got = "ok " + float(pts[0, 0]).hex()
        except oracle.ReadError:
            got = "bad"
        if want.startswith("ok"):
            assert got.startswith("ok"), tok
            g = float.fromhex(got[3:])
            if "nan" in want:
                assert np.isnan(g), tok
            else:
                w = float.fromhex(want[3:])
                assert g == w and np.signbit(g) == np.signbit(w), tok
        else:
            assert got == "bad", tok
    assert truth.count("bad") >= 4 and truth.count("bad") < len(truth) - 8


def test_product_obj_reader_equals_oracle(tmp_path):
    """The product's OBJ reader is host code (no GPU): same points, same polygons."""
    from cybermeshgenerator_amd import files
    rng = np.random.default_rng(4)
    P = rng.normal(size=(300, 3))
    polys = [rng.choice(300, size=rng.integers(3, 7), replace=False) for _ in range(400)]
    p = str(tmp_path / "r.obj")
    write_obj(p, P, polys, extra="# header\nmtllib a.mtl\ng x\n")
    plc = files.read_plc(p)
    pts, opolys = oracle.read_obj(p)
    assert same_bits(plc.points, pts)
    fan = np.array([(q[0], q[i - 1], q[i]) for q in opolys for i in range(2, len(q))])
    assert np.array_equal(plc.triangles, fan)
    for text, msg in (("v 0 0\n", "short vertex record"), ("v 0 0 x\n", "non-numeric"),
                      ("v 0 0 0\nf 1 2\n", "at least 3 corners")):
        open(p, "w").write(text)
        with pytest.raises(RuntimeError, match=msg):
            files.read_plc(p)
    with pytest.raises(RuntimeError, match="unsupported file format"):
        files.read_plc(str(tmp_path / "x.abc"))


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference example files not present")
def test_reference_example_files():
    bunny = os.path.join(REF, "examples/bunny/Stanford_Bunny_sample.stl")
    raw = open(bunny, "rb").read()
    n = struct.unpack("<I", raw[80:84])[0]
    pts, tris, binary = oracle.read_stl(bunny)
    assert binary and len(tris) == n == 112402
    rec = np.frombuffer(raw[84:], dtype=[("nrm", "<f4", 3), ("v", "<f4", (3, 3)),
                                         ("a", "<u2")])
    wp, wi = dict_weld(rec["v"].astype(np.float64))
    assert same_bits(pts, wp) and np.array_equal(tris.ravel(), wi)
    ant = os.path.join(REF, "examples/antenna/Antenna.obj")
    pts, polys = oracle.read_obj(ant)
    text = open(ant).read().splitlines()
    assert len(pts) == sum(1 for l in text if l.startswith("v "))
    assert len(polys) == sum(1 for l in text if l.startswith("f "))
