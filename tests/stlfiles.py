"""Writers of STL and OBJ test files (the layouts src/io/stl.cpp and src/io/obj.cpp read
and write), and seeded surfaces with shared vertices."""
from __future__ import annotations

import os

import numpy as np


def write_binary_stl(path, soup32: np.ndarray, header: bytes = b"binary test") -> None:
    """soup32: float32 (T, 3, 3) triangles; normals written as zeros, attribute 0."""
    soup32 = np.asarray(soup32, dtype="<f4").reshape(-1, 3, 3)
    n = soup32.shape[0]
    rec = np.zeros(n, dtype=[("normal", "<f4", 3), ("v", "<f4", (3, 3)), ("attr", "<u2")])
    rec["v"] = soup32
    with open(path, "wb") as f:
        f.write(header[:80].ljust(80, b"\0"))
        f.write(np.uint32(n).astype("<u4").tobytes())
        f.write(rec.tobytes())


def write_ascii_stl(path, soup: np.ndarray, fmt: str = "{!r}") -> None:
    """soup: (T, 3, 3) triangles, coordinates printed with fmt (repr: round-trips)."""
    soup = np.asarray(soup, dtype=np.float64).reshape(-1, 3, 3)
    lines = ["solid test"]
    for tri in soup:
        lines.append("  facet normal 0 0 1")
        lines.append("    outer loop")
        for v in tri:
            lines.append("      vertex " + " ".join(fmt.format(float(c)) for c in v))
        lines.append("    endloop")
        lines.append("  endfacet")
    lines.append("endsolid test")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def shared_surface(ntris: int, seed: int = 0, dtype=np.float32) -> np.ndarray:
    """A (T, 3, 3) soup over a jittered grid of vertices: most vertices are shared by
    six triangles, a few coordinates are -0.0 and +0.0 of the same point, and some
    triangles repeat (as triangle soups from meshing tools do)."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(ntris / 2))) + 1
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side), indexing="ij"), -1)
    z = rng.integers(-3, 4, (side, side)).astype(np.float64) / 8
    pts = np.concatenate([g / side - 0.5, z[..., None]], -1).astype(dtype)
    a = pts[:-1, :-1].reshape(-1, 3)
    b = pts[1:, :-1].reshape(-1, 3)
    c = pts[:-1, 1:].reshape(-1, 3)
    d = pts[1:, 1:].reshape(-1, 3)
    soup = np.concatenate([np.stack([a, b, c], 1), np.stack([b, d, c], 1)])[:ntris].copy()
    # -0.0 for some zero coordinates: merged with +0.0, first occurrence kept
    zero = soup == 0
    flip = zero & (rng.random(soup.shape) < 0.5)
    soup[flip] = -0.0
    rep = rng.integers(0, len(soup), len(soup) // 50)
    soup[rng.integers(0, len(soup), len(rep))] = soup[rep]
    return soup


def write_obj(path, points, polygons, extra: str = "") -> None:
    with open(path, "w") as f:
        f.write(extra)
        for p in points:
            f.write("v {!r} {!r} {!r}\n".format(*(float(x) for x in p)))
        for poly in polygons:
            f.write("f " + " ".join(str(int(v) + 1) for v in poly) + "\n")


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# examples/bunny/README.md:70-72 (cybermesh.voxelize(read_plc(bunny), resolution=48),
# default pad 1, examples/bunny/run_bunny.py:43,194,297-298)
BUNNY_RES48_CELLS = 24667
BUNNY_RES48_VOLUME = 280129


def bunny_stl(directory) -> str:
    """Decompresses the reference's bunny surface (tests/golden) into `directory` and
    returns the .stl path."""
    import lzma
    path = os.path.join(str(directory), "Stanford_Bunny_sample.stl")
    if not os.path.exists(path):
        with lzma.open(os.path.join(GOLDEN, "stanford_bunny_sample.stl.xz")) as f:
            data = f.read()
        with open(path, "wb") as g:
            g.write(data)
    return path
