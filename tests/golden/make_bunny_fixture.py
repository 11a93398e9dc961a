"""Writes tests/golden/stanford_bunny_sample.stl.xz: the reference's own bunny surface
(examples/bunny/Stanford_Bunny_sample.stl, a binary STL of 112,402 triangles), stored
xz-compressed as a data fixture. Its published occupancy (examples/bunny/README.md:70-72:
24,667 occupied cells, volume 280,129 at resolution 48, default pad 1, the run of
examples/bunny/run_bunny.py:43,194,297-298) pins the voxelizer end to end.

Run once in a checkout that has the reference tree at /root/reference:
    python tests/golden/make_bunny_fixture.py
"""
import lzma
import os

SRC = "/root/reference/examples/bunny/Stanford_Bunny_sample.stl"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "stanford_bunny_sample.stl.xz")

if __name__ == "__main__":
    with open(SRC, "rb") as f:
        data = f.read()
    with lzma.open(DST, "wb", preset=9) as f:
        f.write(data)
    print(f"{DST}: {len(data)} bytes of STL")
