"""Writes ladybug_49_7776.txt: the visibility graph of the BAL ladybug
problem (49 cameras, 7 776 points, 31 843 camera-point pairs) that nelson
embeds as ProblemCollections::ladybug_49_7776
(nelson/samples/Problems.cpp:38-..., nelson/samples/Problems.h:14-16), in the
text format of Problem::load (nelson/samples/Problems.cpp:8-30):
"nCameras nPoints nEdges" then one "camera point" pair per line.

The pairs are read as data from the initializer's text; nothing of the
reference is compiled or run.  Run once where the reference is present:
    python tests/golden/make_ladybug.py [/root/reference]
"""
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = open(os.path.join(ref, "nelson/samples/Problems.cpp")).read()
head = re.search(r"ladybug_49_7776\s*=\s*Problem\{\s*(\d+)\s*,\s*(\d+)\s*,", src)
n_cameras, n_points = int(head.group(1)), int(head.group(2))
pairs = re.findall(r"\{(\d+),(\d+)\}", src[head.end():])
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ladybug_49_7776.txt")
with open(out, "w") as f:
    f.write(f"{n_cameras} {n_points} {len(pairs)}\n")
    for c, p in pairs:
        f.write(f"{c} {p}\n")
print(f"{out}: {n_cameras} cameras, {n_points} points, {len(pairs)} pairs")
