"""Copy the reference's gradient-test scenes (data files) into
tests/golden/reference_scenes/ so that tests can run where the reference is
not mounted.

Copied: Assets/GradientTests/<case>/{<case>.xml, initial.ply, target.ply}
for the cases listed in python/finite_difference/tests.json, and the mesh
files those XMLs name (Assets/meshes/rectangle.obj).  These are scene
descriptions and point clouds the reference's own FD tests load; nothing
executable is copied.  They are CC BY-NC 4.0 data of the reference;
reference_scenes/NOTICE.md names their sources and the licence.

Usage: python tests/golden/collect_reference_scenes.py [/root/reference]
"""
import os
import re
import shutil
import sys

CASES = ("empty", "simple_1")


def main(ref: str = "/root/reference") -> None:
    src_root = os.path.join(ref, "Assets")
    dst_root = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_scenes")
    meshes = set()
    for case in CASES:
        d = os.path.join(dst_root, "GradientTests", case)
        os.makedirs(d, exist_ok=True)
        for f in (f"{case}.xml", "initial.ply", "target.ply"):
            shutil.copyfile(os.path.join(src_root, "GradientTests", case, f), os.path.join(d, f))
        xml = open(os.path.join(d, f"{case}.xml")).read()
        xml = re.sub(r"<!--.*?-->", "", xml, flags=re.S)
        meshes.update(re.findall(r'name="filename" value="([^"]+)"', xml))
    for m in sorted(meshes):
        os.makedirs(os.path.dirname(os.path.join(dst_root, m)), exist_ok=True)
        shutil.copyfile(os.path.join(src_root, m), os.path.join(dst_root, m))
    print("copied", CASES, sorted(meshes))


if __name__ == "__main__":
    main(*sys.argv[1:])
