"""Copy the reference's example inputs (its P1 Gmsh MSH 4.0 meshes and example configs, data
files only) into tests/golden/reference_inputs/, so GPU tests can run them where the reference
tree is absent. Run once from the repository root: python tests/golden/fetch_reference_inputs.py"""
import os
import shutil

SRC = "/root/reference"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_inputs")
FILES = ["doc/1d/line.msh", "doc/2d/square.msh", "doc/2d/room.msh", "doc/3d/cube.msh",
         "doc/config/config.conf", "doc/config/Room_2D.conf"]

if __name__ == "__main__":
    os.makedirs(DST, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(SRC, f), os.path.join(DST, os.path.basename(f)))
        print(f)
