"""Copies the small UCD grids that pi-BEM ships (data files, no code) into
tests/golden/meshes/, so that the ingest tests run where the reference tree
is absent.  Run from the repository root in a container that has the
reference mounted at /root/reference:

    python tests/golden/make_reference_meshes.py
"""
import os
import shutil

SRC = "/root/reference"
FILES = {
    "grids/coarse_cube.inp": "coarse_cube.inp",
    "grids/coarse_cube_double_nodes.inp": "coarse_cube_double_nodes.inp",
    "grids/coarse_sphere.inp": "coarse_sphere.inp",
    "grids/coarse_sphere_double_nodes.inp": "coarse_sphere_double_nodes.inp",
    "examples/example04_Hemisphere/mesh.inp": "hemisphere.inp",
    "examples/example05_DTC_ship/mesh.inp": "dtc_ship.inp",
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "meshes")
    os.makedirs(out, exist_ok=True)
    for src, dst in FILES.items():
        shutil.copyfile(os.path.join(SRC, src), os.path.join(out, dst))
        print(f"{src} -> meshes/{dst}")
