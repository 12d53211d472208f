"""Regenerate the fixtures in this folder from the reference checkout.

The files are data the reference's own tests and docs hold:
  meshes/laplace_unit_3by3.unv.gz   examples/0_GRIDS/laplace_unit_3by3.unv (unit_test_laplace.jl:8)
  meshes/quad40.unv.gz, trig40.unv.gz  test/grids/ (test_mesh_conversion.jl:33,41)
  meshes/OF_cavity_hex/*.gz         test/grids/OF_cavity_hex/polyMesh (test_mesh_conversion.jl:57)
  meshes/backwardFacingStep_10mm.unv.gz  examples/0_GRIDS (2d_incompressible_laminar_BFS.jl)
  2D_laminar_BFS_XCALibre.csv       docs/src/examples/figures/01/ (verification profile)
  laplace_3by3.json                 expected A and b of unit_test_laplace.jl:115-127
  mesh_counts.json                  cell/face/node counts of test_mesh_conversion.jl:33-60
"""
import gzip
import json
import shutil
from pathlib import Path

REF = Path("/root/reference")
HERE = Path(__file__).resolve().parent


def gz(src: Path, dst: Path):
    dst.parent.mkdir(parents=True, exist_ok=True)
    with open(src, "rb") as f, gzip.open(dst, "wb", compresslevel=9) as g:
        shutil.copyfileobj(f, g)


def main():
    gz(REF / "examples/0_GRIDS/laplace_unit_3by3.unv", HERE / "meshes/laplace_unit_3by3.unv.gz")
    gz(REF / "test/grids/quad40.unv", HERE / "meshes/quad40.unv.gz")
    gz(REF / "test/grids/trig40.unv", HERE / "meshes/trig40.unv.gz")
    gz(REF / "examples/0_GRIDS/backwardFacingStep_10mm.unv",
       HERE / "meshes/backwardFacingStep_10mm.unv.gz")
    for f in ("points", "faces", "owner", "neighbour", "boundary"):
        gz(REF / "test/grids/OF_cavity_hex/polyMesh" / f, HERE / f"meshes/OF_cavity_hex/{f}.gz")
    shutil.copy(REF / "docs/src/examples/figures/01/2D_laminar_BFS_XCALibre.csv", HERE)
    # unit_test_laplace.jl:115-127 (A compared with atol=0.1, b with isapprox)
    A = [[4, -1, 0, -1, 0, 0, 0, 0, 0],
         [-1, 5, -1, 0, -1, 0, 0, 0, 0],
         [0, -1, 6, 0, 0, -1, 0, 0, 0],
         [-1, 0, 0, 3, -1, 0, -1, 0, 0],
         [0, -1, 0, -1, 4, -1, 0, -1, 0],
         [0, 0, -1, 0, -1, 5, 0, 0, -1],
         [0, 0, 0, -1, 0, 0, 2, -1, 0],
         [0, 0, 0, 0, -1, 0, -1, 3, -1],
         [0, 0, 0, 0, 0, -1, 0, -1, 4]]
    b = [100.0, 100.0, 120.0, 0.0, 0.0, 20.0, 0.0, 0.0, 20.0]
    json.dump({"source": "test/unit_test_laplace.jl:115-130", "A": A, "b": b, "A_atol": 0.1},
              open(HERE / "laplace_3by3.json", "w"), indent=1)
    counts = {"source": "test/test_mesh_conversion.jl:33-60",
              "trig40": [3484, 5306, 1823], "quad40": [1600, 3280, 1681],
              "OF_cavity_hex": [125, 450, 216]}
    json.dump(counts, open(HERE / "mesh_counts.json", "w"), indent=1)


if __name__ == "__main__":
    main()
