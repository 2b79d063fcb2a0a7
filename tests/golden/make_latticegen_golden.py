"""Writes tests/golden/latticegen_ref.json: SHA-256 digests and line counts of
the node and BC files that the reference generator (latticegen2, compiled
from its own sources into oracle/_ref/latticegen by `make -C oracle ref`)
writes for the STAR-CD geometries in tests/golden/starcd/ (data files of the
reference's STARCDfiles folder), for both complexity settings.

Run from the repository root after `make -C oracle ref`."""
import gzip
import hashlib
import json
import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "latticegen")

# (geometry, nodes in x, seed point)
CASES = [
    ("lid3D", 30, (0.5, 0.5, 0.5)),
    ("Couette3D", 20, (0.5, 0.5, 0.5)),
    ("periodicCube", 17, (0.5, 0.5, 0.5)),
    ("rLid3D", 25, (0.5, 1.5, 0.5)),
    ("channel_8", 40, (0.5, 0.1, 0.1)),
    ("channel3D_4", 60, (0.5, 0.1, 0.1)),
]


def stage(d, name):
    """Copies (and unpacks) one geometry into d/STARCDfiles."""
    os.makedirs(os.path.join(d, "STARCDfiles"), exist_ok=True)
    for ext in ("vrt", "bnd"):
        src = os.path.join(HERE, "starcd", f"{name}.{ext}")
        dst = os.path.join(d, "STARCDfiles", f"{name}.{ext}")
        if os.path.exists(src):
            shutil.copy(src, dst)
        else:
            with gzip.open(src + ".gz", "rb") as fi, open(dst, "wb") as fo:
                fo.write(fi.read())


def digest(path):
    data = open(path, "rb").read()
    return hashlib.sha256(data).hexdigest(), data.count(b"\n")


def main():
    out = []
    for name, n, seed in CASES:
        for c in (1, 2):
            with tempfile.TemporaryDirectory() as d:
                stage(d, name)
                subprocess.run([REF, name, str(n), *map(str, seed), "out", str(c)], cwd=d, check=True,
                               stdout=subprocess.DEVNULL)
                node, bc = digest(os.path.join(d, "out_Node.dat")), digest(os.path.join(d, "out_BC.dat"))
            out.append(dict(geometry=name, n=n, seed=seed, complexity=c, node_sha256=node[0], node_lines=node[1],
                            bc_sha256=bc[0], bc_lines=bc[1]))
    with open(os.path.join(HERE, "latticegen_ref.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
