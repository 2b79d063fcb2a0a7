"""The lattice generator (lbm_latticegen and the latticegen command) against
the reference generator: the reference's own L/lidvalidation files byte for
byte, the SHA-256 digests of the files the reference generator (compiled from
its sources) writes for six STAR-CD geometries of its STARCDfiles folder, and
invariants on a synthetic box with an obstacle: the fluid is every node not
strictly inside the obstacle, and with complexity 2 the boundary links are
exactly the links from fluid nodes to non-fluid or outside nodes."""
import gzip
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from thesis_amd import lattices as LT
from thesis_amd import lbm as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CLI = os.path.join(ROOT, "thesis_amd", "latticegen")


def stage(d, name):
    os.makedirs(os.path.join(d, "STARCDfiles"), exist_ok=True)
    paths = []
    for ext in ("vrt", "bnd"):
        src = os.path.join(GOLDEN, "starcd", f"{name}.{ext}")
        dst = os.path.join(d, "STARCDfiles", f"{name}.{ext}")
        data = open(src, "rb").read() if os.path.exists(src) else gzip.open(src + ".gz", "rb").read()
        open(dst, "wb").write(data)
        paths.append(dst)
    return paths


def test_lid_cavity_equals_reference_files(tmp_path):
    """latticegen lid3D 30 0.5 0.5 0.5 reproduces L/lidvalidation_{Node,BC}.dat."""
    vrt, bnd = stage(str(tmp_path), "lid3D")
    for c in (1, 2):
        out = str(tmp_path / f"lid{c}")
        dims, nl = L.latticegen(vrt, bnd, 30, (0.5, 0.5, 0.5), out, complexity=c)
        assert dims == (30, 30, 30) and nl == 26640
        for kind in ("Node", "BC"):
            want = gzip.open(os.path.join(GOLDEN, f"lidvalidation_{kind}.dat.gz"), "rb").read()
            assert open(f"{out}_{kind}.dat", "rb").read() == want, kind


CASES = json.load(open(os.path.join(GOLDEN, "latticegen_ref.json")))


@pytest.mark.parametrize("case", CASES, ids=[f"{c['geometry']}-c{c['complexity']}" for c in CASES])
def test_matches_reference_generator(tmp_path, case):
    vrt, bnd = stage(str(tmp_path), case["geometry"])
    out = str(tmp_path / "out")
    dims, nl = L.latticegen(vrt, bnd, case["n"], case["seed"], out, complexity=case["complexity"])
    for kind, key in (("Node", "node"), ("BC", "bc")):
        data = open(f"{out}_{kind}.dat", "rb").read()
        assert data.count(b"\n") == case[f"{key}_lines"], kind
        assert hashlib.sha256(data).hexdigest() == case[f"{key}_sha256"], kind
    assert nl == case["bc_lines"]


def test_command_line_matches_library(tmp_path):
    """The latticegen command takes the reference's arguments and reads
    STARCDfiles/<name>.vrt/.bnd from the working directory."""
    vrt, bnd = stage(str(tmp_path), "rLid3D")
    r = subprocess.run([CLI, "rLid3D", "25", "0.5", "1.5", "0.5", "cli", "2"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    L.latticegen(vrt, bnd, 25, (0.5, 1.5, 0.5), str(tmp_path / "lib"), complexity=2)
    for kind in ("Node", "BC"):
        assert open(tmp_path / f"cli_{kind}.dat", "rb").read() == open(tmp_path / f"lib_{kind}.dat", "rb").read()
    assert (tmp_path / "xtecplot.dat").exists()
    r = subprocess.run([CLI, "rLid3D", "25"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage" in r.stdout


def box_geometry(path, boxes):
    """STAR-CD files of axis-aligned boxes: (lo, hi, type) each, faces
    listed W, E, S, N, B, T."""
    verts, faces = [], []
    for lo, hi, typ in boxes:
        base = len(verts)
        for z in (lo[2], hi[2]):
            for y in (lo[1], hi[1]):
                for x in (lo[0], hi[0]):
                    verts.append((x, y, z))
        v = lambda i, j, k: base + 1 + i + 2 * j + 4 * k
        quads = [(v(0, 0, 0), v(0, 1, 0), v(0, 1, 1), v(0, 0, 1)), (v(1, 0, 0), v(1, 1, 0), v(1, 1, 1), v(1, 0, 1)),
                 (v(0, 0, 0), v(1, 0, 0), v(1, 0, 1), v(0, 0, 1)), (v(0, 1, 0), v(1, 1, 0), v(1, 1, 1), v(0, 1, 1)),
                 (v(0, 0, 0), v(1, 0, 0), v(1, 1, 0), v(0, 1, 0)), (v(0, 0, 1), v(1, 0, 1), v(1, 1, 1), v(0, 1, 1))]
        for q in quads:
            faces.append((q, len(faces) + 1, typ))
    with open(path + ".vrt", "w") as f:
        for i, (x, y, z) in enumerate(verts):
            f.write(f"{i + 1:9d}      {x:+.9e}{y:+.9e}{z:+.9e}\n")
    with open(path + ".bnd", "w") as f:
        for i, (q, fid, typ) in enumerate(faces):
            f.write(f"{i + 1:8d}       {q[0]:8d} {q[1]:8d} {q[2]:8d} {q[3]:8d} {fid:6d}      0      {typ}\n")
    return path + ".vrt", path + ".bnd"


def test_obstacle_box_fill_and_boundary_links(tmp_path):
    # delta = 1/16 and faces on nodes (exact binary coordinates): obstacle
    # faces on x nodes 4, 12, y nodes 5, 11, z nodes 6, 10
    vrt, bnd = box_geometry(str(tmp_path / "ob"), [((0, 0, 0), (1, 1, 1), "WALL"),
                                                   ((0.25, 0.3125, 0.375), (0.75, 0.6875, 0.625), "WALL")])
    n = 17
    dims, nl = L.latticegen(vrt, bnd, n, (0.1, 0.1, 0.1), str(tmp_path / "ob"), complexity=2)
    assert dims == (n, n, n)
    lat = L.read_lattice(str(tmp_path / "ob_Node.dat"), str(tmp_path / "ob_BC.dat"))
    fluid = lat["node_class"] != L.SOLID
    zz, yy, xx = np.mgrid[0:n, 0:n, 0:n]
    inside = (xx > 4) & (xx < 12) & (yy > 5) & (yy < 11) & (zz > 6) & (zz < 10)
    assert (fluid == ~inside).all()
    # boundary links = links from fluid nodes to solid or outside nodes
    c = LT.velocities(19)
    rows = np.loadtxt(tmp_path / "ob_BC.dat")
    got = {(int(r[0]), int(r[1]), int(r[2]), int(r[3])) for r in rows}
    assert len(got) == len(rows) == nl
    want = set()
    for d in range(1, 19):
        tx, ty, tz = xx + c[d, 0], yy + c[d, 1], zz + c[d, 2]
        out = (tx < 0) | (tx >= n) | (ty < 0) | (ty >= n) | (tz < 0) | (tz >= n)
        nb_solid = np.zeros_like(out)
        ok = ~out
        nb_solid[ok] = ~fluid[tz[ok], ty[ok], tx[ok]]
        for z, y, x in zip(*np.nonzero(fluid & (out | nb_solid))):
            want.add((int(x), int(y), int(z), d))
    assert got == want
    assert (rows[:, 4] == 1).all()
    # a face through nodes: the nodes on it are fluid, and their links into
    # the obstacle are cut at the node itself (on-node walls)
    ax = rows[(rows[:, 3] == 1) & (rows[:, 0] == 4)]
    assert len(ax) > 0 and np.allclose(ax[:, 5], 0.25)


def test_binary_lattice_round_trip(tmp_path):
    """A binary copy reads back to the same classes, wall links, counts,
    spacing, cyclic axes and wall fractions as the text pair (the reference's
    lid cavity, and a 2D channel with curved-wall points)."""
    node = str(tmp_path / "lid_Node.dat")
    bc = str(tmp_path / "lid_BC.dat")
    for kind, path in (("Node", node), ("BC", bc)):
        open(path, "wb").write(gzip.open(os.path.join(GOLDEN, f"lidvalidation_{kind}.dat.gz"), "rb").read())
    ny, nx = 12, 20
    ch = LT.channel2d(nx, ny, 0.03, obstacles=((8, 6, 2.5),))
    q = np.zeros((ny, nx, 19))
    q[..., 1:9] = np.where((ch["bb"][..., None] >> np.arange(1, 9)) & 1, 0.3, 0.0)
    LT.write_latticegen2(str(tmp_path / "ch"), ch["node_class"], ch["bb"], wall_q=q)
    for np_, bp in ((node, bc), (str(tmp_path / "ch_Node.dat"), str(tmp_path / "ch_BC.dat"))):
        bin_path = np_ + ".lbl"
        L.write_lattice_binary(np_, bp, bin_path)
        a, b = L.read_lattice(np_, bp), L.read_lattice(bin_path)
        for k in ("dims", "nlinks", "delta", "cyclic"):
            assert a[k] == b[k], k
        for k in ("node_class", "bb_links", "type_counts", "dir_counts"):
            assert (a[k] == b[k]).all(), k
        assert (L.read_wall_q(np_, bp) == L.read_wall_q(bin_path, None)).all()
    with open(tmp_path / "bad.lbl", "wb") as f:
        f.write(b"LBMLAT01" + b"\0" * 10)
    with pytest.raises(L.LbmError):
        L.read_lattice(str(tmp_path / "bad.lbl"))


def test_generator_rejects_bad_input(tmp_path):
    vrt, bnd = stage(str(tmp_path), "lid3D")
    out = str(tmp_path / "x")
    with pytest.raises(L.LbmError) as e:
        L.latticegen(str(tmp_path / "missing.vrt"), bnd, 10, (0.5, 0.5, 0.5), out)
    assert e.value.status == 2                       # EIO
    with pytest.raises(L.LbmError) as e:
        L.latticegen(vrt, bnd, 1, (0.5, 0.5, 0.5), out)
    assert e.value.status == 3                       # EINVAL: one node along x
    with pytest.raises(L.LbmError) as e:
        L.latticegen(vrt, bnd, 10, (0.5, 0.5, 0.5), out, complexity=3)
    assert e.value.status == 3
    with pytest.raises(L.LbmError) as e:             # no node nearer the seed than the origin is
        L.latticegen(vrt, bnd, 10, (0.0, 0.0, 0.0), out)
    assert e.value.status == 3 and "seed" in str(e.value)
    bad = tmp_path / "bad.bnd"
    bad.write_text("1 1 2 3 99 1 0 WALL\n")           # vertex 99 does not exist
    with pytest.raises(L.LbmError) as e:
        L.latticegen(vrt, str(bad), 10, (0.5, 0.5, 0.5), out)
    assert e.value.status == 2 and "vertex" in str(e.value)
