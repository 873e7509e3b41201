"""Legacy-VTK output (amr_tree_write_vtk, the reference's
include/ndtree/vtk_print.hpp:53-254): the file is parsed back and checked
against the tree -- header lines, one quad / voxel per interior cell with its
corners at the cell's geometric corners (checked against the oracle's cell
centres and sizes), the connectivity, cell types, cell_index, the is_halo
flags as the reference computes them, and every field bit for bit."""
import sys

import numpy as np
import pytest

from amrlib import load_package
from oracle import OracleTree

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def amr():
    return load_package()


def parse_vtk(path):
    data = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        s = data[pos:end].decode()
        pos = end + 1
        return s

    def array(count, dtype):
        nonlocal pos
        a = np.frombuffer(data, dtype=dtype, count=count, offset=pos)
        pos += a.nbytes
        return a

    out = {"header": [line() for _ in range(4)]}
    tok = line().split()
    assert tok[0] == "POINTS" and tok[2] == "double"
    out["points"] = array(int(tok[1]) * 3, ">f8").reshape(-1, 3)
    tok = line().split()
    assert tok[0] == "CELLS"
    out["cells"] = array(int(tok[2]), ">i4")
    out["ncells"] = int(tok[1])
    tok = line().split()
    assert tok[0] == "CELL_TYPES"
    out["types"] = array(int(tok[1]), ">i4")
    tok = line().split()
    assert tok == ["CELL_DATA", str(out["ncells"])]
    out["scalars"] = {}
    while pos < len(data):
        tok = line().split()
        assert tok[0] == "SCALARS" and tok[3] == "1"
        assert line() == "LOOKUP_TABLE default"
        out["scalars"][tok[1]] = (tok[2], array(out["ncells"], ">i4" if tok[2] == "int" else ">f8"))
    return out


def expected_halo(rank, n, h):
    """is_halo as vtk_print computes it (vtk_print.hpp:217-227): the interior
    element index j read as a padded linear index."""
    ext = (n,) * rank if isinstance(n, int) else tuple(n)
    P = tuple(x + 2 * h for x in ext)
    flags = []
    for j in range(int(np.prod(ext))):
        idx = np.unravel_index(j, P)
        flags.append(int(any(i < h or i >= p - h for i, p in zip(idx, P))))
    return np.array(flags)


def check(tree, path, rank, n, h, names, oracle):
    v = parse_vtk(path)
    assert v["header"] == ["# vtk DataFile Version 3.0", "AMR Tree Structure", "BINARY", "DATASET UNSTRUCTURED_GRID"]
    ext = (n,) * rank if isinstance(n, int) else tuple(n)
    leaves, ni, box = tree.size(), int(np.prod(ext)), 1 << rank
    cells = leaves * ni
    assert v["ncells"] == cells and v["points"].shape == (cells * box, 3)
    # connectivity, types, cell_index
    conn = v["cells"].reshape(cells, box + 1)
    assert np.all(conn[:, 0] == box)
    np.testing.assert_array_equal(conn[:, 1:], np.arange(cells * box).reshape(cells, box))
    assert np.all(v["types"] == (9 if rank == 2 else 11))
    assert v["scalars"]["cell_index"][0] == "int"
    np.testing.assert_array_equal(v["scalars"]["cell_index"][1], np.arange(cells))
    np.testing.assert_array_equal(v["scalars"]["is_halo"][1], np.tile(expected_halo(rank, n, h), leaves))
    # geometry: corners around the oracle's cell centres, VTK quad / voxel order
    cen = oracle.cell_centers().reshape(cells, rank)
    keys = oracle.keys()
    lv = (keys & np.uint64(63)).astype(int)
    L = np.array(tree.lengths)
    per_axis = np.array([ext[rank - 1 - a] for a in range(rank)], dtype=np.float64)  # cells along physical axis a
    cs = np.repeat((L[None, :] / (2.0 ** lv[:, None])) / per_axis[None, :], ni, axis=0)
    pts = v["points"].reshape(cells, box, 3)
    lo = cen - 0.5 * cs
    for k in range(box):
        bits = [(k >> d) & 1 for d in range(rank)] if rank == 3 else [[0, 0], [1, 0], [1, 1], [0, 1]][k]
        exp = lo + np.array(bits)[None, :] * cs
        np.testing.assert_allclose(pts[:, k, :rank], exp, rtol=0, atol=1e-12 * L.max())
    if rank == 2:
        assert np.all(pts[:, :, 2] == 0.0)
    # fields, bit for bit and in order
    u = tree.get_interior().reshape(len(names), -1)
    assert [k for k in v["scalars"] if k not in ("cell_index", "is_halo")] == names
    for f, name in enumerate(names):
        typ, vals = v["scalars"][name]
        assert typ == "double"
        assert np.array_equal(vals.astype(np.float64).view(np.uint64), u[f].view(np.uint64))


def flags_every_third(keys, depth):
    f = np.zeros(keys.size, dtype=np.int8)
    f[::3] = 1
    return f


@pytest.mark.parametrize("rank,n,h,depth", [(2, 16, 1, 7), (2, 10, 2, 7), (3, 8, 1, 5)])
def test_vtk_advection(amr, rank, n, h, depth, tmp_path):
    say(f"case: VTK advection rank={rank} n={n} h={h}")
    L = (2.0, 1.0, 0.5)[:rank]
    g = amr.NdTree(rank, depth, n, h, 600, L)
    o = OracleTree(rank, n, h, depth, 1400, L)
    for _ in range(2 if rank == 2 else 1):
        g.refine_all()
        o.refine_all()
    fl = flags_every_third(g.keys(), depth)
    g.reconstruct_tree(fl)
    o.reconstruct(fl)
    np.testing.assert_array_equal(g.keys(), o.keys())
    g.set_interior(np.random.default_rng(3).standard_normal(g.size() * n**rank))
    path = tmp_path / "tree.vtk"
    g.write_vtk(path)
    check(g, path, rank, n, h, ["Scalar"], o)


def test_vtk_euler(amr, tmp_path):
    say("case: VTK Euler 2D and 3D")
    for rank, n, depth in ((2, 8, 7), (3, 4, 5)):
        s = amr.EulerSolver(rank, depth, n, 1, 600, (1000.0,) * rank)
        o = OracleTree(rank, n, 1, depth, 1400, (1000.0,) * rank)
        s.tree.refine_all()
        o.refine_all()
        fl = flags_every_third(s.tree.keys(), depth)
        s.tree.reconstruct_tree(fl)
        o.reconstruct(fl)

        def ic(c):
            p = np.zeros(c.shape[:-1] + (rank + 2,))
            p[..., 0] = 1.0 + 0.1 * np.sin(c[..., 0] / 100.0)
            p[..., 1] = 0.2
            p[..., rank + 1] = 1.0 + 0.05 * np.cos(c[..., 1] / 100.0)
            return p

        s.initialize(ic)
        path = tmp_path / f"euler{rank}.vtk"
        s.tree.write_vtk(path)
        names = ["Rho", "Rhou", "Rhov", "E2D"] if rank == 2 else ["Rho", "Rhou", "Rhov", "Rhow", "E3D"]
        check(s.tree, path, rank, n, 1, names, o)
