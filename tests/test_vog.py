"""The VOG reader (flowpsi_amd/vog.py, flowpsi_amd/hdf5.py) on the
reference's own grids (quickTest/*.vog, copied under tests/golden/vog), no
GPU: the decoded counts equal the file's numNodes/numFaces/numCells, every
cell is closed with a positive volume, box-shaped domains add up to their
bounding-box volume and keep each boundary surface on one side of the box,
the periodic surfaces of quickTest/viscous/periodic.vog pair with the
translation its case declares, and the varint tables decode hand-made byte
strings."""
import numpy as np
import pytest

from flowpsi_amd import grids, hdf5, vog, workloads

FIXTURES = ["testGrid", "shock_tube", "nozzle", "periodic"]


def test_varint_tables():
    # size 3: first entry -6 (0x40 sign), then +154 (0x9A 0x01) and +1
    t, p = vog._table(np.array([3, 0x46, 0x9A, 0x01, 0x01, 99], np.uint8), 0)
    assert t.tolist() == [-6, 148, 149] and p == 5
    # a multi-byte signed first entry: 6 + (1 << 6) = 70, positive
    t, p = vog._table(np.array([1, 0x86, 0x01], np.uint8), 0)
    assert t.tolist() == [70] and p == 3
    # size byte 0 means 256 entries
    t, p = vog._table(np.array([0, 0] + [1] * 255, np.uint8), 0)
    assert len(t) == 256 and t[-1] == 255


@pytest.mark.parametrize("name", FIXTURES)
def test_reference_grids_decode(name):
    path = workloads.vog_path(name)
    f = hdf5.File(path)
    info = f.attrs("file_info")
    v = vog.read(path)
    assert len(v["pos"]) == int(info["numNodes"][0])
    assert len(v["face_off"]) - 1 == int(info["numFaces"][0])
    assert v["nc"] == int(info["numCells"][0])
    g = vog.to_grid(v, {k: grids.SYMMETRY for k in v["surfaces"]}, name=name)
    g.check()
    assert np.all(g.vol > 0)
    # every cell has at least four faces; every node is used
    nfc = np.bincount(g.cl, minlength=g.nc) + np.bincount(g.cr[g.cr >= 0], minlength=g.nc)
    assert nfc.min() >= 4
    assert len(np.unique(v["face_nodes"])) == len(v["pos"])
    if name in ("shock_tube", "testGrid"):
        lo, hi = v["pos"].min(0), v["pos"].max(0)
        assert g.vol.sum() == pytest.approx(np.prod(hi - lo), rel=1e-8)
        tag = g.meta["face_tag"]
        for t in np.unique(tag[g.btype != grids.INTERIOR]):
            fc = g.fc[:, tag == t]
            on_side = [(np.abs(fc[d] - lo[d]).max() < 1e-7 or np.abs(fc[d] - hi[d]).max() < 1e-7) for d in range(3)]
            assert any(on_side), f"surface {t} is not one side of the box"


def test_shock_tube_is_a_row_of_hexes():
    g, q, prm = workloads.vog_case("shock_tube")
    assert g.nc == 200 and np.allclose(g.vol, g.vol[0], rtol=1e-12)
    assert np.all(np.bincount(g.cl, minlength=g.nc) + np.bincount(g.cr[g.cr >= 0], minlength=g.nc) == 6)
    assert (q[4] == 1e6).sum() == 100


def test_periodic_surfaces_pair_with_the_declared_translation():
    # quickTest/viscous/periodic*.vars: top = periodic(name="A", translate=[0,-2,0]), bottom = periodic(name="A")
    g, q, prm = workloads.vog_case("periodic")
    per = np.flatnonzero(g.btype == grids.PERIODIC)
    assert len(per) == 52
    top = g.meta["surfaces"]["top"]
    for f in per:
        # translate carries top onto bottom: a top face's ghost is its bottom partner's cell moved by -translate
        want = np.array([0.0, 2.0, 0.0]) if g.meta["face_tag"][f] == top else np.array([0.0, -2.0, 0.0])
        assert np.allclose(g.pshift[:, f], want, atol=1e-12)
        assert np.allclose(g.area[f], g.area[g.pface[f]], rtol=1e-12)
