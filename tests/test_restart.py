"""Restart files and the HDF5 readers, no GPU: flowpsi_amd/hdf5.py reads
every dataset of the VOG fixtures exactly as libhdf5 does; restart files
written through libhdf5 read back bit for bit (by both readers), carry the
time level n-1 for BDF2 only, and shift the gage pressure by the difference
of the ambient pressures as src/solverIC.loci:594-601 does."""
import numpy as np
import pytest

from flowpsi_amd import h5lib, hdf5, restart, workloads

try:
    h5lib.lib()
    HAVE_LIB = True
except OSError:
    HAVE_LIB = False
need_lib = pytest.mark.skipif(not HAVE_LIB, reason="libhdf5 not installed")


@need_lib
@pytest.mark.parametrize("name", ["testGrid", "shock_tube", "nozzle", "periodic"])
def test_reader_agrees_with_libhdf5(name):
    path = workloads.vog_path(name)
    f = hdf5.File(path)
    for ds, dt in (("face_info/cluster_info", np.uint8), ("face_info/cluster_sizes", np.uint16)):
        assert np.array_equal(f[ds], h5lib.read_dataset(path, ds, dt))
    P = f["node_info/positions"]
    Q = h5lib.read_dataset(path, "node_info/positions", np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8")]))
    for k in "xyz":
        assert np.array_equal(P[k], Q[k])


class _Stub:
    """The Solver calls restart uses, on host arrays."""

    def __init__(self, g, prm, q, qm1=None):
        self.grid, self.prm, self.q, self.qm1 = g, prm, q, qm1

    def get_state(self):
        return self.q.copy()

    def get_time_levels(self):
        return self.q.copy(), (self.qm1 if self.qm1 is not None else self.q).copy()

    def set_state(self, q):
        self.q, self.qm1 = np.array(q), None

    def set_time_levels(self, qn, qnm1):
        self.qm1 = np.array(qnm1)


@need_lib
@pytest.mark.parametrize("turb,unsteady", [("none", False), ("sa", True), ("sst", True), ("sst", False)])
def test_round_trip(tmp_path, turb, unsteady):
    if turb == "none":
        g, q, prm = workloads.box(4)
    else:
        g, q, prm = (workloads.channel_sa if turb == "sa" else workloads.channel_sst)(nx=4, ny=6, nz=4)
    prm = dict(prm, unsteady=unsteady, time_scheme="bdf2", Pambient=0.0)
    qm1 = q * (1 + 1e-3 * np.random.default_rng(1).standard_normal(q.shape))
    restart.write(str(tmp_path), "case", _Stub(g, prm, q, qm1), ncycle=17, stime=0.25)
    files = sorted(p.name for p in tmp_path.iterdir())
    prev = [x for x in files if "Prev" in x]
    assert bool(prev) == unsteady
    dst = _Stub(g, prm, np.zeros_like(q))
    info = restart.read(str(tmp_path), "case", dst)
    assert info == {"ncycle": 17, "stime": 0.25}
    assert np.array_equal(dst.q, q)
    if unsteady:
        assert np.array_equal(dst.qm1, qm1)
    # libhdf5 reads the same values
    T = h5lib.read_dataset(str(tmp_path / "flowVars_case"), "temperature/data", np.float64)
    u = h5lib.read_dataset(str(tmp_path / "flowVars_case"), "u/data", np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8")]))
    assert np.array_equal(T, q[0]) and np.array_equal(u["y"], q[2])
    iv = h5lib.read_dataset(str(tmp_path / "flowVars_case"), "temperature/Interval Set", np.int64)
    assert iv.tolist() == [[0, g.nc - 1]]


@need_lib
def test_ambient_pressure_shift(tmp_path):
    g, q, prm = workloads.box(3)
    restart.write(str(tmp_path), "c", _Stub(g, dict(prm, Pambient=1000.0), q))
    dst = _Stub(g, dict(prm, Pambient=250.0), np.zeros_like(q))
    restart.read(str(tmp_path), "c", dst)
    assert np.allclose(dst.q[4], q[4] + 750.0, rtol=0, atol=1e-9)
    assert np.array_equal(dst.q[:4], q[:4])
