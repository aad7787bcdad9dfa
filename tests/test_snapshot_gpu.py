"""GPU snapshot writer (scihpc_vtk_* through the C ABI) against the files
vtkWriter itself wrote (golden bytes) and the oracle's writer.

Bar: byte-identical files.  Covers 1-D, 2-D and 3-D grids, tile-ragged
sizes, several snapshots from one handle, host and device fields, files
larger than one pinned chunk, and the error paths (no open file, an
unwritable path).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from golden_util import load, spacing
from scihpc_amd import ScihpcError, VtkWriter, padded_shape, read_vtk

VTK = load("vtk")


class _Hip:
    """hipMalloc / hipMemcpy / hipFree through ctypes, for device fields."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), a.nbytes) == 0
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def free(self, p):
        self.lib.hipFree(C.c_void_p(p))


def _origin_spacing(c):
    ndim = int(c["ndim"])
    ext = c["ext"]
    return [ext[2 * a] for a in range(ndim)], list(spacing(c))[:ndim]


@pytest.mark.parametrize("case", sorted(VTK), ids=str)
def test_vtk_host_fields_match_reference_bytes(tmp_path, case):
    c = VTK[case]
    ndim, n = int(c["ndim"]), [int(x) for x in c["n"]]
    org, sp = _origin_spacing(c)
    print(f"vtk host {case}", flush=True)
    w = VtkWriter(ndim, n, org, sp, str(tmp_path / "snap"))
    path = w.begin(3)
    w.add_scalar("phi", c["phi"])
    w.add_scalar("pressure", c["p"])
    w.add_vector("nvel", c["u"], c["v"], c["w"])
    w.close()
    assert open(path, "rb").read() == c["file"].tobytes()


@pytest.mark.parametrize("case", sorted(VTK), ids=str)
def test_vtk_device_fields_match_reference_bytes(tmp_path, case):
    c = VTK[case]
    ndim, n = int(c["ndim"]), [int(x) for x in c["n"]]
    org, sp = _origin_spacing(c)
    print(f"vtk device {case}", flush=True)
    hip = _Hip()
    ptrs = [hip.upload(c[k]) for k in ("phi", "p", "u", "v", "w")]
    try:
        w = VtkWriter(ndim, n, org, sp, str(tmp_path / "snap"))
        path = w.begin(3)
        w.add_scalar_device("phi", ptrs[0])
        w.add_scalar_device("pressure", ptrs[1])
        w.add_vector_device("nvel", *ptrs[2:])
        w.close()
    finally:
        for p in ptrs:
            hip.free(p)
    assert open(path, "rb").read() == c["file"].tobytes()


@pytest.mark.parametrize("ndim,n", [(2, (37, 45)), (3, (33, 7, 41)), (2, (512, 1024)), (3, (96, 96, 96))],
                         ids=["2d_ragged", "3d_ragged", "2d_512x1024", "3d_96"])
def test_vtk_matches_oracle_bytes(oracle, tmp_path, ndim, n):
    """Ragged tiles and files past one 16 MiB pinned chunk (the 3-D vector is
    21 MiB) against the oracle's writer; read back through read_vtk."""
    print(f"vtk oracle {ndim}d {n}", flush=True)
    rng = np.random.default_rng(5)
    shp = padded_shape(ndim, *n)
    phi, u, v, w_ = (rng.standard_normal(shp).ravel() for _ in range(4))
    org, sp = [0.0, -1.0, 0.5][:ndim], [1.0 / n[0], 2.0 / n[1], 0.75][:ndim]
    w = VtkWriter(ndim, n, org, sp, str(tmp_path / "g"))
    path = w.begin(0)
    w.add_scalar("phi", phi)
    w.add_vector("nvel", u, v, w_)
    w.close()
    want = str(tmp_path / "o.vtk")
    oracle.vtk_write(want, ndim, n, org + [0.0] * (3 - ndim), sp + [0.0] * (3 - ndim), [("phi", phi)],
                     [("nvel", (u, v, w_))])
    assert open(path, "rb").read() == open(want, "rb").read()
    got = read_vtk(path)
    sl = tuple(slice(3, 3 + n[a]) if a < ndim else slice(0, 1) for a in range(3))
    np.testing.assert_array_equal(got["phi"], np.transpose(phi.reshape(shp)[sl], (2, 1, 0)))


def test_vtk_several_snapshots_one_handle(tmp_path):
    """Snapshots 0..3 from one handle, the next begin closing the previous
    file, as the reference's driver loop does (create/add/close per plot)."""
    n = (24, 40)
    shp = padded_shape(2, *n)
    w = VtkWriter(2, n, (0.0, 0.0), (0.1, 0.05), str(tmp_path / "dam"))
    paths = []
    for i in range(4):
        paths.append(w.begin(i))
        w.add_scalar("phi", np.full(shp, float(i)))
        if i % 2:
            w.close()
    w.close()
    for i, p in enumerate(paths):
        got = read_vtk(p)
        assert np.all(got["phi"] == float(i))


def test_vtk_errors(tmp_path):
    n = (8, 8)
    shp = padded_shape(2, *n)
    w = VtkWriter(2, n, (0.0, 0.0), (1.0, 1.0), str(tmp_path / "missing_dir" / "x"))
    with pytest.raises(ScihpcError):
        w.add_scalar("phi", np.zeros(shp))  # no open file
    with pytest.raises(ScihpcError) as e:
        w.begin(0)  # directory does not exist
    assert "cannot open" in str(e.value)
    with pytest.raises(ValueError):
        VtkWriter(2, n, (0.0, 0.0), (1.0, 1.0), str(tmp_path / "y")).add_scalar("phi", np.zeros(10))
