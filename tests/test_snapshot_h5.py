"""HDF5 field snapshot files (Bubble::OutField, src/bubble.cpp:883-954) written
from host arrays through cb_write_snapshot_h5 and read back through HDF5's
own filter pipeline: datasets, shapes, chunking, filters and attributes as
the reference's writer declares them (:889-948)."""
import os

import numpy as np
import pytest

import coolbubble_amd as cb
import h5read

pytestmark = pytest.mark.skipif(h5read.lib() is None, reason="HDF5 C library not installed")

NAMES5 = ["phi", "pi", "Ekin", "gradphi", "potEnergy"]


def expected_chunk(D, n):
    """Whole rows per chunk: 1D the field, 2D min(n, max(1, 65536 // n)) rows,
    3D one plane (src/bubble.cpp:920-925)."""
    if D == 1:
        return (n,)
    if D == 2:
        return (min(n, max(1, 65536 // n)), n)
    return (1, n, n)


@pytest.mark.parametrize("D,N,reduc,n", [(1, 4096, 1, 4096), (2, 2048, 1, 2048), (2, 2049, 1, 2049),
                                         (2, 4100, 2, 2050), (2, 8192, 4, 2048), (2, 10000, 4, 2500),
                                         (3, 512, 1, 512), (3, 513, 1, 513), (3, 1025, 2, 513),
                                         (3, 2048, 4, 512)])
def test_saved_points(D, N, reduc, n):
    assert cb.snapshot_points(D, N) == (reduc, n)
    # thinning keeps lattice points 0, reduc, 2 reduc, ... below N on each axis
    assert len(range(0, N, reduc)) == n


@pytest.mark.parametrize("D,N,nf", [(1, 300, 5), (2, 64, 5), (2, 700, 2), (2, 4100, 2), (3, 20, 5), (3, 33, 2)])
def test_write_read_back(tmp_path, D, N, nf):
    reduc, n = cb.snapshot_points(D, N)
    rng = np.random.default_rng(D * 1000 + N)
    fields = [rng.standard_normal((n,) * D).astype(np.float32) * (f + 1) for f in range(nf)]
    fields[0][(0,) * D] = np.float32(-0.0)
    fields[1].reshape(-1)[-1] = np.float32(np.finfo(np.float32).tiny / 4)   # a subnormal
    path = str(tmp_path / "_phi_t=0003.h5")
    cb.write_snapshot_h5(path, fields, D, N, L=32.0, c=4.0, t=1.25, a=1.5)
    with h5read.File(path) as f:
        assert sorted(f.names()) == sorted(NAMES5[:nf])
        assert f.attrs() == {"D": D, "L": 32.0, "nGrids": N, "reduc": reduc, "c": 4.0, "t": 1.25, "a": 1.5}
        for name, want in zip(NAMES5, fields):
            got, (esize, tclass), chunk, filters = f.dataset(name)
            assert got.shape == (n,) * D
            assert (esize, tclass) == (4, 1)                 # 4-byte H5T_FLOAT
            np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
            assert chunk == expected_chunk(D, n)
            assert filters == [(2, (4,)), (1, (9,))]         # shuffle (4-byte elements), then deflate 9


def test_deflate_level_and_compression(tmp_path):
    D, N = 2, 512
    x = np.linspace(0, 2 * np.pi, N, endpoint=False)
    smooth = np.sin(x)[:, None] * np.cos(x)[None, :]
    fields = [smooth.astype(np.float32), (0.5 * smooth).astype(np.float32)]
    sizes = {}
    for level in (1, 9):
        path = str(tmp_path / ("lvl%d.h5" % level))
        cb.write_snapshot_h5(path, fields, D, N, L=8.0, c=2.0, t=0.0, a=1.0, deflate=level)
        with h5read.File(path) as f:
            got, _, _, filters = f.dataset("pi")
            np.testing.assert_array_equal(got, fields[1])
            assert filters == [(2, (4,)), (1, (level,))]
        sizes[level] = os.path.getsize(path)
    assert sizes[9] <= sizes[1] < 2 * N * N * 4


def test_invalid_arguments(tmp_path):
    f = np.zeros((8, 8), dtype=np.float32)
    with pytest.raises(ValueError):
        cb.write_snapshot_h5(str(tmp_path / "x.h5"), [f] * 3, 2, 8, 1.0, 1.0, 0.0, 1.0)
    with pytest.raises(ValueError):
        cb.write_snapshot_h5(str(tmp_path / "x.h5"), [f, f[:4]], 2, 8, 1.0, 1.0, 0.0, 1.0)
    with pytest.raises(cb.CBError):
        cb.write_snapshot_h5(str(tmp_path / "x.h5"), [f, f], 2, 8, 1.0, 1.0, 0.0, 1.0, deflate=10)
    with pytest.raises(cb.CBError):
        cb.write_snapshot_h5(str(tmp_path / "no_such_dir" / "x.h5"), [f, f], 2, 8, 1.0, 1.0, 0.0, 1.0)
