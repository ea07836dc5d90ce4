"""CPU pins for the marching-cubes case table, the oracle's marching cubes and
the reproducible exp / sin / cos used by the float path."""
import os
import re

import numpy as np
import pytest

import oracle as O

REF_TABLE = "/root/reference/src/SLAM/VisualOdometry/DeviceSpecific/OpenCL/NuiMarchingCubeTable.h"


def table():
    tri = np.empty((256, 16), np.int8)
    nv = np.empty(256, np.uint8)
    O.lib().or_mc_table.argtypes = [np.ctypeslib.ndpointer(np.int8), np.ctypeslib.ndpointer(np.uint8)]
    O.lib().or_mc_table(tri, nv)
    return tri, nv


CORNERS = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def test_mc_table_invariants():
    tri, nv = table()
    assert nv[0] == 0 and nv[255] == 0
    for c in range(256):
        n = int(nv[c])
        assert n % 3 == 0 and n <= 15
        assert (tri[c, :n] >= 0).all() and (tri[c, n:] == -1).all()
        # a triangle vertex sits on an edge whose two corners have opposite signs
        for e in tri[c, :n]:
            a, b = CORNERS[e]
            assert ((c >> a) & 1) != ((c >> b) & 1), (c, e)
        # every sign-changing edge carries at least one vertex (the surface is not dropped)
        crossed = {e for e, (a, b) in enumerate(CORNERS) if ((c >> a) & 1) != ((c >> b) & 1)}
        assert crossed == set(tri[c, :n].tolist()), c
    # one sign pattern with a single corner inside: one triangle around that corner
    for k in range(8):
        assert nv[1 << k] == 3 and nv[255 - (1 << k)] == 3


@pytest.mark.skipif(not os.path.exists(REF_TABLE), reason="reference sources not mounted")
def test_mc_table_equals_reference_table():
    src = open(REF_TABLE).read()
    block = src[src.index("sTriTable"):src.index("sNumVertsTable")]
    rows = re.findall(r"\{([-0-9,\s]+)\}", block)
    ref = np.array([[int(v) for v in r.replace(" ", "").split(",") if v] for r in rows], np.int8)
    nvb = src[src.index("sNumVertsTable"):src.index("sEdgeTable")]
    ref_nv = np.array([int(v) for v in re.findall(r"-?\d+", nvb[nvb.index("{"):])][:256])
    tri, nv = table()
    np.testing.assert_array_equal(tri, ref)
    np.testing.assert_array_equal(nv, ref_nv)


@pytest.mark.skipif(not os.path.exists(REF_TABLE), reason="reference sources not mounted")
def test_edge_masks_equal_reference_edge_table():
    """The hash mesher takes a cube's edge mask from the edges its triangles use;
    that is the reference's sEdgeTable (OCL/NuiMarchingCubeTable.h:542) entry by entry."""
    src = open(REF_TABLE).read()
    blk = src[src.index("sEdgeTable"):]
    ref = [int(v, 16) for v in re.findall(r"0x[0-9a-fA-F]+", blk[blk.index("{"):blk.index("}")])]
    edge, _ = O.mc_tables()
    assert len(ref) == 256
    np.testing.assert_array_equal(edge, np.array(ref, np.int32))


def sphere_volume(N=32, r=0.3, dim=1.0):
    c = dim / N
    y, x, z = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    X = (x - N // 2 + 0.5) * c
    Y = (y - N // 2 + 0.5) * c
    Z = (z + 0.5) * c - dim / 2
    f = np.clip((np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - r) / (3 * c), -1, 1)
    vol = np.empty((N ** 3, 2), np.int16)
    vol[:, 0] = np.trunc(f.reshape(-1) * 32766).astype(np.int16)
    vol[:, 1] = 1
    return vol


def test_mc_sphere_count_is_closed_surface_sum():
    N = 32
    vol = sphere_volume(N)
    total = O.lib().or_mc_count_s16(vol.reshape(-1), N)
    # brute force: sum of the table counts over all cubes
    tri, nv = table()
    f = vol[:, 0].reshape(N, N, N)
    neg = (f < 0).astype(np.int32)
    ci = (neg[:-1, :-1, :-1] + 2 * neg[:-1, 1:, :-1] + 4 * neg[1:, 1:, :-1] + 8 * neg[1:, :-1, :-1] +
          16 * neg[:-1, :-1, 1:] + 32 * neg[:-1, 1:, 1:] + 64 * neg[1:, 1:, 1:] + 128 * neg[1:, :-1, 1:])
    assert total == int(nv[ci].sum()) > 0
    # a zero weight anywhere removes the 8 cubes around that voxel
    vol2 = vol.copy()
    vol2[((16 * N) + 16) * N + 16 - 5, 1] = 0
    assert O.lib().or_mc_count_s16(vol2.reshape(-1), N) <= total


def test_mc_emit_on_sphere_surface():
    import ctypes as C
    N, r, dim = 48, 0.3, 1.0
    vol = sphere_volume(N, r, dim)
    lib = O.lib()
    lib.or_mc_emit_s16.restype = C.c_longlong
    lib.or_mc_emit_s16.argtypes = [np.ctypeslib.ndpointer(np.int16), C.c_void_p, C.c_int, C.c_float,
                                   np.ctypeslib.ndpointer(np.float32), C.c_void_p, C.c_longlong]
    total = lib.or_mc_count_s16(vol.reshape(-1), N)
    verts = np.empty((total, 3), np.float32)
    n = lib.or_mc_emit_s16(vol.reshape(-1), None, N, dim, verts, None, total)
    assert n == total and n % 3 == 0
    # vertices lie on the sphere (centre (0, 0, dim/2) in the mesher's frame, y flipped)
    d = np.sqrt(verts[:, 0] ** 2 + verts[:, 1] ** 2 + (verts[:, 2] - dim / 2) ** 2)
    assert np.abs(d - r).max() < dim / N * 0.3


def test_mc_dedup_mesh_counts():
    import ctypes as C
    N = 32
    vol = sphere_volume(N)
    lib = O.lib()
    lib.or_mc_mesh_dedup_s16.restype = C.c_longlong
    lib.or_mc_mesh_dedup_s16.argtypes = [np.ctypeslib.ndpointer(np.int16), C.c_int, C.c_float,
                                         np.ctypeslib.ndpointer(np.float32), C.c_longlong,
                                         np.ctypeslib.ndpointer(np.int32), C.c_longlong,
                                         C.POINTER(C.c_longlong)]
    total = lib.or_mc_count_s16(vol.reshape(-1), N)
    pts = np.empty((total, 3), np.float32)
    idx = np.empty(total, np.int32)
    ni = C.c_longlong()
    npts = lib.or_mc_mesh_dedup_s16(vol.reshape(-1), N, 1.0, pts, total, idx, total, C.byref(ni))
    assert ni.value == total  # one index per triangle corner
    assert 0 < npts < total  # shared vertices are merged
    assert idx.max() == npts - 1 and idx.min() == 0


# ---------------------------------------------------------------- reproducible math

def test_expf_det_is_correctly_rounded():
    import ctypes as C
    lib = O.lib()
    lib.or_expf_det.restype = C.c_float
    lib.or_expf_det.argtypes = [C.c_float]
    rng = np.random.default_rng(5)
    xs = np.concatenate([-np.abs(rng.normal(0, 3, 4000)), rng.uniform(-87, 88, 2000), [0.0, -1e-8, 1e-8, -50.0]])
    xs = xs.astype(np.float32)
    for x in xs:
        got = np.float32(lib.or_expf_det(float(x)))
        true = np.exp(np.float64(x))
        # within half an ulp of the exact value (the correctly rounded float)
        ulp = np.float64(np.spacing(np.float32(true)))
        assert abs(np.float64(got) - true) <= 0.5 * ulp * (1 + 1e-6), (x, got, true)


def test_sincosf_det_is_correctly_rounded():
    import ctypes as C
    lib = O.lib()
    lib.or_sincosf_det.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    rng = np.random.default_rng(6)
    xs = np.concatenate([rng.normal(0, 0.05, 3000), rng.uniform(-10, 10, 2000), [0.0, 1e-9, -0.7853982]])
    s, c = C.c_float(), C.c_float()
    for x in xs.astype(np.float32):
        lib.or_sincosf_det(float(x), C.byref(s), C.byref(c))
        for got, true in ((s.value, np.sin(np.float64(x))), (c.value, np.cos(np.float64(x)))):
            tol = 0.5 * np.float64(np.spacing(np.float32(abs(true)))) * (1 + 1e-6) + 1e-45
            assert abs(np.float64(np.float32(got)) - true) <= tol, (x, got, true)
