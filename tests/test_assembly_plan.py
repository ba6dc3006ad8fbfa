"""Host-side checks of the element assembly plan (no GPU): coloured patches, coloured elements, atomics.

Invariants (sf2d_plan_validate): every listed element appears once in the device order; no two units
(patches or elements) of one colour share a GLL point; patch slots address their patch's point list;
patches hold at most 10 elements and 256 points; the patch mode's internal point numbering is a
permutation in which every patch owns one contiguous id range (ranges disjoint) and lists, sorted, only
points owned by other patches as its halo.  On structured meshes a 10-element strip owns its upper four
GLL rows (160 points) and reads 45 halo points (the row below and the column to its left)."""
import ctypes as C

import numpy as np
import pytest

from oracle.oracle import OracleSim
from specfem2d_amd import config, driver


def validate(ibool, nglob, lists, mode):
    gpu, _ = driver.load_libraries()
    nspec = ibool.shape[0]
    np_ = lists.shape[1]
    ib = np.ascontiguousarray(ibool.reshape(-1), dtype=np.int32)
    ph = np.ascontiguousarray(lists.reshape(-1), dtype=np.int32)
    stats = np.zeros(6, np.int32)
    f = gpu.sf2d_plan_validate
    f.restype = C.c_int
    rc = f(C.c_int(nspec), C.c_int(nglob), ib.ctypes.data_as(C.c_void_p), C.c_int(np_),
           ph.ctypes.data_as(C.c_void_p), C.c_int(mode), stats.ctypes.data_as(C.c_void_p))
    return rc, stats


def inner_only(nspec):
    lists = np.zeros((2, nspec), np.int32)
    lists[1] = np.arange(1, nspec + 1)
    return lists


MESHES = {
    "tromp": lambda: config.tromp2005_kernel(nstep=2),
    "c1": lambda: config.plain_elastic_c1(nstep=2),
    "layers": lambda: config.small_test_mesh(nx=13, nz=10, nstep=2).replace(
        interfaces=[-1000.0, -620.0, -330.0, 0.0], nz_layers=[3, 4, 3], xmax=1390.0),
}


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_plan_invariants_structured(name, mode):
    cfg = MESHES[name]()
    o = OracleSim(cfg, "sp")
    rc, stats = validate(o.ibool(), o.nglob, inner_only(o.nspec), mode)
    assert rc == 0
    if mode in (0, 1):
        assert stats[0] == 0 and stats[1] == 4  # 2 x 2 checkerboard of elements or of row patches
    if mode == 0:
        npatch = sum((cfg.nx + 9) // 10 for _ in range(cfg.nz))  # row segments of <= 10 elements
        assert stats[2] == npatch
        assert stats[3] <= 41 * 5
        assert stats[5] == o.nglob  # every point owned by exactly one patch
        # halo of each strip: the GLL row below it and the column to its left, where those exist
        expect = 0
        for ez in range(cfg.nz):
            for x0 in range(0, cfg.nx, 10):
                w = min(10, cfg.nx - x0)
                if ez > 0:
                    expect += 4 * w + 1 + (4 if x0 > 0 else 0)  # row below (with its corner), left column
                elif x0 > 0:
                    expect += 5  # bottom strip: the whole left column
        assert stats[4] == expect


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_plan_invariants_shuffled_and_split_lists(mode):
    # an arbitrary element order and an outer/inner split, as a partitioned mesh would pass
    cfg = config.small_test_mesh(nx=17, nz=11, nstep=2)
    o = OracleSim(cfg, "sp")
    rng = np.random.default_rng(4)
    perm = rng.permutation(o.nspec) + 1
    nout = 37
    lists = np.zeros((2, o.nspec - nout), np.int32)
    lists[0, :nout] = perm[:nout]
    lists[1, :] = perm[nout:]
    rc, stats = validate(o.ibool(), o.nglob, lists, mode)
    assert rc == 0
    if mode != 2:
        assert stats[0] >= 1 and stats[1] >= 4


def chains(ibool, nglob, lists, first=None):
    gpu, _ = driver.load_libraries()
    ib = np.ascontiguousarray(ibool.reshape(-1), dtype=np.int32)
    ph = np.ascontiguousarray(lists.reshape(-1), dtype=np.int32)
    f = gpu.sf2d_plan_chains
    f.restype = C.c_int
    nf = C.c_int(-1)
    rc = f(C.c_int(ibool.shape[0]), C.c_int(nglob), ib.ctypes.data_as(C.c_void_p), C.c_int(lists.shape[1]),
           ph.ctypes.data_as(C.c_void_p), C.byref(nf))
    if first is not None:
        first.append(nf.value)
    return rc


@pytest.mark.parametrize("nx,expect", [(80, 1), (12, 1), (13, 1), (10, 1), (8, 0), (1, 0)])
def test_chain_patches_on_structured_meshes(nx, expect):
    """Rows of at least ten elements split into strips that are chains (the row-lane force kernel applies);
    narrower meshes make patches that wrap into the next row (a 2-D block) and fall back to the general
    patch kernel.  The chain property itself is re-derived here from ibool: along a strip of consecutive
    elements of one row, element k's points (4, j) are element k+1's points (0, j)."""
    cfg = config.small_test_mesh(nx=nx, nz=4, nstep=2)
    o = OracleSim(cfg, "sp")
    ib = o.ibool()
    first = []
    assert chains(ib, o.nglob, inner_only(o.nspec), first) == expect
    # first-touch assembly: exactly one update per point is the step's first
    assert first[0] == (o.nglob if expect else 0)
    if expect:
        e = np.arange(o.nspec).reshape(cfg.nz, cfg.nx)
        left, right = e[:, 1:].reshape(-1), e[:, :-1].reshape(-1)
        b = ib.reshape(-1, 5, 5)
        np.testing.assert_array_equal(b[right][:, :, 4], b[left][:, :, 0])
    # a shuffled element list breaks the strips
    rng = np.random.default_rng(3)
    lists = inner_only(o.nspec)
    lists[1] = rng.permutation(lists[1])
    assert chains(ib, o.nglob, lists) == 0


def row_records(ibool, nglob, lists):
    gpu, _ = driver.load_libraries()
    ib = np.ascontiguousarray(ibool.reshape(-1), dtype=np.int32)
    ph = np.ascontiguousarray(lists.reshape(-1), dtype=np.int32)
    f = gpu.sf2d_plan_row_records
    f.restype = C.c_int
    n = C.c_int(-1)
    rc = f(C.c_int(ibool.shape[0]), C.c_int(nglob), ib.ctypes.data_as(C.c_void_p), C.c_int(lists.shape[1]),
           ph.ctypes.data_as(C.c_void_p), C.byref(n))
    return rc, n.value


@pytest.mark.parametrize("nx,expect", [(80, 1), (12, 1), (13, 1), (10, 1), (8, -1)])
def test_chain_row_records(nx, expect):
    """The column-lane kernel's point ids as one record per chain row: in the plan's numbering every GLL row of a
    chain is numbered consecutively (column x > 0 at base + x), so {base, id of column 0, first-touch mask} per row
    reproduces the id and first-touch flag of every writing lane (checked point by point on the host); meshes whose
    patches are not chains have no records."""
    cfg = config.small_test_mesh(nx=nx, nz=4, nstep=2)
    o = OracleSim(cfg, "sp")
    rc, n = row_records(o.ibool(), o.nglob, inner_only(o.nspec))
    assert rc == expect
    if expect == 1:
        chains_per_row = -(-nx // 12) if nx >= 12 else 1  # rows narrower than 12 elements: one chain of 10
        assert n == 5 * chains_per_row * cfg.nz
