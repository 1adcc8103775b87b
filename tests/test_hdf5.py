"""Chombo AMR hierarchy files (chombo_hdf5.py): the reader against files written by
Chombo with the HDF5 library (the reference's own test fixtures), the writer against
the reader and against the structure of those files.

Pins, each failing for a plausible decode mistake:
- lib/test/BoxTools/meshRefineTest.{1,2,3}d[.FLOAT].H5 (written by meshRefineTest
  through CH_HDF5's write of a BoxLayout): every level parses, SpaceDim matches the
  file name, the FLOAT and double files hold the same grids, boxes are disjoint and
  each finer level lies inside the refined coarser one (ratio 2).
- lib/util/SelectRegion/in.2d.H5 (an AMR checkpoint of four conserved components on
  three levels, refinement ratios 4 and 2; 4-byte HDF5 lengths): every covered coarse
  cell equals the average of its fine cells to round-off, for every component and
  both level pairs -- a swapped axis, a wrong component stride or a wrong box offset
  breaks it.
- Round trips: the writer's files read back bitwise, 2-D and 3-D, with ghost cells;
  rewriting in.2d.H5 gives the same tree (groups, attribute names and types, dataset
  types and shapes, data) as the file Chombo wrote.
"""
import os

import numpy as np
import pytest

import chombo_hdf5 as C
import hierarchy as H

REF = "/root/reference"
MESH = os.path.join(REF, "lib", "test", "BoxTools")
SELECT = os.path.join(REF, "lib", "util", "SelectRegion", "in.2d.H5")
needs_ref = pytest.mark.skipif(not os.path.isdir(REF), reason="the reference checkout is not on this machine")


def _boxes(g, dim):
    return [C._box_from(r, dim) for r in g.members["boxes"].read()]


def _disjoint(boxes):
    for a in range(len(boxes)):
        for b in range(a + 1, len(boxes)):
            if all(max(boxes[a][0][d], boxes[b][0][d]) <= min(boxes[a][1][d], boxes[b][1][d])
                   for d in range(len(boxes[a][0]))):
                return False
    return True


@needs_ref
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_read_mesh_refine_fixtures(dim):
    root = C.read_file(os.path.join(MESH, f"meshRefineTest.{dim}d.H5"))
    rootf = C.read_file(os.path.join(MESH, f"meshRefineTest.{dim}d.FLOAT.H5"))
    assert root.members["Chombo_global"].attrs["SpaceDim"] == dim
    levels = sorted(k for k in root.members if k.startswith("level_"))
    assert levels == ["level_0", "level_1", "level_2"]
    prev = None
    for lev in levels:
        b = _boxes(root.members[lev], dim)
        assert b == _boxes(rootf.members[lev], dim)
        assert _disjoint(b)
        assert len(root.members[lev].members["Processors"].read()) == len(b)
        if prev is not None:
            # proper containment in the refined coarser level (ratio 2)
            for (lo, hi) in b:
                clo, chi = [v // 2 for v in lo], [v // 2 for v in hi]
                cells = np.zeros([chi[d] - clo[d] + 1 for d in range(dim)], dtype=bool)
                for (plo, phi) in prev:
                    sl = tuple(slice(max(clo[d], plo[d]) - clo[d], min(chi[d], phi[d]) - clo[d] + 1) for d in range(dim))
                    if all(s.start < s.stop for s in sl):
                        cells[sl] = True
                assert cells.all()
        prev = b


def _dense(L, comp):
    (lo, hi) = L.domain
    g = np.full(tuple(hi[d] - lo[d] + 1 for d in range(len(lo))), np.nan)
    for b, (blo, bhi) in enumerate(L.boxes):
        g[tuple(slice(blo[d] - lo[d], bhi[d] - lo[d] + 1) for d in range(len(lo)))] = L.data[b][comp]
    return g


@needs_ref
def test_read_amr_checkpoint_fine_averages_equal_coarse():
    a = C.read_amr_hierarchy(SELECT)
    assert a.space_dim == 2 and len(a.levels) == 3
    assert a.names == ["density", "x-momentum", "y-momentum", "energy-density"]
    assert [L.ref_ratio for L in a.levels] == [4, 4, 2]
    assert [len(L.boxes) for L in a.levels] == [1, 4, 16]
    for l in range(2):
        r = a.levels[l].ref_ratio
        assert a.levels[l + 1].dx == a.levels[l].dx / r
        for comp in range(4):
            c, f = _dense(a.levels[l], comp), _dense(a.levels[l + 1], comp)
            nx, ny = f.shape
            avg = f.reshape(nx // r, r, ny // r, r).mean(axis=(1, 3))
            m = ~np.isnan(avg)
            assert m.sum() > 100
            rel = np.abs(avg[m] - c[m]) / np.abs(c[m])
            assert rel.max() < 1e-14, (l, comp, rel.max())
    for L in a.levels:  # conserved density and energy stay positive
        assert min(d[0].min() for d in L.data) > 0 and min(d[3].min() for d in L.data) > 0


def _tree(g):
    """(attribute name -> dtype kind/size/fields, dataset name -> (dtype, shape), groups recursively)"""
    out = {}
    for k, v in g.attrs.items():
        dt = np.asarray(v).dtype if not isinstance(v, str) else "str"
        out["@" + k] = str(dt.names if getattr(dt, "names", None) else dt)
    for k, m in g.members.items():
        if isinstance(m, C.H5Group):
            out[k] = _tree(m)
        else:
            out[k] = (str(m.dtype.names or m.dtype), m.shape)
    return out


@needs_ref
def test_rewrite_checkpoint_same_tree_and_data(tmp_path):
    a = C.read_amr_hierarchy(SELECT)
    p = str(tmp_path / "rewritten.2d.hdf5")
    C.write_amr_hierarchy(p, a)
    b = C.read_amr_hierarchy(p)
    assert b.names == a.names and b.space_dim == 2
    for La, Lb in zip(a.levels, b.levels):
        assert La.boxes == Lb.boxes and La.domain == Lb.domain and La.ref_ratio == Lb.ref_ratio
        assert (La.dx, La.dt, La.time) == (Lb.dx, Lb.dt, Lb.time)
        for x, y in zip(La.data, Lb.data):
            np.testing.assert_array_equal(x, y)
    ta, tb = _tree(C.read_file(SELECT)), _tree(C.read_file(p))
    # the root attributes this file's writer added (time, iteration, max_level) are not
    # WriteAMRHierarchyHDF5's; everything else matches, names and types
    for k in ("@time", "@iteration", "@max_level"):
        ta.pop(k)
    tb.pop("@filetype")
    for lev in ("level_0", "level_1", "level_2"):
        # this 2001 file predates the per-level time attribute and the Processors
        # dataset that the current writeLevel / write(BoxLayout) add (CH_HDF5.H:1863,
        # CH_HDF5.cpp:190); the rest is the same
        assert set(tb[lev]) - set(ta[lev]) == {"@time", "Processors"}
        assert set(ta[lev]) <= set(tb[lev])
        for k in ta[lev]:
            if k == "data_attributes":
                assert ta[lev][k] == tb[lev][k]
            elif k.startswith("@"):
                assert ta[lev][k] == tb[lev][k], (lev, k)
            else:
                assert ta[lev][k][1] == tb[lev][k][1], (lev, k)
    assert ta.keys() == tb.keys()


@pytest.mark.parametrize("ghost", [0, 1])
def test_write_read_round_trip_3d(tmp_path, ghost):
    h = H.amr_hierarchy(32, fracs=(0.25, 0.06), max_box=16, block=4)
    rng = np.random.default_rng(3)
    f1, f2 = rng.standard_normal(h.total_cells()), rng.standard_normal(h.total_cells())
    amr = C.file_from_hierarchy(h, {"phi": f1, "rhs": f2}, dt=0.5, time=1.25)
    if ghost:
        for L in amr.levels:
            L.ghost = (1, 1, 1)
            L.data = [np.pad(d, ((0, 0), (1, 1), (1, 1), (1, 1)), constant_values=7.0) for d in L.data]
    p = str(tmp_path / "h.3d.hdf5")
    C.write_amr_hierarchy(p, amr)
    back = C.read_amr_hierarchy(p)
    assert back.names == ["phi", "rhs"] and back.space_dim == 3
    h2 = C.hierarchy_from_file(back)
    assert h2.levels == h.levels and h2.domain == h.domain and h2.dx0 == h.dx0
    np.testing.assert_array_equal(C.field_from_file(back, 0), f1)
    np.testing.assert_array_equal(C.field_from_file(back, 1), f2)
    assert [L.dt for L in back.levels] == [0.5, 0.25, 0.125]
    assert all(L.time == 1.25 for L in back.levels)
    root = C.read_file(p)
    assert root.attrs["filetype"] == "VanillaAMRFileType"
    assert root.members["Chombo_global"].attrs["SpaceDim"] == 3
    da = root.members["level_1"].members["data_attributes"].attrs
    assert da["objectType"] == "FArrayBox" and da["comps"] == 2
    assert int(da["outputGhost"]["intvecti"]) == ghost


def test_reader_refuses_other_files(tmp_path):
    p = tmp_path / "x.h5"
    p.write_bytes(b"not hdf5" * 20)
    with pytest.raises(ValueError):
        C.read_file(str(p))
