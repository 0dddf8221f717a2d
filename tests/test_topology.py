"""Product topology (host C++ in libcpgpu.so, host-only context) against the
oracle restatement: panel order, Kutta folding table, clusters, tip flags.
Integer outputs must be identical; coefficients bit-identical.  CPU only."""
import os

import numpy as np
import pytest

from cpanel_amd import meshgen
from cpanel_amd.case import Context
from oracle import oracle as O


@pytest.fixture(scope="module")
def host_ctx():
    return Context(device=-1)


def meshes(golden):
    yield "4412", meshgen.read_tri(os.path.join(golden, "4412.tri"))
    yield "4412AR6", meshgen.read_tri(os.path.join(golden, "4412AR6.tri"))
    yield "c1-small", meshgen.wing("4412", 1.0, 6.0, 40, 10, "uniform", 20.0)
    yield "c2-small", meshgen.wing("0012", 1.0, 10.0, 40, 16, "cosine", 30.0)
    yield "tapered-swept", meshgen.wing("2412", 1.0, 8.0, 30, 12, "cosine", 15.0, taper=0.5, sweep_deg=20.0)


def test_topology_matches_oracle(host_ctx, golden):
    for name, m in meshes(golden):
        og = O.Geom.from_arrays(m["xyz"], m["tri"], m["ids"])
        pg = host_ctx.geometry(m)
        assert (pg.nb, pg.nw) == (og.nb, og.nw), name
        ob, pb = og.body_info(), pg.body()
        assert np.array_equal(ob["tri_index"], pb["tri_index"]), name
        assert np.array_equal(ob["center"], pb["center"]), name
        assert np.array_equal(ob["normal"], pb["normal"]), name
        assert np.array_equal(ob["area"], pb["area"]), name
        ow, pw = og.wake_info(), pg.wake()
        assert np.array_equal(ow["tri_index"], pw["tri_index"]), name
        assert np.array_equal(ow["idx4"], pw["idx4"]), name
        assert np.array_equal(ow["coef"], pw["coef"]), name
        ocl, ots = og.clusters()
        pcl, pts, ptip = pg.clusters()
        assert np.array_equal(ocl, pcl), name
        assert np.array_equal(ots, pts), name
        assert np.array_equal(ob["tip"], ptip), name


def test_generated_meshes_are_closed_and_outward():
    for m in (meshgen.config_c1(), meshgen.small_wing()):
        body = m["ids"] <= 1000
        tri = m["tri"][body]
        x = m["xyz"]
        vol = np.einsum("ij,ij->i", x[tri[:, 0]], np.cross(x[tri[:, 1]], x[tri[:, 2]])).sum() / 6
        assert vol > 0
        # watertight: every body edge shared by exactly two body triangles
        e = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        assert np.all(cnt == 2)


def test_config_sizes():
    m1 = meshgen.config_c1()
    assert int((m1["ids"] <= 1000).sum()) == 4956
    m2 = meshgen.config_c2()
    assert int((m2["ids"] <= 1000).sum()) == 25916


def test_bad_mesh_reports_error(host_ctx):
    from cpanel_amd.case import CPanelError
    m = meshgen.small_wing()
    tri = m["tri"].copy()
    tri[0, 0] = len(m["xyz"]) + 5  # missing node
    with pytest.raises(CPanelError):
        host_ctx.geometry(dict(xyz=m["xyz"], tri=tri, ids=m["ids"]))
