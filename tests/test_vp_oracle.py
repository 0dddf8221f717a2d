"""The vortex-particle part of the FP64 oracle, pinned without the reference.

* octree: the reference's own known-answer test (a 3x3x3 grid gives a root
  centred at (1,1,1), octree_test.cpp:316-356) and the tree that the split rule
  (> 10 members) forces on it;
* pair kernels: closed-form limits (Gaussian and algebraic kernels tend to the
  singular Biot-Savart kernel far away, the Gaussian to a finite value at the
  core), the stretching kernel against finite differences of the Gaussian
  velocity field, and the conservation/antisymmetry of the PSE diffusion;
* Barnes-Hut: with opening angle 0 no cell is far and the traversal must
  equal the direct sum with the same leaf kernels; with theta > 0 the error
  against the direct sum shrinks as theta does;
* panel -> point velocity: a closed surface of constant doublet strength
  induces no velocity (its vortex-ring edges cancel pairwise).
"""
import numpy as np
import pytest

from oracle import vp_oracle as V

PI = np.pi


def cloud(n, seed=3, spread=1.0):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3)) * spread
    a = rng.normal(size=(n, 3)) * 0.1
    r = 0.05 + 0.05 * rng.random(n)
    return x, a, r


def test_octree_reference_kat():
    g = np.array([[i, j, k] for i in range(3) for j in range(3) for k in range(3)], dtype=float)
    t = V.BHTree(g, np.ones_like(g), np.full(27, 0.1))
    info = t.info()
    assert np.all(info["origin"] == 1.0)
    assert np.allclose(info["half"], 0.51 * 2.0)
    # 27 > 10 members: the root splits once; octant counts 8,4,4,4,2,2,2,1
    # all <= 10, so the tree is the root plus 8 leaves on level 2
    assert info["n_nodes"] == 9 and info["levels"] == 2


def test_root_leaf_gives_zero():
    """With <= 10 particles the root is a leaf and has no children; the
    reference's traversal starts at the root's children, so every sum is 0."""
    x, a, r = cloud(10)
    t = V.BHTree(x, a, r)
    assert t.info()["n_nodes"] == 1
    for mode in (1, 2, 3):
        assert np.all(t.eval(mode) == 0)
    assert np.all(t.eval(0, x + 0.3) == 0)


def test_kernel_far_limits():
    xs, a = np.zeros(3), np.array([0.3, -0.2, 0.5])
    for d in (np.array([40.0, 3.0, -7.0]), np.array([-5.0, 60.0, 2.0])):
        bs = -np.cross(d, a) / (4 * PI * np.linalg.norm(d) ** 3)   # singular Biot-Savart
        g = V.pair(1, xs, a, 0.1, d, np.zeros(3), 0.1)
        assert np.allclose(g, bs, rtol=1e-13, atol=0)
        alg = V.pair(0, xs, a, 0.1, d)
        sig = 1.5 * 0.1
        r2 = d @ d
        # (r^2 + 2.5 s^2) / (r^2 + s^2)^2.5 -> 1/r^3 (1 + O(s^4/r^4))
        assert np.allclose(alg, bs, rtol=2 * (sig * sig / r2) ** 2 * 10, atol=0)


def test_gaussian_core_limit():
    """K(rho) -> (2 pi)^-1.5 / 3 as rho -> 0: u ~ -K/s^3 d x alpha."""
    a = np.array([0.0, 0.0, 1.0])
    r = 0.1
    sig = np.sqrt(2 * (1.5 * r) ** 2) / np.sqrt(2)
    d = np.array([1e-2, 0.0, 0.0]) * sig
    u = V.pair(1, np.zeros(3), a, r, d, np.zeros(3), r)
    rho = 1e-2
    K = (2 * PI) ** -1.5 * (1.0 / 3 - rho * rho / 10)
    ref = -K / sig ** 3 * np.cross(d, a)
    assert np.allclose(u, ref, rtol=1e-6)


def test_stretching_is_gradient_of_gaussian_velocity():
    """source->vortexStretching(target) (the Barnes-Hut orientation) equals
    (alpha_s . grad) u_t(x_s): the stretching of the source by the Gaussian
    field of a particle (x_t, alpha_t), by central differences."""
    rng = np.random.default_rng(11)
    for _ in range(5):
        xs, xt = rng.normal(size=3) * 0.1, rng.normal(size=3) * 0.1
        as_, at = rng.normal(size=3), rng.normal(size=3)
        rs, rt = 0.08, 0.11
        st = V.pair(2, xs, as_, rs, xt, at, rt)
        h = 1e-6
        grad = np.zeros((3, 3))
        for l in range(3):
            e = np.zeros(3)
            e[l] = h
            up = V.pair(1, xt, at, rt, xs + e, np.zeros(3), rs)
            um = V.pair(1, xt, at, rt, xs - e, np.zeros(3), rs)
            grad[:, l] = (up - um) / (2 * h)
        fd = grad @ as_
        assert np.allclose(st, fd, rtol=1e-6, atol=1e-8 * np.abs(fd).max())


def test_stretching_cutoff():
    xs, xt = np.zeros(3), np.array([0.51, 0.0, 0.0])
    assert np.all(V.pair(2, xs, np.ones(3), 0.1, xt, np.ones(3), 0.1) == 0)   # beyond 5 r_source
    assert np.any(V.pair(2, xs, np.ones(3), 0.11, xt, np.ones(3), 0.1) != 0)


def test_diffusion_antisymmetric_and_conservative():
    x, a, r = cloud(60, seed=5, spread=0.2)
    d01 = V.pair(3, x[0], a[0], r[0], x[1], a[1], r[1])
    d10 = V.pair(3, x[1], a[1], r[1], x[0], a[0], r[0])
    assert np.allclose(d01, -d10, rtol=1e-14, atol=0)
    tot = V.direct(3, x, a, r).sum(axis=0)
    scale = np.abs(V.direct(3, x, a, r)).sum()
    assert np.all(np.abs(tot) <= 1e-13 * scale)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_bh_theta0_equals_direct(mode):
    x, a, r = cloud(700, seed=7, spread=0.5)
    pts = np.random.default_rng(8).normal(size=(50, 3)) * 0.5
    t = V.BHTree(x, a, r, theta=0.0)
    assert t.info()["levels"] >= 3
    bh = t.eval(mode, pts)
    dr = V.direct(mode, x, a, r, pts)
    assert np.allclose(bh, dr, rtol=1e-11, atol=1e-13 * np.abs(dr).max())


@pytest.mark.parametrize("mode", [0, 1])
def test_bh_error_shrinks_with_theta(mode):
    x, a, r = cloud(3000, seed=9, spread=1.0)
    pts = np.random.default_rng(10).normal(size=(200, 3))
    dr = V.direct(mode, x, a, r, pts)
    errs = []
    for th in (0.8, 0.5, 0.25):
        bh = V.BHTree(x, a, r, theta=th).eval(mode, pts)
        errs.append(np.linalg.norm(bh - dr) / np.linalg.norm(dr))
    # monopole pseudo-particles of random strengths are crude; only the trend
    # is a property of the scheme
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 0.5 * errs[0]


def icosphere():
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=float)
    v /= np.linalg.norm(v, axis=1)[:, None]
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    return v, f


def test_closed_doublet_surface_induces_no_velocity():
    v, f = icosphere()
    verts = v[f]
    n = len(f)
    pts = np.array([[0.3, 0.2, -0.1], [1.5, 0.4, 0.2], [-0.7, 1.9, 0.5]])
    u = V.panels_to_points(verts, np.zeros(n), np.ones(n), None, None, pts)
    # vortex-ring edges cancel pairwise; every point here is in the near field
    assert np.abs(u).max() < 1e-14
    # and a single panel does induce a velocity
    u1 = V.panels_to_points(verts[:1], np.zeros(1), np.ones(1), None, None, pts)
    assert np.abs(u1).max() > 1e-3


# ---------------------------------------------------------------- stretching
def _fd(f, p, h=1e-6):
    J = np.zeros((3, 3))
    for i in range(3):
        e = np.zeros(3)
        e[i] = h
        J[i] = (f(p + e) - f(p - e)) / (2 * h)
    return J  # J[i, j] = d f_j / d x_i, the layout of the gradient matrices


def _generic_triangle(seed):
    rng = np.random.default_rng(seed)
    tri = rng.normal(size=(3, 3))
    c = tri.mean(0)
    n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    n /= np.linalg.norm(n)
    l = (tri[0] - c) / np.linalg.norm(tri[0] - c)
    R = np.array([l, np.cross(n, l), n])
    L = max(np.linalg.norm(tri[(k + 1) % 3] - tri[k]) for k in range(3))
    return tri, c, n, R, L


def test_panel_doublet_gradient_is_derivative_of_doublet_velocity():
    """panel::velocityGradientDoublet (sum of gradDoub over the edges) is the
    exact derivative of panel::dubVInf's near-field velocity."""
    from oracle import oracle as O
    for seed in (1, 2, 3):
        tri, c, n, R, L = _generic_triangle(seed)
        p = c + 0.8 * L * n + 0.4 * L * R[0]
        M = V.velgrad(tri, p, V.VG_PANEL_DOUBLET, mu=1.0)
        J = _fd(lambda q: O.dub_v_inf(tri, q), p)
        assert np.abs(M - J).max() <= 1e-7 * np.abs(J).max()


def test_panel_source_gradient_is_local_frame_derivative():
    """bodyPanel::velocityGradientTriSource is minus the derivative of the
    source velocity (bodyPanel::panelVInf) taken in the panel's local frame
    (local coordinates in, local velocity components out)."""
    from oracle import oracle as O
    for seed in (3, 4, 5):
        tri, c, n, R, L = _generic_triangle(seed)
        for p in (c + 0.7 * L * n + 0.3 * L * R[0], c + np.array([1.0, 2.0, 0.5]) * L):
            M = V.velgrad(tri, p, V.VG_PANEL_SOURCE, sigma=1.0)
            J = R @ _fd(lambda q: O.panel_v_inf(tri, q)[0], p) @ R.T
            assert np.abs(M + J).max() <= 1e-6 * np.abs(J).max()


def test_point_gradients_far_field():
    """Point doublet and point source gradients (panel.cpp:813-840,
    bodyPanel.cpp:733-763) are minus the derivatives of the far-field
    velocities for a panel normal to z and a point at z = 0 (where the
    formulas' use of the centre's z is exact); the source's d v / d y entry
    is the reference's ((x-x0)^2 - (y-y0)^2 + z^2) / r^5, which differs from
    the derivative by (y-y0)^2 / r^5 and is kept as written."""
    from oracle import oracle as O
    tri = np.array([[1.0, 0, 0.3], [-1, 1, 0.3], [-1, -1, 0.3]])
    c = tri.mean(0)
    for p in (np.array([9.0, -7.0, 0.0]), np.array([-6.0, 10.0, 0.0])):
        Md = V.velgrad(tri, p, V.VG_POINT_DOUBLET, mu=1.0)
        Jd = _fd(lambda q: O.panel_v_inf(tri, q)[1], p)
        assert np.abs(Md + Jd).max() <= 1e-7 * np.abs(Jd).max()
        Ms = V.velgrad(tri, p, V.VG_POINT_SOURCE, sigma=1.0)
        Js = _fd(lambda q: O.panel_v_inf(tri, q)[0], p)
        mask = np.ones((3, 3), bool)
        mask[1, 1] = False
        assert np.abs((Ms + Js)[mask]).max() <= 1e-7 * np.abs(Js).max()
        d = c - p
        r5 = (d[0] ** 2 + d[1] ** 2 + c[2] ** 2) ** 2.5
        area = 2.0
        assert abs(Ms[1, 1] - (-Js[1, 1] + d[1] ** 2 / r5 * area / (4 * np.pi))) <= 1e-7 * np.abs(Js).max()


def test_part_stretching_dispatch():
    """partStretching: body panels add source + doublet gradients (point
    forms beyond 5 longest sides), wake panels the doublet only; the
    stretching is M alpha."""
    tri, c, n, R, L = _generic_triangle(7)
    al = np.array([0.3, -0.2, 0.5])
    for p in (c + 0.6 * L * n, c + 7 * L * n):
        far = np.linalg.norm(p - c) / L > 5
        Mb = V.velgrad(tri, p, V.VG_PART, body=True, sigma=0.7, mu=-0.4)
        ks, kd = (V.VG_POINT_SOURCE, V.VG_POINT_DOUBLET) if far else (V.VG_PANEL_SOURCE, V.VG_PANEL_DOUBLET)
        assert np.allclose(Mb, V.velgrad(tri, p, ks, sigma=0.7) + V.velgrad(tri, p, kd, mu=-0.4), rtol=1e-14,
                           atol=0)
        Mw = V.velgrad(tri, p, V.VG_PART, body=False, mu=-0.4)
        assert np.array_equal(Mw, V.velgrad(tri, p, kd, mu=-0.4))
        s = V.panels_stretching(tri[None], [0.7], [-0.4], tri[None], [-0.4], p[None], al[None])[0]
        assert np.allclose(s, Mb @ al + Mw @ al, rtol=1e-13, atol=0)


def test_doublet_gradient_exclusions():
    """Edges whose filament core (5 % of the edge) holds the point add
    nothing; a point over the panel within 5 % of the longest side keeps
    only the d w / d x_i column (panel.cpp:842-874)."""
    tri, c, n, R, L = _generic_triangle(8)
    on = c + 0.01 * L * n
    M = V.velgrad(tri, on, V.VG_PANEL_DOUBLET, mu=1.0)
    assert np.all(M[:, :2] == 0) and np.any(M[:, 2] != 0)
    # a point just above vertex 0: edges 0 -> 1 and 2 -> 0 are skipped, only
    # the edge 1 -> 2 term remains (gradDoub of that edge)
    out = (tri[0] - c) / np.linalg.norm(tri[0] - c)  # just outside the panel's projection
    near_vertex = tri[0] + 0.002 * L * out + 0.001 * L * n
    Mv = V.velgrad(tri, near_vertex, V.VG_PANEL_DOUBLET, mu=4 * np.pi)
    e12 = V.velgrad(np.roll(tri, -1, axis=0), near_vertex, V.VG_EDGE)
    assert np.allclose(Mv, e12, rtol=1e-13, atol=0)


def test_core_kernel_velocity_is_alg_smooth_with_one_core():
    """The configured case's velocity kernel is the reference's
    velInflAlgSmooth (particle.cpp:31-38) with sigma = core for every
    particle: equal to the oracle's mode-0 kernel of a particle whose radius
    is core / 1.5 (coreOverlap)."""
    rng = np.random.default_rng(3)
    core = 0.07
    for _ in range(20):
        xs, a, xt = rng.normal(size=3), rng.normal(size=3), rng.normal(size=3) * 0.2
        v1 = V.pair_core(0, core, xs, a, xt)
        v0 = V.pair(0, xs, a, core / 1.5, xt)
        assert np.allclose(v1, v0, rtol=1e-13, atol=1e-16)


def test_core_kernel_stretching_is_directional_derivative():
    """(alpha_t . grad) u of the core kernel: central differences of the
    velocity along alpha_t (h = 1e-6, error O(h^2))."""
    rng = np.random.default_rng(4)
    core = 0.05
    for _ in range(30):
        xs, a = rng.normal(size=3), rng.normal(size=3)
        xt = xs + rng.normal(size=3) * rng.choice([0.02, 0.1, 1.0])
        at = rng.normal(size=3)
        h = 1e-6
        fd = (V.pair_core(0, core, xs, a, xt + h * at) - V.pair_core(0, core, xs, a, xt - h * at)) / (2 * h)
        st = V.pair_core(2, core, xs, a, xt, at)
        assert np.allclose(st, fd, rtol=2e-6, atol=1e-9 * np.abs(fd).max() + 1e-12)


def test_core_direct_sums_excludes_self_and_matches_pairs():
    rng = np.random.default_rng(5)
    X, A = rng.normal(size=(40, 3)), rng.normal(size=(40, 3))
    core = 0.1
    vel = V.direct_core(V.CORE_VEL, core, X, A)
    st = V.direct_core(V.CORE_STRETCH, core, X, A)
    for t in (0, 17, 39):
        ev = sum(V.pair_core(0, core, X[j], A[j], X[t]) for j in range(40) if j != t)
        es = sum(V.pair_core(2, core, X[j], A[j], X[t], A[t]) for j in range(40) if j != t)
        assert np.allclose(vel[t], ev, rtol=1e-12) and np.allclose(st[t], es, rtol=1e-12)
    pts = rng.normal(size=(5, 3))
    vp = V.direct_core(V.CORE_VEL_POINTS, core, X, A, pts)
    assert np.allclose(vp[2], sum(V.pair_core(0, core, X[j], A[j], pts[2]) for j in range(40)), rtol=1e-12)
