"""Synthetic wing meshes in CPanel's .tri convention.

A .tri file (read by geometry::readTri, CPanel/geometry.cpp:155-239) holds
node coordinates, 1-based triangle connectivity and one surface ID per
triangle; IDs above 1000 are wake surfaces.  Meshes built here have one body
ID (1) and one wake ID (1001), outward normals, a sharp closed trailing edge
shared by the upper and lower panels, flat tip caps of npt-2 triangles, and
a flat prescribed wake whose first node row coincides with the trailing edge
(welded by the reader, geometry.cpp:631-667).
"""
import numpy as np


def naca4(code, n_per_surface, closed_te=True):
    """Closed airfoil loop (x, z) for chord 1: trailing edge, lower surface to
    the leading edge, upper surface back (2*n_per_surface distinct points),
    cosine spacing.  Standard NACA 4-digit thickness/camber formulas."""
    m = int(code[0]) / 100.0
    p = int(code[1]) / 10.0
    t = int(code[2:]) / 100.0
    beta = np.linspace(0.0, np.pi, n_per_surface + 1)
    x = 0.5 * (1.0 - np.cos(beta))
    a4 = -0.1036 if closed_te else -0.1015
    yt = 5 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x**2 + 0.2843 * x**3 + a4 * x**4)
    yt[-1] = 0.0
    yc = np.zeros_like(x)
    dyc = np.zeros_like(x)
    if m > 0:
        f = x < p
        yc[f] = m / p**2 * (2 * p * x[f] - x[f] ** 2)
        dyc[f] = 2 * m / p**2 * (p - x[f])
        g = ~f
        yc[g] = m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * x[g] - x[g] ** 2)
        dyc[g] = 2 * m / (1 - p) ** 2 * (p - x[g])
    th = np.arctan(dyc)
    xu, zu = x - yt * np.sin(th), yc + yt * np.cos(th)
    xl, zl = x + yt * np.sin(th), yc - yt * np.cos(th)
    # loop: TE (x=1), lower TE->LE, LE, upper LE->TE (TE not repeated)
    lx = np.concatenate([[1.0], xl[::-1][1:-1], [0.0], xu[1:-1]])
    lz = np.concatenate([[yc[-1] if m > 0 else 0.0], zl[::-1][1:-1], [0.0], zu[1:-1]])
    lz[0] = 0.0
    return lx, lz


def span_stations(span, n_strips, spacing="uniform"):
    if spacing == "uniform":
        return np.linspace(-span / 2, span / 2, n_strips + 1)
    if spacing == "cosine":  # dense toward both tips
        th = np.linspace(-np.pi / 2, np.pi / 2, n_strips + 1)
        y = span / 2 * np.sin(th)
        y[0], y[-1] = -span / 2, span / 2
        return y
    raise ValueError(spacing)


def wing(code="4412", chord=1.0, span=6.0, n_chord=80, n_span=30, spacing="uniform", wake_len=20.0,
         taper=1.0, sweep_deg=0.0, with_wake=True, wake_rows=1, noise=1e-8, seed=1234):
    """Returns dict(xyz (nn,3), tri (nt,3) 0-based, ids (nt,)).
    n_chord = panels around the section (n_chord/2 per surface).

    noise: relative amplitude (times the chord) of a seeded perturbation of
    the x and z coordinates of interior-station nodes, the level of the
    digits CAD exporters write.  Without it every panel of an untwisted wing
    has exactly coplanar neighbours and bodyPanel::setTipFlag
    (bodyPanel.cpp:58-95, angle < 1e-8) marks the whole wing as tip patches.
    The trailing-edge and leading-edge nodes, and the end stations (flat tip
    caps), are left exact."""
    assert n_chord % 2 == 0
    lx, lz = naca4(code, n_chord // 2)
    npt = len(lx)
    ys = span_stations(span, n_span, spacing)
    nst = len(ys)
    xyz = np.zeros((nst * npt, 3))
    for j, y in enumerate(ys):
        eta = abs(y) / (span / 2)
        c = chord * (1 - (1 - taper) * eta)
        xqc = abs(y) * np.tan(np.radians(sweep_deg))
        xle = xqc + 0.25 * chord - 0.25 * c
        xyz[j * npt:(j + 1) * npt, 0] = xle + c * lx
        xyz[j * npt:(j + 1) * npt, 1] = y
        xyz[j * npt:(j + 1) * npt, 2] = c * lz
    if noise > 0:
        rng = np.random.default_rng(seed)
        for j in range(1, nst - 1):
            sl = slice(j * npt, (j + 1) * npt)
            d = rng.uniform(-1.0, 1.0, size=(npt, 2)) * noise * chord
            d[0] = 0.0            # trailing edge
            d[npt // 2] = 0.0     # leading edge
            xyz[sl, 0] += d[:, 0]
            xyz[sl, 2] += d[:, 1]
    tris = []
    for j in range(nst - 1):
        for k in range(npt):
            a = j * npt + k
            b = j * npt + (k + 1) % npt
            c_ = (j + 1) * npt + (k + 1) % npt
            d = (j + 1) * npt + k
            tris.append((a, b, c_))
            tris.append((a, c_, d))
    # tips: the section polygon (npt points) closed by npt-2 triangles laid
    # strip-wise between the lower and upper surface points (a fan from one
    # vertex would make the tip panels' centroids nearly collinear, which
    # leaves the 2-D CHTLS fits of tip panels singular); outward normals
    for j, sgn in ((0, -1.0), (nst - 1, 1.0)):
        base = j * npt
        h = npt // 2
        poly_tris = [(0, 1, npt - 1)]
        for q in range(1, h - 1):
            lo0, lo1 = q, q + 1
            up0, up1 = npt - q, npt - q - 1
            poly_tris.append((lo0, lo1, up0))
            poly_tris.append((lo1, up1, up0))
        poly_tris.append((h - 1, h, h + 1))
        for a, b, c_ in poly_tris:
            t = (base + a, base + b, base + c_)
            p = xyz[list(t)]
            nrm = np.cross(p[1] - p[0], p[2] - p[0])
            if nrm[1] * sgn < 0:
                t = (t[0], t[2], t[1])
            tris.append(t)
    tri = np.array(tris, dtype=np.int64)
    ids = np.ones(len(tri), dtype=np.int64)
    if with_wake:
        nb_nodes = len(xyz)
        te = xyz[[j * npt for j in range(nst)]].copy()
        rows = [te]
        for r in range(1, wake_rows + 1):
            q = te.copy()
            q[:, 0] += wake_len * r / wake_rows
            rows.append(q)
        wn = np.concatenate(rows)
        wtri = []
        for r in range(wake_rows):
            o0 = nb_nodes + r * nst
            o1 = nb_nodes + (r + 1) * nst
            for j in range(nst - 1):
                wtri.append((o0 + j, o1 + j, o1 + j + 1))
                wtri.append((o0 + j, o1 + j + 1, o0 + j + 1))
        xyz = np.concatenate([xyz, wn])
        tri = np.concatenate([tri, np.array(wtri, dtype=np.int64)])
        ids = np.concatenate([ids, np.full(len(wtri), 1001, dtype=np.int64)])
    return dict(xyz=xyz, tri=tri, ids=ids)


def write_tri(path, mesh):
    xyz, tri, ids = mesh["xyz"], mesh["tri"], mesh["ids"]
    with open(path, "w") as f:
        f.write("%d %d\n" % (len(xyz), len(tri)))
        for p in xyz:
            f.write("%.17g %.17g %.17g\n" % tuple(p))
        for t in tri:
            f.write("%d %d %d\n" % (t[0] + 1, t[1] + 1, t[2] + 1))
        for i in ids:
            f.write("%d\n" % i)


def read_tri(path):
    v = open(path).read().split()
    nn, nt = int(v[0]), int(v[1])
    xyz = np.array(v[2:2 + 3 * nn], float).reshape(-1, 3)
    o = 2 + 3 * nn
    tri = np.array(v[o:o + 3 * nt], np.int64).reshape(-1, 3) - 1
    ids = np.array(v[o + 3 * nt:o + 4 * nt], np.int64)
    return dict(xyz=xyz, tri=tri, ids=ids)


# ---- the workloads --------------------------------------------------------
def config_c1():
    """NACA 4412, chord 1, AR 6, 80 x 30, flat wake 20 c (about 4,956 body tris)."""
    return wing("4412", 1.0, 6.0, 80, 30, "uniform", 20.0)


def config_c2():
    """NACA 0012, AR 10, 160 cosine x 80 cosine-to-tips, wake 30 c (25,916 body tris)."""
    return wing("0012", 1.0, 10.0, 160, 80, "cosine", 30.0)


def config_c4(with_wake=True):
    """Swept (20 deg at the quarter chord), tapered (0.5) NACA 4412, span 8,
    60 panels around x 50 strips (6,116 body tris, 50 trailing-edge edges);
    one buffer wake row of one chord behind the trailing edge, or none
    (with_wake=False: the vortex-particle geometry, whose buffer rows follow
    from the time step)."""
    return wing("4412", 1.0, 8.0, 60, 50, "uniform", 1.0, taper=0.5, sweep_deg=20.0, with_wake=with_wake)


def c4_particles(mesh, steps=400, per_edge=4, dt=0.02, seed=0):
    """Synthetic vortex-particle wake of the C4 case: per_edge particles per
    trailing-edge edge per step, convected downstream at |V| = 1 for `steps`
    steps (dt chord/|V|), with seeded jitter; strengths of a trailing vortex
    sheet; radius 0.1 x mean panel size.  Returns xyz, alpha, radius."""
    rng = np.random.default_rng(seed)
    xyz = np.asarray(mesh["xyz"])
    tri = np.asarray(mesh["tri"])
    ids = np.asarray(mesh["ids"])
    w = tri[ids > 1000]
    # wake panels (TE_j, far_j, far_j+1) and (TE_j, far_j+1, TE_j+1): TE edge = (a, c) of the second kind
    te = [(t[0], t[2]) for t in w[1::2]]
    p0 = np.array([xyz[a] for a, b in te])
    p1 = np.array([xyz[b] for a, b in te])
    body = tri[ids < 1000]
    e = xyz[body[:, 1]] - xyz[body[:, 0]]
    f = xyz[body[:, 2]] - xyz[body[:, 0]]
    mean_size = float(np.mean(np.sqrt(0.5 * np.linalg.norm(np.cross(e, f), axis=1))))
    n_edges = len(te)
    out_x, out_a = [], []
    for s in range(steps):
        for q in range(per_edge):
            u = (q + 0.5) / per_edge
            x = p0 + u * (p1 - p0)
            x = x + np.array([1.0, 0.0, 0.0]) * (dt * (steps - s - 0.5))
            x = x + rng.normal(scale=0.01, size=x.shape)
            seg = p1 - p0
            g = np.sin(np.linspace(0, np.pi, n_edges))[:, None]
            a = seg * g * 0.05 + rng.normal(scale=1e-3, size=x.shape)
            out_x.append(x)
            out_a.append(a)
    X = np.concatenate(out_x)
    A = np.concatenate(out_a)
    return X, A, np.full(len(X), 0.1 * mean_size)


def small_wing(code="4412", n_chord=20, n_span=6, span=4.0, spacing="uniform", taper=1.0, sweep_deg=0.0):
    return wing(code, 1.0, span, n_chord, n_span, spacing, 10.0, taper, sweep_deg)
