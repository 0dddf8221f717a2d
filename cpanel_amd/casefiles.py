"""Per-case output files in the reference's formats.

cpCase::writeFiles (cpCase.cpp:667-709) writes, for every case, a directory
named V<V>_Mach<M>_alpha<a>_beta<b> holding
  * surfaceData.vtu   body cell data (cpCase::writeBodyData, :711-758);
  * wakeData.vtu      wake doublet strengths and potentials (writeWakeData,
                      :883-908);
  * spanwiseData_Wake<i>.csv  2y/b, Cl, Cdi per Trefftz station
                      (writeSpanwiseData, :910-935).
The VTU layout follows VTUfile::write (VTUfile.cpp:20-131): ASCII
UnstructuredGrid, one Piece, Float64 cell arrays with the reference's names
and order (including its "zPostition" spelling), values in default stream
formatting (6 significant digits) each followed by a tab, one row per line.
Cell order is the body-panel order of the solver (the reference's panel
order); the points are the input mesh nodes and the connectivity is each
panel's node indices in that list (0-based).

The arrays come from Geometry.case_fields, which evaluates them on the
device from the solved doublet strengths.
"""
import os

import numpy as np


def _g(v):
    # std::ostream default formatting of a double
    return "%g" % v


def _label(V, mach, alpha, beta):
    return "V%s_Mach%s_alpha%s_beta%s" % (_g(V), _g(mach), _g(alpha), _g(beta))


def _double_array(name, a):
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    out = ['\t\t\t\t<DataArray type="Float64" Name="%s" NumberOfComponents="%d" Format="ascii">\n' % (name, a.shape[1])]
    out.extend("".join(_g(v) + "\t" for v in row) + "\n" for row in a)
    out.append("\t\t\t\t</DataArray>\n")
    return "".join(out)


def _int_array(name, a):
    a = np.asarray(a)
    if a.ndim == 1:
        a = a[:, None]
    out = ['\t\t\t\t<DataArray type="Int32" Name="%s" Format="ascii">\n' % name]
    out.extend("".join("%d\t" % v for v in row) + "\n" for row in a)
    out.append("\t\t\t\t</DataArray>\n")
    return "".join(out)


def vtu_text(points, conn, cell_data):
    """One-piece VTU text; cell_data: list of (name, array)."""
    conn = np.asarray(conn)
    nv = conn.shape[1]
    ctype = {1: 1, 2: 3, 3: 5, 4: 9, 8: 11}[nv]
    s = ['<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n', "\t<UnstructuredGrid>\n",
         '\t\t<Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (len(points), len(conn))]
    if cell_data:
        s.append('\t\t\t<CellData Scalars="scalars">\n')
        s.extend(_double_array(n, a) for n, a in cell_data)
        s.append("\t\t\t</CellData>\n")
    s.append("\t\t\t<Points>\n")
    s.append(_double_array("Position", points))
    s.append("\t\t\t</Points>\n")
    s.append("\t\t\t<Cells>\n")
    s.append(_int_array("connectivity", conn))
    s.append(_int_array("offsets", (np.arange(len(conn)) + 1) * nv))
    s.append(_int_array("types", np.full(len(conn), ctype)))
    s.append("\t\t\t</Cells>\n")
    s.append("\t\t</Piece>\n")
    s.append("\t</UnstructuredGrid>\n")
    s.append("</VTKFile>\n")
    return "".join(s)


def surface_vtu(points, conn, f, center, bez):
    """surfaceData.vtu text for one condition's fields f (Geometry.case_fields)."""
    return vtu_text(points, conn, [("Doublet Strengths", f["mu"]), ("Source Strengths", f["sigma"]),
                                   ("Velocity Potential", f["potential"]), ("Velocity", f["velocity"]),
                                   ("Cp", f["cp"]), ("bezNormals", bez), ("xPosition", center[:, 0]),
                                   ("yPosition", center[:, 1]), ("zPostition", center[:, 2])])


def wake_vtu(points, conn, f):
    return vtu_text(points, conn, [("Doublet Strengths", f["wake_mu"]), ("Velocity Potential", f["wake_potential"])])


def spanwise_csv(span, bref):
    """spanwiseData_Wake<i>.csv text of one wake: span is (stations, 3) with
    y, Cl/PG, Cdi/PG^2 per station."""
    lines = ["2y/b,Cl,Cdi"]
    for y, cl, cd in span:
        lines.append("%s,%s,%s" % (_g(2 * y / bref), _g(cl), _g(cd)))
    return "\n".join(lines) + "\n"


def write_case(root, geom, mesh, condition, mach, fields, bref):
    """Writes one case's directory under root (cpCase::writeFiles) and
    returns its path.  geom: the Geometry the case was solved on; mesh: its
    input mesh dict (xyz, tri); condition: (V, alpha_deg, beta_deg); fields:
    that condition's entry of Geometry.case_fields."""
    V, alpha, beta = condition
    d = os.path.join(root, _label(V, mach, alpha, beta))
    os.makedirs(d, exist_ok=True)
    xyz = np.asarray(mesh["xyz"], dtype=np.float64)
    tri = np.asarray(mesh["tri"])
    b = geom.body()
    with open(os.path.join(d, "surfaceData.vtu"), "w") as fh:
        fh.write(surface_vtu(xyz, tri[b["tri_index"]], fields, b["center"], b["normal"]))
    if geom.nw > 0:
        w = geom.wake()
        with open(os.path.join(d, "wakeData.vtu"), "w") as fh:
            fh.write(wake_vtu(xyz, tri[w["tri_index"]], fields))
        for i, sp in enumerate(fields["spanwise"]):
            with open(os.path.join(d, "spanwiseData_Wake%d.csv" % (i + 1)), "w") as fh:
                fh.write(spanwise_csv(sp, bref))
    return d


def read_vtu_arrays(path):
    """name -> array of a VTU written by vtu_text (or by the reference)."""
    out = {}
    name, rows = None, None
    with open(path) as fh:
        for line in fh:
            t = line.strip()
            if t.startswith("<DataArray"):
                name = t.split('Name="')[1].split('"')[0]
                rows = []
            elif t.startswith("</DataArray>"):
                out[name] = np.array(rows, dtype=np.float64)
                name = None
            elif name is not None and t:
                rows.append([float(v) for v in t.split()])
    return out
