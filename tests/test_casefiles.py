"""Per-case output files (cpanel_amd.casefiles): text layout of the
reference's VTUfile::write (VTUfile.cpp:20-131) and writeSpanwiseData
(cpCase.cpp:910-935), host only."""
import numpy as np

from cpanel_amd import casefiles


def test_label_matches_reference_directory_names():
    # the reference's case directories, e.g. Connor Cases/4412/V1_Mach0.1_alpha5_beta0
    assert casefiles._label(1.0, 0.1, 5.0, 0.0) == "V1_Mach0.1_alpha5_beta0"
    assert casefiles._label(1.0, 0.1, 5.5, -2.25) == "V1_Mach0.1_alpha5.5_beta-2.25"


def test_vtu_layout_and_round_trip():
    pts = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]])
    conn = np.array([[0, 1, 2], [1, 3, 2]])
    mu = np.array([0.078061612345, -1.5e-7])
    vel = np.array([[1.0, 2.0, 3.0], [0.5, 0.25, 0.125]])
    t = casefiles.vtu_text(pts, conn, [("Doublet Strengths", mu), ("Velocity", vel)])
    lines = t.split("\n")
    assert lines[0] == '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">'
    assert lines[2] == '\t\t<Piece NumberOfPoints="4" NumberOfCells="2">'
    assert lines[3] == '\t\t\t<CellData Scalars="scalars">'
    assert lines[4] == '\t\t\t\t<DataArray type="Float64" Name="Doublet Strengths" NumberOfComponents="1" Format="ascii">'
    # default stream formatting, 6 significant digits, a tab after each value
    assert lines[5] == "0.0780616\t"
    assert lines[6] == "-1.5e-07\t"
    assert "1\t2\t3\t" in lines
    assert '\t\t\t\t<DataArray type="Int32" Name="offsets" Format="ascii">' in lines
    i = lines.index('\t\t\t\t<DataArray type="Int32" Name="types" Format="ascii">')
    assert lines[i + 1:i + 3] == ["5\t", "5\t"]  # VTK triangles
    j = lines.index('\t\t\t\t<DataArray type="Int32" Name="offsets" Format="ascii">')
    assert lines[j + 1:j + 3] == ["3\t", "6\t"]
    assert t.endswith("</VTKFile>\n")


def test_vtu_reader(tmp_path):
    pts = np.random.default_rng(0).normal(size=(5, 3))
    conn = np.array([[0, 1, 2], [2, 3, 4]])
    a = np.array([1.25, -3.5])
    p = tmp_path / "x.vtu"
    p.write_text(casefiles.vtu_text(pts, conn, [("A", a)]))
    v = casefiles.read_vtu_arrays(str(p))
    assert np.array_equal(v["A"][:, 0], a)
    assert np.allclose(v["Position"], pts, rtol=1e-5, atol=1e-6)
    assert np.array_equal(v["connectivity"], conn)


def test_spanwise_csv():
    span = np.array([[-3.0, 0.0, 0.0], [-2.9268, 0.0483945, 0.0162582], [0.0, 0.135118, 0.00508061]])
    t = casefiles.spanwise_csv(span, bref=6.0)
    assert t.split("\n")[:4] == ["2y/b,Cl,Cdi", "-1,0,0", "-0.9756,0.0483945,0.0162582", "0,0.135118,0.00508061"]
