"""Case summary in the reference's .CPout text format.

The RESULTS blocks follow caseMgr::writeCase (CPanelMgr.cpp:119-167): left
justified, fixed notation, 3 decimals for the flow conditions and 5 for the
coefficients, the same field widths (11, 7, 8, 8 for the flow conditions; 15
and 10 for the force table; 12 for moments and stability derivatives) and the
same tab structure, so a summary written here diffs line by line against one
the reference wrote.  The INPUTS block mirrors inputParams::print's layout
(labels right-aligned to 27 columns, then " -> ").
"""
import math


def _f(v, w, p):
    return ("%.*f" % (p, v)).ljust(w)


def _g(v):
    # default ostream formatting of a double (6 significant digits, no
    # trailing zeros)
    return "%g" % v


def inputs_block(geometry_file, Sref, bref, cref, cg, V, alpha, beta, mach, stab=False, vortex_particles=False,
                 unsteady=False):
    rows = [("Geometry File", geometry_file),
            ("Reference Area", "%s ft^2" % _g(Sref)),
            ("Reference Span", "%s ft" % _g(bref)),
            ("Reference Chord", "%s ft" % _g(cref)),
            ("Center of Gravity", "[%s] ft" % " ".join(_g(c) for c in cg)),
            ("Velocity", "%s ft/s" % _g(V)),
            ("Alpha", "%s degrees" % _g(alpha)),
            ("Beta", "%s degrees" % _g(beta)),
            ("Mach #", "%s " % _g(mach)),
            ("Surface Streamlines", "OFF"),
            ("Stability Derivatives", "ON" if stab else "OFF"),
            ("Write Inf Coeff to File", "OFF"),
            ("Vortex Particle Wake", "ON" if vortex_particles else "OFF"),
            ("Accelerate Code", "ON"),
            ("Unsteady Mode", "ON" if unsteady else "OFF")]
    return "".join("%s -> %s\n" % (k.rjust(27), v) for k, v in rows)


def case_block(number, V, mach, alpha, beta, res, stab=None):
    """One case (caseMgr::writeCase).  res: dict with CL, CDi, Fbody[3],
    Fwind[3], CM[3]; stab: dict from Geometry.stability_derivatives or None."""
    L = []
    L.append("")
    L.append("---Case #%d---" % number)
    L.append("\n\t--Flow Conditions--\n")
    L.append("\t\t" + "Velocity".ljust(11) + "Mach".ljust(7) + "Alpha".ljust(8) + "Beta".ljust(8))
    L.append("\t\t" + _f(V, 11, 3) + _f(mach, 7, 3) + _f(alpha, 8, 3) + _f(beta, 8, 3))
    L.append("\n\t--Force Coefficients--\n")
    L.append("\t\t-Trefftz Plane-")
    L.append("\t\t\tCL = %.5f\tCDi = %.5f" % (res["CL"], res["CDi"]))
    L.append("\n\t\t-Surface Integrated Force Coefficients-")
    fb, fw, cm = res["Fbody"], res["Fwind"], res["CM"]
    L.append("\t\t\t" + "Body Axis".ljust(15) + "CN".ljust(10) + "CA".ljust(10) + "CY".ljust(10))
    L.append("\t\t\t" + " ".ljust(15) + _f(fb[2], 10, 5) + _f(fb[0], 10, 5) + _f(fb[1], 10, 5))
    L.append("\t\t\t" + "Wind Axis".ljust(15) + "CL".ljust(10) + "CD".ljust(10) + "CY".ljust(10))
    L.append("\t\t\t" + " ".ljust(15) + _f(fw[2], 10, 5) + _f(fw[0], 10, 5) + _f(fw[1], 10, 5))
    L.append("\n\t--Moment Coefficients--\n")
    L.append("\t\t" + "Cm (pitch)".ljust(12) + "Cl (roll)".ljust(12) + "Cn (yaw)".ljust(12))
    L.append("\t\t" + _f(cm[1], 12, 5) + _f(cm[0], 12, 5) + _f(cm[2], 12, 5))
    if stab is not None:
        L.append("\n\t--Stability Derivatives--\n")
        da, db = stab["dF_dAlpha"], stab["dF_dBeta"]
        ma, mb = stab["dM_dAlpha"], stab["dM_dBeta"]
        L.append("\t\t" + "".join(s.ljust(12) for s in ("CL_alpha", "CY_alpha", "Cm_alpha", "Cl_alpha", "Cn_alpha")))
        L.append("\t\t" + "".join(_f(v, 12, 5) for v in (da[2], da[1], ma[1], ma[0], ma[2])))
        L.append("\t\t" + "".join(s.ljust(12) for s in ("CL_beta", "CY_beta", "Cm_beta", "Cl_beta", "Cn_beta")))
        L.append("\t\t" + "".join(_f(v, 12, 5) for v in (db[2], db[1], mb[1], mb[0], mb[2])))
        L.append("\n\n")
    return "\n".join(L) + "\n"


def summary(inputs, cases):
    """Whole .CPout text: inputs (kwargs of inputs_block) and a list of
    (V, mach, alpha, beta, res, stab) tuples."""
    out = "-----------INPUTS-----------\n\n"
    out += inputs_block(**inputs)
    out += "\n\n"
    out += "-----------RESULTS-----------\n\n"
    for i, (V, mach, a, b, res, stab) in enumerate(cases):
        out += case_block(i + 1, V, mach, a, b, res, stab)
    return out


def write(path, inputs, cases):
    with open(path, "w") as f:
        f.write(summary(inputs, cases))


def stability_derivatives(r0, ra, rb, delta_deg=0.5):
    """cpCase::stabilityDerivatives (cpCase.cpp:635-665): finite differences
    over delta_deg in alpha and beta of F = (CDi, CY_wind, CL) and CM."""
    drad = delta_deg * math.pi / 180

    def F(r):
        return [r["CDi"], r["Fwind"][1], r["CL"]]

    f0, fa, fb = F(r0), F(ra), F(rb)
    return dict(dF_dAlpha=[(a - b) / drad for a, b in zip(fa, f0)],
                dF_dBeta=[(a - b) / drad for a, b in zip(fb, f0)],
                dM_dAlpha=[(a - b) / drad for a, b in zip(ra["CM"], r0["CM"])],
                dM_dBeta=[(a - b) / drad for a, b in zip(rb["CM"], r0["CM"])])
