"""The .CPout summary writer reproduces the reference's RESULTS text
(4412AR6.CPout, written by caseMgr::writeCase) from the same numbers."""
import os

from cpanel_amd import cpout


def test_case_block_matches_reference_text(golden, fixtures_json):
    fx = fixtures_json["4412AR6"]
    p = fx["params"]
    res = dict(CL=fx["CL"], CDi=fx["CDi"], Fbody=[fx["CA"], fx["CY"], fx["CN"]],
               Fwind=[fx["CD_surface"], fx["CY"], fx["CL_surface"]], CM=[fx["Cl"], fx["Cm"], fx["Cn"]])
    stab = dict(dF_dAlpha=[0.0, fx["CY_alpha"], fx["CL_alpha"]], dF_dBeta=[0.0, fx["CY_beta"], fx["CL_beta"]],
                dM_dAlpha=[fx["Cl_alpha"], fx["Cm_alpha"], fx["Cn_alpha"]],
                dM_dBeta=[fx["Cl_beta"], fx["Cm_beta"], fx["Cn_beta"]])
    text = "-----------RESULTS-----------\n\n" + cpout.case_block(1, p["V"], p["mach"], p["alpha"], p["beta"], res,
                                                                  stab)
    ref = open(os.path.join(golden, "4412AR6_results.txt")).read()
    assert text == ref


def test_stability_derivative_differences():
    r0 = dict(CL=0.5, CDi=0.02, Fwind=[0.03, 0.001, 0.5], CM=[0.0, -0.07, 0.0])
    ra = dict(CL=0.54, CDi=0.022, Fwind=[0.031, 0.0011, 0.54], CM=[0.001, -0.069, 0.0002])
    rb = dict(CL=0.4999, CDi=0.02, Fwind=[0.03, 0.004, 0.4999], CM=[0.004, -0.07, -0.0003])
    d = cpout.stability_derivatives(r0, ra, rb, 0.5)
    import math
    dr = 0.5 * math.pi / 180
    assert abs(d["dF_dAlpha"][2] - 0.04 / dr) < 1e-12
    assert abs(d["dF_dBeta"][1] - 0.003 / dr) < 1e-12
    assert abs(d["dM_dAlpha"][1] - 0.001 / dr) < 1e-12
