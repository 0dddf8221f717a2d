"""Extracts the reference-held fixtures used by the tests into tests/golden/.

Reads data files from the reference checkout (Connor Cases/, CPanelTests/);
writes small fixtures:
  4412.tri, 4412AR6.tri             - meshes (copied data files)
  4412_surface.npz                  - per-panel mu, sigma, phi, V, Cp, centre (surfaceData-0.vtu)
  4412_wake_mu.npy                  - wake doublet strengths (wakeData-0.vtu)
  4412_spanwise.npy                 - 2y/b, Cl, Cdi (spanwiseData_Wake1.csv)
  4412AR6_surface.npz               - per-panel arrays at alpha 5, beta 0.5 (see fixtures.json)
  fixtures.json                     - .CPout numbers and the kernel known-answer table
  4412AR6_results.txt               - the RESULTS section of 4412AR6.CPout (output text format)
  steadyTest.tri                    - vortex-particle case mesh (copied data file)
  steadyTest_vp.npz                 - per-step arrays of that case (surfaceData-k, wakeData-k,
                                      bufferWake2Data-k, particleData-k .vtu; FP32, 6 digits):
                                      buffer-wake cell vertices, body mu/sigma, wake rows' mu,
                                      particle positions/strengths/shed steps

Run: python tests/golden/make_golden.py [/root/reference]
"""
import json
import os
import re
import shutil
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
CASES = os.path.join(REF, "Connor Cases")


def vtu_arrays(path, names):
    txt = open(path).read()
    out = {}
    for n in names:
        m = re.search('Name="%s"[^>]*>(.*?)</DataArray>' % n, txt, re.S)
        out[n] = np.array(m.group(1).split(), float)
    return out


VP_STEPS = [1, 2, 3, 4, 5, 6, 8, 10, 20, 40]


def steady_test_vp():
    d = os.path.join(CASES, "steadyTest")
    shutil.copy(os.path.join(d, "steadyTest.tri"), os.path.join(OUT, "steadyTest.tri"))
    r = os.path.join(d, "V1_Mach0.1_alpha10_beta0")
    out = {}
    for row, fname in (("w1", "wakeData-1.vtu"), ("w2", "bufferWake2Data-1.vtu")):
        a = vtu_arrays(os.path.join(r, fname), ["Position", "connectivity"])
        pts = a["Position"].reshape(-1, 3)
        con = a["connectivity"].astype(int).reshape(-1, 4)
        out[row + "_cells"] = pts[con].astype(np.float32)
    for k in VP_STEPS:
        s = vtu_arrays(os.path.join(r, "surfaceData-%d.vtu" % k), ["Doublet Strengths", "Source Strengths"])
        out["mu_%d" % k] = s["Doublet Strengths"].astype(np.float32)
        out["sigma_%d" % k] = s["Source Strengths"].astype(np.float32)
        out["w1mu_%d" % k] = vtu_arrays(os.path.join(r, "wakeData-%d.vtu" % k),
                                        ["Doublet Strengths"])["Doublet Strengths"].astype(np.float32)
        out["w2mu_%d" % k] = vtu_arrays(os.path.join(r, "bufferWake2Data-%d.vtu" % k),
                                        ["Doublet Strengths"])["Doublet Strengths"].astype(np.float32)
        if k >= 3:
            p = vtu_arrays(os.path.join(r, "particleData-%d.vtu" % k), ["Strength", "Time Step Shed", "Position"])
            out["ppos_%d" % k] = p["Position"].reshape(-1, 3).astype(np.float32)
            out["pstr_%d" % k] = p["Strength"].reshape(-1, 3).astype(np.float32)
            out["pshed_%d" % k] = p["Time Step Shed"].astype(np.int16)
    np.savez_compressed(os.path.join(OUT, "steadyTest_vp.npz"), **out)


def main():
    steady_test_vp()
    shutil.copy(os.path.join(CASES, "4412", "4412.tri"), os.path.join(OUT, "4412.tri"))
    shutil.copy(os.path.join(CASES, "4412AR6", "4412AR6.tri"), os.path.join(OUT, "4412AR6.tri"))
    txt = open(os.path.join(CASES, "4412AR6", "4412AR6.CPout")).read()
    with open(os.path.join(OUT, "4412AR6_results.txt"), "w") as f:
        f.write(txt[txt.index("-----------RESULTS-----------"):])
    names = ["Doublet Strengths", "Source Strengths", "Velocity Potential", "Velocity", "Cp", "xPosition",
             "yPosition", "zPostition"]
    s = vtu_arrays(os.path.join(CASES, "4412", "V1_Mach0.1_alpha5_beta0", "surfaceData-0.vtu"), names)
    np.savez_compressed(os.path.join(OUT, "4412_surface.npz"), mu=s["Doublet Strengths"],
                        sigma=s["Source Strengths"], phi=s["Velocity Potential"],
                        vel=s["Velocity"].reshape(-1, 3), cp=s["Cp"],
                        center=np.stack([s["xPosition"], s["yPosition"], s["zPostition"]], 1))
    w = vtu_arrays(os.path.join(CASES, "4412", "V1_Mach0.1_alpha5_beta0", "wakeData-0.vtu"),
                   ["Doublet Strengths"])
    np.save(os.path.join(OUT, "4412_wake_mu.npy"), w["Doublet Strengths"])
    csv = np.loadtxt(os.path.join(CASES, "4412", "V1_Mach0.1_alpha5_beta0", "spanwiseData_Wake1.csv"),
                     delimiter=",", skiprows=1)
    np.save(os.path.join(OUT, "4412_spanwise.npy"), csv)
    s6 = vtu_arrays(os.path.join(CASES, "4412AR6", "V1_Mach0.1_alpha5_beta0", "surfaceData.vtu"),
                    ["Doublet Strengths", "Source Strengths", "Cp"])
    np.savez_compressed(os.path.join(OUT, "4412AR6_surface.npz"), mu=s6["Doublet Strengths"],
                        sigma=s6["Source Strengths"], cp=s6["Cp"])
    fixtures = {
        "4412": {
            "source": "Connor Cases/4412/4412.CPin, 4412.CPout:34-45",
            "params": {"V": 1, "alpha": 5, "beta": 0, "mach": 0.1, "Sref": 6, "bref": 6, "cref": 1,
                       "cg": [0.25, 0, 0]},
            "CL": 0.67565, "CDi": 0.03186, "CN": 0.67724, "CA": -0.02922, "CY": 0.00003,
            "CL_surface": 0.67720, "CD_surface": 0.02991, "Cm": -0.08044, "Cl": 0.00359, "Cn": -0.00019,
            "print_decimals": 5,
        },
        "4412AR6": {
            "source": "Connor Cases/4412AR6/4412AR6.CPin, 4412AR6.CPout:33-51",
            "params": {"V": 1, "alpha": 5, "beta": 0, "mach": 0.1, "Sref": 6, "bref": 6, "cref": 1,
                       "cg": [0.25, 0, 0]},
            "CL": 0.65503, "CDi": 0.02911, "CN": 0.65709, "CA": -0.02743, "CY": 0.00022,
            "CL_surface": 0.65698, "CD_surface": 0.02994, "Cm": -0.07303, "Cl": -0.00032, "Cn": 0.00047,
            "CL_alpha": 4.23776, "CY_alpha": 0.00391, "Cm_alpha": 0.06230, "Cl_alpha": 0.00132,
            "Cn_alpha": 0.01123, "CL_beta": -0.00297, "CY_beta": 0.03141, "Cm_beta": 0.00068,
            "Cl_beta": 0.04724, "Cn_beta": -0.00311,
            "vtu_state": "alpha 5, beta 0.5 (stability-derivative run overwrites panel state, cpCase.cpp:635-665)",
            "print_decimals": 5,
        },
        "steadyTest": {
            "source": "Connor Cases/steadyTest/steadyTest.CPin, steadyTest.CPout (vortex-particle wake)",
            "params": {"V": 1, "alpha": 10, "beta": 0, "mach": 0.1, "Sref": 6, "bref": 6, "cref": 1,
                       "cg": [0.416, 0, 0.03], "steps": 40, "time_step": 0, "accelerate": 1},
            "steps_stored": VP_STEPS,
            "CL": 1.10020, "CDi": 0.06869, "CN": 1.11715, "CA": -0.10571, "CY": 0.00034,
            "CL_surface": 1.11853, "CD_surface": 0.08989, "Cm": 0.11546, "Cl": -0.00117, "Cn": -0.00052,
            "print_decimals": 5,
        },
        "kat_triangle": {
            "source": "CPanelTests/influenceTests.cpp:598-668 (commented-out testPntClose/testPntFar/"
                      "testCollocationPnt); triangle (1,0,0),(-1,1,0),(-1,-1,0)",
            "verts": [[1, 0, 0], [-1, 1, 0], [-1, -1, 0]],
            "points": {"close": {"poi": [2, 2, 2], "phi_src": 0.043411, "phi_dub": 0.006645},
                       "far": {"poi": [7, 7, 7], "phi_src": 0.012919, "phi_dub": 0.0005958},
                       "collocation_dub": -0.5},
        },
    }
    with open(os.path.join(OUT, "fixtures.json"), "w") as f:
        json.dump(fixtures, f, indent=1)
    print("golden fixtures written to", OUT)


if __name__ == "__main__":
    main()
