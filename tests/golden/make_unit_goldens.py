"""Extracts the golden vectors of the reference's unit tests into tests/golden/unit_goldens.json.

Source (data only): period_search_pure_cpu_win/UnitTest_pure_cpu_win/arrangers.cpp of the reference,
the numeric arrays its MSVC unit tests (InnerFunctionsTests.cpp) compare against, dumped by the
reference authors from the production code for the 118-point fixture input in the same directory
(Lmax = Mmax = 3, nrows = 6).  Only the numbers are kept; index layouts follow the fixture functions:

  get_t/get_f        vertex directions, 1-based [1..146]           (TestAreanorm inputs)
  get_at/get_af      facet normal angles, 1-based [1..288]          (TestAreanorm)
  get_ifp            trifac vertex triplets, 1-based                (TestTrifac)
  get_dsph           Dsph[i][k], facets 1..146, k = 1..16           (TestSphfunc)
  get_pleg           Pleg[i][l][m], facets 1..288                   (TestSphfunc)
  get_cg_first       ellfit result, 1-based [1..16]                 (TestEllfit)
  get_blmat/get_dblm Blmat[1..3][1..3], Dblm[1..2][1..3][1..3]      (TestBlmatrix)
  get_tmat/get_dtm   tmat[1..3][1..3], dtm[1..3][1..3][1..3]        (TestMatrix)
  get_dphp           dphp[1..3] (+ Scale from the test body)        (TestPhasec)
  get_area, get_dg   Area[0..287], Dg[0..287][0..15] inputs          (TestBright)
  get_dyda           bright() dyda, 0-based [0..23], ymod from test (TestBright)
  get_dg_curv, get_darea, get_nor_0, get_dyda_after_conv            (TestConv)
  get_xx1/get_xx2/get_tim/get_cg24                                  (shared inputs)

Run in a container that has the reference checkout:  python tests/golden/make_unit_goldens.py
"""
import json
import os
import re
import sys

REF = "/root/reference/period_search_pure_cpu_win/UnitTest_pure_cpu_win"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "unit_goldens.json")


def parse_arrays(path):
    src = open(path, encoding="latin-1").read()
    funcs = {}
    for chunk in re.split(r"\n(?=void get_)", src):
        m = re.match(r"void (get_\w+)\(", chunk)
        if not m:
            continue
        arrays = {}
        for am in re.finditer(r"(?:double|int)\s+(\w+)\s*\[[^\]]*\]\s*=\s*\{([^}]*)\}", chunk):
            vals = [v.strip() for v in am.group(2).replace("\n", " ").split(",") if v.strip()]
            arrays[am.group(1)] = [float(v) for v in vals]
        funcs[m.group(1)] = arrays
    return funcs


def main():
    f = parse_arrays(os.path.join(REF, "arrangers.cpp"))
    g = {}
    g["vertex_t"] = f["get_t"]["_t"][1:147]
    g["vertex_f"] = f["get_f"]["_f"][1:147]
    g["areanorm_at"] = f["get_at"]["_at"][1:289]
    g["areanorm_af"] = f["get_af"]["_af"][1:289]
    ifp = f["get_ifp"]
    g["trifac_ifp"] = [[int(ifp["_ifp_%d" % v][i]) for v in (1, 2, 3)] for i in range(1, 289)]
    ds = f["get_dsph"]
    g["sphfunc_dsph_1_146"] = [[ds["_dsph_%d" % k][i] for k in range(1, 17)] for i in range(1, 147)]
    pl = f["get_pleg"]
    g["sphfunc_pleg"] = {"%d,%d" % (l, m): pl["_pleg_j%d_k%d" % (l, m)][1:289] for l in range(4) for m in range(4)}
    g["ellfit_cg_first"] = f["get_cg_first"]["_cg_first_local"][1:17]
    # TestEllfit/TestCurv inputs: escl from ave, al0 and onlyrel = 0
    g["ellfit_inputs"] = {"ave": 0.382008772881355651573898057904, "al0": 0.069555451868686835048549710336,
                          "onlyrel": 0}
    cg = f["get_cg24"]["_cg"]
    g["cg24"] = cg[1:25]
    bm = f["get_blmat_m"]
    g["blmat"] = [[bm["_blmat_%d" % k][i] for k in (1, 2, 3)] for i in (1, 2, 3)]
    dm = f["get_dblm_m"]
    g["dblm"] = [[[dm["_dblm_%d_%d" % (j, k)][i] for k in (1, 2, 3)] for j in (1, 2, 3)] for i in (1, 2)]
    tm = f["get_tmat"]
    g["tmat"] = [[tm["_tmat_%d" % k][i] for k in (1, 2, 3)] for i in (1, 2, 3)]
    dt = f["get_dtm"]
    g["dtm"] = [[[dt["_dtm_%d_%d" % (j, k)][i] for k in (1, 2, 3)] for j in (1, 2, 3)] for i in (1, 2, 3)]
    g["dphp"] = f["get_dphp"]["__dphp"][1:4]
    g["phasec_scale"] = 1.214621163362573641464337015350  # InnerFunctionsTests.cpp TestPhasec
    g["xx1"] = f["get_xx1"]["__xx1"][1:4]
    g["xx2"] = f["get_xx2"]["__xx2"][1:4]
    g["tim1"] = f["get_tim"]["_tim"][1]
    g["bright_area"] = f["get_area"]["_area_a"][0:288]
    dg = f["get_dg"]
    g["bright_dg"] = [[dg["_dg_%d" % k][i] for k in range(16)] for i in range(288)]
    g["bright_dyda"] = f["get_dyda"]["__dyda"][0:24]
    g["bright_ymod"] = 0.34207493045125292  # InnerFunctionsTests.cpp TestBright
    dgc = f["get_dg_curv"]
    g["conv_dg"] = [[dgc["__Dg_conv_%d" % k][i] for k in range(16)] for i in range(288)]
    g["conv_darea"] = f["get_darea"]["__Darea"][0:288]
    g["conv_nor_x"] = f["get_nor_0"]["__Nor"][0:288]
    g["conv_dres"] = f["get_dyda_after_conv"]["__dyda"][0:16]
    g["conv_res"] = 6.2124784483419404e-17  # InnerFunctionsTests.cpp TestConv
    g["mrqcof_chisq_commented_test"] = 7.8099025458229479  # InnerFunctionsTests.cpp (commented-out test)
    json.dump(g, open(OUT, "w"), indent=0)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    sys.exit(main())
