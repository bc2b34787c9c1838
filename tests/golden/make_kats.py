"""Writes the known-answer fixtures in this folder (inputs and expected
outputs, JSON).  Each case is one of the reference's own test inputs; its
expected output is fixed by the mathematics, not by running any code:

  prefix_sum_all_ones.json   test/algorithm/PrefixSumIntegerTest.java:30-36:
                             200 draws of Random().ints(n, 1, 2), i.e. all
                             ones; the exclusive scan is 0..199, total 200.
  simplegrid_impulse.json    test/lightgrid/TestSimpleGrid.java:39-56: a
                             16 x 16 grid, 0.01f at row 11, column 10; its
                             inclusive 2-D prefix sum is 0.01f at rows >= 11
                             and columns >= 10, 0 elsewhere.
  lightgrid_all_ones.json    test/lightgrid/TestLightGrid.java:41-54: every
                             16 x 32 tile of the 1600 x 1600 grid all ones;
                             the tile SAT is (x + 1) * (y + 1) in every tile
                             (the 512 values of one tile are stored).
  distribution1d_case.json   test/sampling/Test.java:17: f = {0,1,0,1,1,1,0,0};
                             piecewise-constant sampling (pbrt) has CDF
                             {0,0,1/4,1/4,1/2,3/4,1,1,1}, picks cell i for u in
                             [cdf[i], cdf[i+1]), continuous pdf f[i] / mean(f)
                             = 2 on the picked cells, u' = (i + du) / 8.

Run from the repository root: python tests/golden/make_kats.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def write(name, obj):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, indent=1)
        f.write("\n")


def main():
    n = 200
    write("prefix_sum_all_ones.json", {"input": [1] * n, "exclusive_scan": list(range(n)), "total": n})

    func = [[0.0] * 16 for _ in range(16)]
    func[11][10] = 0.01
    sat = [[0.01 if (y >= 11 and x >= 10) else 0.0 for x in range(16)] for y in range(16)]
    write("simplegrid_impulse.json", {"nu": 16, "nv": 16, "rx": 16, "ry": 16, "func_rows": func, "sat_rows": sat})

    tile_sat = [[(x + 1) * (y + 1) for x in range(16)] for y in range(32)]
    write("lightgrid_all_ones.json", {"nu": 1600, "nv": 1600, "rx": 16, "ry": 32, "tile_value": 1.0,
                                      "tile_sat_rows": tile_sat})

    f = [0, 1, 0, 1, 1, 1, 0, 0]
    total = sum(f)
    cdf = [0.0]
    for v in f:
        cdf.append(cdf[-1] + v / total)
    cases = []
    for u in [0.0, 0.1, 0.2, 0.3, 0.45, 0.55, 0.7, 0.8, 0.95]:
        i = max(k for k in range(8) if cdf[k] <= u and cdf[k + 1] > u)
        du = (u - cdf[i]) / (cdf[i + 1] - cdf[i])
        cases.append({"u": u, "offset": i, "u_out": (i + du) / 8, "pdf": f[i] / (total / 8)})
    write("distribution1d_case.json", {"func": f, "cdf": cdf, "samples": cases})


if __name__ == "__main__":
    main()
