"""Writes triangle_integrals.json from the reference's test data file
UnitTest_PZ/TestIntegNum/Triangle.txt (read from /root/reference, run once in
the build container): the exact integrals over the reference triangle
(0,0), (1,0), (0,1) of sum_{i+j<=p} (i+1)(xi+0.1)^i (j+1)(eta+0.1)^j for
order p = line index, as UnitTest_PZ/TestIntegNum/IntegNumUnitTest.cpp:327-420
evaluates them with TPZIntTriang."""
import json
import os

SRC = "/root/reference/UnitTest_PZ/TestIntegNum/Triangle.txt"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "triangle_integrals.json")

if __name__ == "__main__":
    with open(SRC) as f:
        vals = [ln.strip() for ln in f if ln.strip()]
    with open(OUT, "w") as f:
        json.dump({"source": "UnitTest_PZ/TestIntegNum/Triangle.txt (exact integrals over the reference "
                             "triangle of sum_{i+j<=p} (i+1)(xi+0.1)^i (j+1)(eta+0.1)^j for order p = "
                             "line index; checked in UnitTest_PZ/TestIntegNum/IntegNumUnitTest.cpp:"
                             "153-163, 371-420 with tol 6e-13, x10 above 10, x100 above 100)",
                   "values": vals}, f, indent=0)
