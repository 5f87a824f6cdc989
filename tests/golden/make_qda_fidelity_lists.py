"""Copy the reference's QDA fidelity golden lists (data) into qda_fidelity_lists.json.

Source: test/CPUTest/CommonTest/CorrectnessTest_QDA_CompareList.inl:1-4 of the
reference repository (4 lists of 371 fidelities: nqubit 3, step_rate 0.01,
p 1.3, alpha 2, beta -1/+1, tridiagonal and QRAM block encodings).
Run in a checkout that has the reference mounted at /root/reference.
"""
import json
import os
import re

SRC = "/root/reference/test/CPUTest/CommonTest/CorrectnessTest_QDA_CompareList.inl"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "qda_fidelity_lists.json")

lists = {}
for line in open(SRC):
    m = re.match(r"#define\s+CorrectnessTest_QDA_CompareList(\w+)\s+(.*)", line)
    if m:
        lists[m.group(1)] = [float(x) for x in m.group(2).replace(" ", "").split(",") if x]
meta = {"nqubit": 3, "step_rate": 0.01, "p": 1.3, "alpha": 2.0,
        "beta": {"Neg_Tri": -1.0, "Pos_Tri": 1.0, "Neg_QRAM": -1.0, "Pos_QRAM": 1.0},
        "tolerance": 1e-5, "source": "test/CPUTest/CommonTest/CorrectnessTest_QDA_CompareList.inl:1-4, driver :105-202"}
json.dump({"meta": meta, "lists": lists}, open(OUT, "w"), indent=0)
print({k: len(v) for k, v in lists.items()})
