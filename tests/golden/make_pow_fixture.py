"""Writes tests/golden/pow_crown_even.json from the reference's own golden file
solver/tests/data_64/pow_test_data (plain tensors, loaded with
weights_only=True): entries 15 and 17 are the FP64 CROWN bounds of x^4 and
x^6 over the five intervals of solver/tests/test_pow.py:83-88.  Run in a
container where the reference is mounted; the JSON is what the tests read."""
import json
import sys

import torch

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
d = torch.load(f"{REF}/solver/tests/data_64/pow_test_data", weights_only=True)
lbs = [-7.681, -6.174, -2.0, 0.2, 3.919]
ubs = [-4.414, -1.939, 3.0, 2.5, 5.997]
fx = {"source": "solver/tests/data_64/pow_test_data entries 15 (x^4, CROWN) and 17 (x^6, CROWN); "
                "inputs from solver/tests/test_pow.py:83-88",
      "input_lb": lbs, "input_ub": ubs,
      "cases": [{"exponent": 4, "lb": d[15][0].flatten().tolist(), "ub": d[15][1].flatten().tolist()},
                {"exponent": 6, "lb": d[17][0].flatten().tolist(), "ub": d[17][1].flatten().tolist()}]}
json.dump(fx, open(__file__.replace("make_pow_fixture.py", "pow_crown_even.json"), "w"), indent=1)
