"""Writes tests/golden/sshape_crown.json from the reference's own golden
files (plain tensors, loaded with weights_only=True):

* solver/tests/data_64/s_shape_test_data entries 0 and 2: the FP64 CROWN
  bounds of sin(x) over 19 intervals of width pi and 19 of width pi/2; the
  inputs are those of solver/tests/test_s_shaped.py:70-80
  (linspace(-10, 9, 19) pi .. linspace(-9, 10, 19) pi, then halved);
* solver/tests/data_64/pow_test_data entries 0-11: IBP (even entries) and
  CROWN (odd entries) bounds of x^p for non-integer p over the six boxes of
  solver/tests/test_pow.py:596-607.

Run in a container where the reference is mounted; the JSON is what the
tests read."""
import json
import math
import sys

import torch

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
s = torch.load(f"{REF}/solver/tests/data_64/s_shape_test_data", weights_only=True)
p = torch.load(f"{REF}/solver/tests/data_64/pow_test_data", weights_only=True)
lb = (torch.linspace(-10, 9, 19, dtype=torch.float64) * math.pi).tolist()
ub = (torch.linspace(-9, 10, 19, dtype=torch.float64) * math.pi).tolist()
cases = [(0.5, 1, 4), (0.3, 1, 5), (1.5, 1, 4), (2.5, 1, 3), (0.5, 0, 4), (1.5, 0, 4)]
fx = {
    "source": "solver/tests/data_64/s_shape_test_data entries 0 (width pi) and 2 (width pi/2), "
              "CROWN of sin, inputs solver/tests/test_s_shaped.py:70-80; "
              "solver/tests/data_64/pow_test_data entries 0-11 (IBP, CROWN of x^p, "
              "solver/tests/test_pow.py:596-607)",
    "sin": [{"input_lb": lb, "input_ub": ub,
             "lb": s[0][0].flatten().tolist(), "ub": s[0][1].flatten().tolist()},
            {"input_lb": [v / 2 for v in lb], "input_ub": [v / 2 for v in ub],
             "lb": s[2][0].flatten().tolist(), "ub": s[2][1].flatten().tolist()}],
    "pow_noninteger": [{"exponent": e, "x_low": a, "x_high": b,
                        "ibp": [p[2 * k][0].flatten()[0].item(), p[2 * k][1].flatten()[0].item()],
                        "crown": [p[2 * k + 1][0].flatten()[0].item(),
                                  p[2 * k + 1][1].flatten()[0].item()]}
                       for k, (e, a, b) in enumerate(cases)],
}
json.dump(fx, open(__file__.replace("make_sshape_fixture.py", "sshape_crown.json"), "w"), indent=1)
