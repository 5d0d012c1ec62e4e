"""Copy the reference's whole-run regression targets into a JSON fixture.

The targets are data held by the reference's own tests
(src/tests/test_targets.yml): for each test case, the float32 sums of the
state arrays rho, rhou, rhov, rhow, rhoY, rhoX and p2_nodes at the end of the
run, checked with numpy.isclose by src/tests/diagnostics.py:42-58.  Run in a
container that holds the reference tree:

    python tests/golden/make_targets.py /root/reference/src/tests/test_targets.yml
"""
import json
import os
import sys

import yaml


def main(path):
    with open(path) as f:
        targets = yaml.safe_load(f)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_targets.json")
    with open(out, "w") as f:
        json.dump({"source": "src/tests/test_targets.yml", "targets": targets}, f, indent=1, sort_keys=True)
    print(out)


if __name__ == "__main__":
    main(sys.argv[1])
