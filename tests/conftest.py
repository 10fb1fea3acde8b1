import json
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "2d-turbulent-navier-stokes_amd")
GOLDEN = os.path.join(REPO, "tests", "golden")
for p in (PKG, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running")


class Golden:
    """One fixture written by tests/golden/make_golden.py."""

    def __init__(self, name):
        from tdt import caseio
        self.name = name
        with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
            self.z = {k: z[k] for k in z.files}
        if "mesh_workload" in self.z:   # a generated grid, named by its digest
            from tdt import gridgen
            self.mesh = gridgen.WORKLOADS[str(self.z["mesh_workload"])][0]()
            assert self.mesh.digest() == str(self.z["mesh_digest"]), name + ": grid differs"
        else:
            self.mesh = caseio.Mesh.from_arrays({k[2:]: v for k, v in self.z.items()
                                                 if k.startswith("m_")})
        self.flow = caseio.Flow(**json.loads(str(self.z["flow_json"])))
        self.nrun = int(self.z["nrun"])

    def has(self, key):
        return key in self.z

    def __getitem__(self, key):
        return self.z[key]

    def dump(self, tag):
        return self.z["d_" + tag]


DUMP_CASES = ["t2d_chan_cold", "t2d_chan", "t2d_ramp", "taxi_duct", "t7_plate", "t6_port",
              "t6_port_axi"]
RUN_CASES = ["case1", "case1_axi", "duct_385x129"]


@pytest.fixture(scope="session")
def golden():
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = Golden(name)
        return cache[name]
    return get
