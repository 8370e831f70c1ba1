"""Writes tests/golden/testsaver.json from the reference's own test fixture
(tests/testSaver/state.npy + scalars.dat: two pentagons, maxEdgeLength 0.2) and the
outputs its notebook recorded for it (tests/neighborTesting.ipynb: neighbours of edge 1
are [7, 8] with inside flags [True, False]).  Run in the container that holds the
reference; only the data file is read (numpy.load without pickles)."""
import json
import os

import numpy as np

REF = "/root/reference/tests/testSaver"
HERE = os.path.dirname(os.path.abspath(__file__))

state = np.load(os.path.join(REF, "state.npy"), allow_pickle=False)
scalars = {}
with open(os.path.join(REF, "scalars.dat")) as f:
    for line in f:
        k, v = line.rstrip("\n").split("\t")
        scalars[k.rstrip(":")] = v
V = int(scalars["numVertices"])
out = {
    "source": "reference tests/testSaver/state.npy, scalars.dat; tests/neighborTesting.ipynb cells 3-6, 9, 14, 20",
    "nArray": [int(x) for x in state[:-2 * V]],
    "positions": [float(x) for x in state[-2 * V:]],
    "maxEdgeLength_saved": float(scalars["maxEdgeLength"]),
    "maxEdgeLength_notebook": 0.25,
    "expected_neighbors": {"1": [7, 8]},
    "expected_inside": {"1": [True, False]},
}
with open(os.path.join(HERE, "testsaver.json"), "w") as f:
    json.dump(out, f, indent=1)
print("wrote", os.path.join(HERE, "testsaver.json"))
