"""Writes the golden fixtures of the widening rows: a Hittable scene, an HDRI
sky and a metric given as a tensor, each traced by the FP64 oracle.  Run
from the repository root: python tests/golden/make_golden_extras.py

Each case names its extra by a key that tests/golden/extras.py rebuilds
(world, sky texture or tensor source); the scene is JSON, the outputs one
.npz, as for make_golden.py."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import Oracle  # noqa: E402
from tests.golden.extras import CASES, applied  # noqa: E402


def main():
    ora = Oracle(threads=1)
    manifest = {}
    for name, (scene, extra) in CASES().items():
        with applied(ora, extra):
            o = ora.render(scene, rgb=True)
        keys = ["cls", "r", "theta", "phi", "intensity", "attempts", "accepted", "rgb"] + (["color"] if "color" in o else [])
        np.savez_compressed(os.path.join(HERE, f"x_{name}.npz"), **{k: o[k] for k in keys})
        manifest[name] = {"scene": scene.to_dict(), "extra": extra}
        print(name, np.bincount(o["cls"], minlength=5))
    with open(os.path.join(HERE, "extras_cases.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
