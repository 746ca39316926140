"""Writes the golden fixtures: small scenes of every mode traced by the FP64
oracle.  Run from the repository root: python tests/golden/make_golden.py

The reference (Rust) cannot run here, so these vectors come from the oracle;
they pin the oracle against regressions and give the GPU tests fixed
expected outputs.  Each case stores its scene as JSON and the per-ray
outputs in one .npz."""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import gargantua_amd as gg  # noqa: E402
from oracle import Oracle  # noqa: E402


def cases():
    cam = gg.configs.inclination_camera(30.0, math.radians(75.0), math.radians(45.0))
    ref = gg.RK45Integrator(max_radius=200.0, max_steps=5000)
    out = {}
    for mname, metric in (("schw", gg.Schwarzschild(1.0)), ("kerr09", gg.Kerr(1.0, 0.9))):
        out[f"{mname}_none_ref"] = gg.Scene(metric=metric, camera=cam, integrator=ref, width=16, height=16)
        out[f"{mname}_volumetric_ref"] = gg.Scene(metric=metric, camera=cam, disk=gg.AccretionDisk(), integrator=ref,
                                                  width=16, height=16)
    out["schw_thin_rtol"] = gg.configs.schwarzschild_256(16, 16)
    out["kerr09_thin_rtol_jitter"] = gg.configs.kerr_1024(16, 16)
    out["kerr0998_thin_rtol"] = gg.configs.kerr_thin_scene(16, 16, spin=0.998, r_camera=60.0, inclination_deg=85.0,
                                                           fov_deg=30.0, max_radius=300.0)
    sanity = gg.Scene(metric=gg.Schwarzschild(1.0),
                      camera=gg.Camera(position=(0.0, 30.0, math.pi / 2.0, 0.0), look_at=(-1.0, 0.0, 0.0),
                                       up=(0.0, 0.0, 1.0), fov_y_radians=math.radians(60.0), aspect=12 / 8),
                      disk=gg.AccretionDisk(), integrator=gg.RK45Integrator(max_radius=200.0, max_steps=5000),
                      width=12, height=8, samples_per_axis=2)
    out["schw_volumetric_spa2_rgb"] = sanity
    return out


def main():
    ora = Oracle(threads=1)
    manifest = {}
    for name, sc in cases().items():
        o = ora.render(sc, rgb=True)
        arrays = {k: o[k] for k in ("cls", "r", "theta", "phi", "intensity", "attempts", "accepted", "rgb")}
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
        manifest[name] = sc.to_dict()
        print(name, np.bincount(o["cls"], minlength=4))
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
