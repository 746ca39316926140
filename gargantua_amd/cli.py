"""Command-line renderer with the reference's arguments (crates/gr-renderer/src/main.rs:17-169).

    python -m gargantua_amd.cli --metric kerr --spin 0.9 --disk --scene lensed-spheres --output out.png

Same options and defaults as the reference CLI: metric (schwarzschild | kerr), mass, spin, width, height,
disk, scene (none | lensed-spheres), r-inner, r-outer, scale-height, samples, quiet, hdri, hdri-intensity,
output.  The camera is the CLI's fixed one (r = 30M in the equatorial plane looking at the hole, FOV 60
deg, aspect width / height); the four render drivers are chosen as render_dispatch does (main.rs:153-169).
Every ray is traced on the GPU through libgargantua; the image is written as PNG.
"""
from __future__ import annotations

import argparse
import math
import sys
import time

from . import (AccretionDisk, Camera, HdriSky, Kerr, RenderOptions, Schwarzschild, configs, png, render,
               render_with_disk, render_with_disk_and_scene, render_with_scene)


def parse_args(argv=None) -> argparse.Namespace:
    ap = argparse.ArgumentParser(prog="gargantua_amd.cli", description="General-relativity ray renderer")
    ap.add_argument("--metric", default="schwarzschild")
    ap.add_argument("--mass", type=float, default=1.0)
    ap.add_argument("--spin", type=float, default=0.9)
    ap.add_argument("--width", type=int, default=800)
    ap.add_argument("--height", type=int, default=600)
    ap.add_argument("--disk", action="store_true")
    ap.add_argument("--scene", default="none", help="none (default) or lensed-spheres")
    ap.add_argument("--r-inner", type=float, default=6.0)
    ap.add_argument("--r-outer", type=float, default=20.0)
    ap.add_argument("--scale-height", type=float, default=1.0)
    ap.add_argument("--samples", type=int, default=1, help="per-axis supersampling")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--hdri", default=None, help="equirectangular Radiance .hdr environment map")
    ap.add_argument("--hdri-intensity", type=float, default=1.0)
    ap.add_argument("--output", default="out.png")
    return ap.parse_args(argv)


def run(args: argparse.Namespace):
    """Render as main() does and save the PNG; returns the (height, width, 3) image."""
    camera = Camera(position=(0.0, 30.0, math.pi / 2.0, 0.0), look_at=(-1.0, 0.0, 0.0), up=(0.0, 0.0, 1.0),
                    fov_y_radians=math.radians(60.0), aspect=args.width / args.height)
    t0 = time.perf_counter()
    sky = None
    if args.hdri:
        try:
            sky = HdriSky.from_file(args.hdri).with_intensity(args.hdri_intensity)
        except (OSError, ValueError) as e:
            print(f"failed to load HDRI '{args.hdri}': {e}", file=sys.stderr)
            raise SystemExit(1)
    opts = RenderOptions(samples_per_axis=max(1, args.samples), show_progress=not args.quiet, sky=sky)
    disk = AccretionDisk(r_inner=args.r_inner, r_outer=args.r_outer, scale_height=args.scale_height)
    if args.scene == "none":
        scene = None
    elif args.scene == "lensed-spheres":
        scene = configs.lensed_spheres_world()
    else:
        print(f"Unknown scene '{args.scene}'. Supported: none, lensed-spheres", file=sys.stderr)
        raise SystemExit(1)
    if args.metric == "schwarzschild":
        metric = Schwarzschild(args.mass)
    elif args.metric == "kerr":
        metric = Kerr(args.mass, args.spin)
    else:
        print(f"Unknown metric '{args.metric}'. Supported: schwarzschild, kerr", file=sys.stderr)
        raise SystemExit(1)
    if args.disk and scene is not None:
        img = render_with_disk_and_scene(metric, camera, disk, scene, args.width, args.height, opts)
    elif args.disk:
        img = render_with_disk(metric, camera, disk, args.width, args.height, opts)
    elif scene is not None:
        img = render_with_scene(metric, camera, scene, args.width, args.height, opts)
    else:
        img = render(metric, camera, args.width, args.height, opts)
    png.write(args.output, img)
    tags = (["disk"] if args.disk else []) + ([args.scene] if scene is not None else [])
    suffix = "+" + "+".join(tags) if tags else ""
    print(f"Rendered {args.width}x{args.height} ({args.metric}{suffix}) in {time.perf_counter() - t0:.2f}s "
          f"→ {args.output}")
    return img


def main(argv=None) -> None:
    run(parse_args(argv))


if __name__ == "__main__":
    main()
