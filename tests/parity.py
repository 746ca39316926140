"""Comparison helpers for GPU-vs-oracle parity (test infrastructure)."""
from __future__ import annotations

import math

import numpy as np


def state_rel_err(a: dict, b: dict) -> np.ndarray:
    """Per-ray final-state error: max(|dr|/r, |dtheta|, |dphi| wrapped), angles in rad."""
    dr = np.abs(a["r"] - b["r"]) / np.maximum(np.abs(b["r"]), 1e-300)
    dth = np.abs(a["theta"] - b["theta"])
    dph = np.abs(a["phi"] - b["phi"])
    dph = np.minimum(dph, 2 * math.pi - dph)
    return np.maximum(dr, np.maximum(dth, dph))


def boundary_band(cls_img: np.ndarray, width: int = 2) -> np.ndarray:
    """Pixels within `width` px (Chebyshev) of a class change in `cls_img` (2-D)."""
    h, w = cls_img.shape
    edge = np.zeros((h, w), bool)
    edge[:, 1:] |= cls_img[:, 1:] != cls_img[:, :-1]
    edge[:, :-1] |= cls_img[:, 1:] != cls_img[:, :-1]
    edge[1:, :] |= cls_img[1:, :] != cls_img[:-1, :]
    edge[:-1, :] |= cls_img[1:, :] != cls_img[:-1, :]
    band = edge.copy()
    for dy in range(-width, width + 1):
        for dx in range(-width, width + 1):
            sh = np.zeros_like(edge)
            ys = slice(max(dy, 0), h + min(dy, 0))
            yd = slice(max(-dy, 0), h + min(-dy, 0))
            xs = slice(max(dx, 0), w + min(dx, 0))
            xd = slice(max(-dx, 0), w + min(-dx, 0))
            sh[yd, xd] = edge[ys, xs]
            band |= sh
    return band


def psnr(a: np.ndarray, b: np.ndarray) -> float:
    peak = float(max(np.max(np.abs(b)), 1e-300))
    mse = float(np.mean((a - b) ** 2))
    return float("inf") if mse == 0 else 10.0 * math.log10(peak * peak / mse)


def compare(gpu: dict, ref: dict, shape=None, band_px: int = 2) -> dict:
    cls_eq = gpu["cls"] == ref["cls"]
    err = state_rel_err(gpu, ref)
    same = cls_eq
    out = {
        "n": int(len(cls_eq)),
        "class_agree": float(np.mean(cls_eq)),
        "state_bit_identical": float(np.mean((gpu["r"] == ref["r"]) & (gpu["theta"] == ref["theta"])
                                             & (gpu["phi"] == ref["phi"]))),
        "attempts_equal": float(np.mean(gpu["attempts"] == ref["attempts"])),
        "err_le_1e-6": float(np.mean(err[same] <= 1e-6)) if same.any() else 1.0,
        "err_median": float(np.median(err[same])) if same.any() else 0.0,
        "psnr_intensity": psnr(gpu["intensity"], ref["intensity"]),
    }
    if shape is not None:
        band = boundary_band(ref["cls"].reshape(shape), band_px).ravel()
        out["band_frac"] = float(np.mean(band))
        out["class_agree_outside_band"] = float(np.mean(cls_eq[~band])) if (~band).any() else 1.0
        out["class_agree_inside_band"] = float(np.mean(cls_eq[band])) if band.any() else 1.0
        ok = same & ~band
        out["err_le_1e-6_outside_band"] = float(np.mean(err[ok] <= 1e-6)) if ok.any() else 1.0
    return out
