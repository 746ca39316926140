"""Central finite differences of the differentiable thin-disk loss, computed
with the oracle (the checker of every gradient test).

L(theta) = mean_p (I_p(theta) - target_p)^2 over the pixel-centre rays of the
gradient scene.  The render is piecewise smooth in theta: an accept/reject
decision of the step-size controller (or the number of attempts before the
disk crossing) can flip between theta - h and theta + h, and the ray's
intensity then jumps by the size of one step's truncation error.  A central
difference straddling such a flip is not a derivative.  `central_fd`
therefore reports, per parameter, both the plain central difference of L and
the central difference over the rays whose class, attempt count and
sequence of step decisions (accept/reject, err == 0, clamps of h, invalid
stages, pole reflections) are equal at theta - h and theta + h (`smooth`),
together with that ray mask, so a test compares a gradient with the latter
over the same rays.
"""
import numpy as np

FIELDS = ("spin", "inclination", "emissivity_q", "r_inner")


def central_fd(oracle, make_scene, target, h):
    """make_scene() -> a fresh gg_grad_scene at theta; returns a dict with
    'fd' [4] (all rays), 'fd_smooth' [4] (rays without a decision flip),
    'mask' [4, N] (those rays), 'flips' [4] (rays excluded)."""
    base = make_scene()
    n = base.width * base.height
    t = np.asarray(target, dtype=np.float64).reshape(n)
    fd = np.zeros(4)
    fd_s = np.zeros(4)
    mask = np.zeros((4, n), dtype=bool)
    for k, f in enumerate(FIELDS):
        gs = []
        for sgn in (1.0, -1.0):
            g = make_scene()
            setattr(g, f, getattr(base, f) + sgn * h)
            gs.append(g)
        outs = [oracle.grad_render(g) for g in gs]
        lp = (outs[0]["intensity"] - t) ** 2
        lm = (outs[1]["intensity"] - t) ** 2
        same = ((outs[0]["cls"] == outs[1]["cls"]) & (outs[0]["attempts"] == outs[1]["attempts"]) &
                (oracle.grad_decisions(gs[0]) == oracle.grad_decisions(gs[1])))
        d = (lp - lm) / (2.0 * h * n)
        fd[k] = d.sum()
        fd_s[k] = d[same].sum()
        mask[k] = same
    return {"fd": fd, "fd_smooth": fd_s, "mask": mask, "flips": (~mask).sum(axis=1)}


def masked_gradient(dI, intensity, target, mask):
    """2 mean_p r_p dI_p over the rays of mask[k] for each parameter k."""
    n = len(intensity)
    r = np.asarray(intensity) - np.asarray(target)
    return np.array([2.0 * np.sum((r * dI[:, k])[mask[k]]) / n for k in range(4)])
