"""Ratio tracking of shadow rays in the oracle (CPU only).

Ratio tracking (Novak, Selle and Jarosz 2014, section 3) replaces delta
tracking's 0/1 outcome by the product of the null-collision probabilities
1 - sigma(x) / majorant at the tentative collisions, an estimate of the
transmittance with the same expectation and less variance.  The reference
has delta tracking only (volume/Volume.h:168-203); here it is an opt-in mode
(BNX_RENDER_RATIO_TRACKING).

Pinned in a volume whose rays see a constant density sigma = 0.75 sigma_bar
below a majorant sigma_bar: every macrocell holds planes of voxels at the
majorant's density, which the rays stay a voxel away from.

* Each estimate is exactly p^N in float32, p = 1 - sigma / sigma_bar = 0.25,
  N the ray's tentative collisions (a mistake in the factor, 1 - sigma_bar /
  sigma or sigma / sigma_bar, fails it).
* The mean estimate equals the product over the ray's macrocell segments of
  E[p^N(L)], N(L) the number of renewals of -fastLog(1 - r) / sigma_bar
  steps (over the 2^24 equally likely r) within the segment, solved from the
  renewal equation; walking without the restart at every macrocell entry,
  or the exponential law exp(-sigma L), are told apart.
* Its variance is below delta tracking's E[T](1 - E[T])."""
import math

import numpy as np
import pytest

from barney_amd import scenes
from tests.test_volume_pins_cpu import NQ, _fastlog_q, _mc_lengths

SIG_BAR = 3.0
FIELD = 0.25            # opacity ramp 0 -> 1 over [0, 1]: density 0.25 sigma_bar on the rays
P_NULL = 1.0 - FIELD


def _volume():
    n = 33
    f = np.full((n, n, n), FIELD, np.float32)
    f[:, :, 0::8] = 1.0                        # x-planes of majorant density in every macrocell
    tf = np.array([[0.9, 0.6, 0.3, 0.0], [0.9, 0.6, 0.3, 1.0]], np.float32)
    return scenes.StructuredVolumeScene(scalars=f, grid_origin=(0.0, 0.0, 0.0), grid_spacing=(1.0 / 32,) * 3, tf=tf,
                                        tf_domain=(0.0, 1.0), base_density=SIG_BAR, cells_per_mc=8)


@pytest.fixture(scope="module")
def qdist():
    return _fastlog_q()


def _renewal_g(q, x_max, dx=1e-3):
    """g(x) = E[p^N(x)], N(x) = renewals of steps q (unit majorant) in
    [0, x): g(x) = P(q >= x) + p * sum_y P(q in dy) g(x - y)"""
    n = int(math.ceil(x_max / dx)) + 2
    edges = np.arange(n + 1) * dx
    cnt = np.histogram(q, bins=edges)[0].astype(np.float64) / NQ     # step mass per bin
    surv = 1.0 - np.concatenate([[0.0], np.cumsum(cnt)])[:n]          # P(q >= x_k)
    g = np.zeros(n)
    for k in range(n):
        # steps landing in bin j < k (midpoint rule on the step's bin)
        conv = np.dot(cnt[:k], g[k - 1::-1][:k]) if k > 0 else 0.0
        g[k] = surv[k] + P_NULL * conv
    return edges[:n], g


def test_ratio_tracking_estimates(orc, qdist):
    s = orc.Scene()
    s.add_volume(_volume())
    s.build()
    dims, maj = s.volume_majorants()
    assert tuple(dims) == (4, 4, 4)
    assert np.allclose(np.asarray(maj).ravel(), SIG_BAR)
    rng = np.random.default_rng(17)
    n = 40000
    ang = rng.uniform(0, 2 * math.pi, n)
    org = np.stack([np.full(n, 0.14), 0.5 + 1.5 * np.cos(ang), 0.5 + 1.5 * np.sin(ang)], 1)
    tgt = np.stack([np.full(n, 0.14), rng.uniform(0.1, 0.9, n), rng.uniform(0.1, 0.9, n)], 1)
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 4:7], rays[:, 7] = org, d, 1e30
    T, N = s.volume_ratio_n(rays)

    # every estimate is p^N in float32
    pw = np.ones(n, np.float32)
    for k in range(int(N.max())):
        pw = np.where(N > k, (pw * np.float32(P_NULL)).astype(np.float32), pw)
    assert np.array_equal(T, pw)
    assert N.max() >= 4 and np.mean(N) > 0.5

    # the mean against the renewal solution over the exact macrocell segments
    xs, g = _renewal_g(qdist, SIG_BAR * 1.6)
    exp_T, exp_norestart, exp_exp = [], [], []
    keep = []
    for i in range(n):
        L = _mc_lengths(rays[i, 0:3], rays[i, 4:7], 0.25, (4, 4, 4))
        if L is None or len(L) == 0:
            continue
        keep.append(i)
        gi = np.interp(SIG_BAR * L, xs, g)
        exp_T.append(np.prod(gi))
        exp_norestart.append(np.interp(SIG_BAR * L.sum(), xs, g))
        exp_exp.append(math.exp(-FIELD * SIG_BAR * L.sum()))
    keep = np.array(keep)
    exp_T = np.array(exp_T)
    assert len(keep) > 0.9 * n
    obs = float(T[keep].mean())
    e = float(exp_T.mean())
    var = float(np.var(T[keep].astype(np.float64)))
    sd = math.sqrt(var / len(keep))
    print(f"ratio tracking: mean T {obs:.5f} expected {e:.5f} (sd {sd:.5f}); without restarts "
          f"{np.mean(exp_norestart):.5f}, exp law {np.mean(exp_exp):.5f}")
    assert abs(obs - e) < 4.5 * sd
    assert abs(np.mean(exp_norestart) - obs) > 8 * sd
    # less variance than delta tracking's Bernoulli outcome of the same mean
    assert var < 0.8 * float(np.mean(exp_T * (1 - exp_T)) + np.var(exp_T))
