"""Plane elastic fronts in an ice block, for the wave-speed pins.

Linearising the Kirchhoff stress of kernels.cuh:24-34 about F = I (with the
two-dimensional deviator dev(A) = A - tr(A)/2 I and kappa = 2 mu/3 + lambda of
parameters_sim.cpp:128-133) gives tau_xx = (mu + kappa) e for uniaxial strain
F = diag(1 + e, 1), and tau_xy = mu g for simple shear F = [[1, g], [0, 1]].
P-waves therefore travel at sqrt((mu + kappa)/rho) and S-waves at
sqrt(mu/rho).  The force P2G applies, dt V (4/h^2) tau (kernels.cuh:88), enters
both speeds, while mass and momentum sums do not see it.

The left part of a block moves at V (along x for a P front, along y for an S
front) and the rest is at rest, so a front carrying particle velocity V/2 runs
into the resting part.  Its speed is taken between two probes on the
centreline, from the times their mean velocity crosses V/4; a difference of
two arrivals cancels the time the front takes to form.  The probes lie closer
to the interface than 0.34 block heights, so the front reaches them before the
P-waves that the moving part sends along the free top and bottom surfaces
(c_P/c_S = 1.78 here).  Grid dispersion makes MPM fronts about 1.3% slower
than the continuum speed on this 4-points-per-cell block.
"""
import math

import numpy as np

from icempm_multi_amd import scenes

H = 0.25            # cell size, m
V = 0.01            # m/s; stresses stay far below the failure surface
L_MOVING = 10.0     # m of the block's left part set in motion
PROBES = (2.0, 8.0)  # m from the interface
CHUNK = 10          # substeps between samples


def block():
    """A 30 x 30 m block (120 x 120 cells, 57,600 points), no gravity, no indenter
    on the grid, lifted 4 cells so that no point reaches the wall nodes."""
    P, hs, n = scenes.block_scene(nx=120, ny=120, h=H, grid=(192, 160), dt=1e-5, Gravity=0.0)
    hs[5, :n] += 4 * H
    return P, hs, n


def expected_speed(P, comp):
    return math.sqrt(((P.mu + P.kappa) if comp == 0 else P.mu) / P.Density)


def start(P, hs, n, comp):
    """Sets the moving part's velocity component comp; returns the probe masks
    (over the first n points, in the current order) and the substeps to run."""
    x, y = hs[4, :n], hs[5, :n]
    xs = x.min() + L_MOVING
    ymid = 0.5 * (y.min() + y.max())
    vel = hs[6 + comp, :n]
    vel[x < xs] = V
    probes = [(np.abs(x - (xs + d)) < 0.5 * H) & (np.abs(y - ymid) < 2 * H) for d in PROBES]
    t_end = 1.25 * PROBES[1] / expected_speed(P, comp)
    nsub = int(math.ceil(t_end / P.InitialTimeStep / CHUNK)) * CHUNK
    return probes, nsub


def arrival(times, values):
    """First time the series reaches V/4, interpolated linearly; None if never."""
    t0, v0 = 0.0, 0.0
    for t, v in zip(times, values):
        if v >= V / 4:
            return t0 + (V / 4 - v0) * (t - t0) / (v - v0)
        t0, v0 = t, v
    return None


def measured_speed(times, series):
    a = [arrival(times, s) for s in series]
    assert None not in a, a
    return (PROBES[1] - PROBES[0]) / (a[1] - a[0])
