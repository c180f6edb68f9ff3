"""Pins of the FP64 CPU oracle (oracle/mpm_oracle.cpp) against facts it does
not restate: closed forms, identities, invariants, library routines and
independent brute-force formulations.  The reference ships no tests or golden
vectors for this path, so these are the pins.
"""
import json
import math
import os

import numpy as np
import pytest

import _oracle as O
from icempm_multi_amd.params import SimParams, FIELD_NAMES, STATUS_CRUSHED, STATUS_DISABLED

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
rng = np.random.default_rng(12345)


def p1_params():
    """Reset + the keys of the reference's build/p1.json + derived constants."""
    cfg = json.load(open(os.path.join(GOLDEN, "configs", "build_p1.json")))
    P = O.reset_params()
    for k, v in cfg.items():
        if k in ("InputRawPoints",):
            continue
        if k == "DP_phi":
            P.DP_tan_phi = math.tan(v * math.pi / 180)
        elif k == "GridX":
            P.GridXTotal = v
        else:
            setattr(P, k, v)
    return O.derive(P)


def test_struct_layout_matches_c():
    import ctypes
    offs = O.simparams_offsets()
    assert len(offs) == len(FIELD_NAMES) + 1
    for name, off in zip(FIELD_NAMES, offs):
        assert getattr(SimParams, name).offset == off, name
    assert ctypes.sizeof(SimParams) == offs[-1]


def test_params_reset_defaults_and_identities():
    P = O.reset_params()
    # parameters_sim.cpp:5-60 spot values
    assert P.n_indenter_subdivisions == 100 and P.GridXTotal == 128 and P.GridY == 55
    assert P.InitialTimeStep == 3e-5 and P.Density == 916 and P.GridHaloSize == 5
    # elastic identities (independent of the formulas used): E = mu(3lambda+2mu)/(lambda+mu),
    # bulk modulus kappa = E/(3(1-2nu)), nu = lambda/(2(lambda+mu))
    E, nu = P.YoungsModulus, P.PoissonsRatio
    assert math.isclose(P.mu * (3 * P.lambda_ + 2 * P.mu) / (P.lambda_ + P.mu), E, rel_tol=1e-14)
    assert math.isclose(P.kappa, E / (3 * (1 - 2 * nu)), rel_tol=1e-14)
    assert math.isclose(P.lambda_ / (2 * (P.lambda_ + P.mu)), nu, rel_tol=1e-14)
    assert math.isclose(P.DP_tan_phi, 1 / math.sqrt(3), rel_tol=1e-15)


def test_params_p1_hand_values():
    P = p1_params()
    h = 3.3 / 256
    assert P.cellsize == h and P.cellsize_inv == 1 / h
    assert P.UpdateEveryNthStep == 500          # 1/(200*1e-5)
    assert math.isclose(P.vmax, 0.5 * h / 1e-5, rel_tol=1e-15)
    assert math.isclose(P.Dp_inv, 4 / h ** 2, rel_tol=1e-15)
    assert math.isclose(P.IndRSq, 0.324 ** 2 / 4, rel_tol=1e-15)
    assert P.gbOffset == 110 * 8
    assert math.isclose(P.NACC_beta, 0.1)
    assert math.isclose(P.DP_tan_phi, math.tan(math.radians(70)), rel_tol=1e-15)


@pytest.mark.parametrize("fx", list(np.linspace(0.5, 1.4999, 23)) + list(rng.uniform(0.5, 1.5, 20)))
def test_weights_moments(fx):
    w = O.weights(fx)
    d = np.arange(3) - fx
    assert abs(w.sum() - 1) < 1e-15
    assert abs((w * d).sum()) < 1e-15
    assert abs((w * d * d).sum() - 0.25) < 1e-15    # justifies Dp_inv = 4/h^2
    # independent piecewise form of the quadratic B-spline N(r)
    def N(r):
        r = abs(r)
        return 0.75 - r * r if r < 0.5 else (0.5 * (1.5 - r) ** 2 if r < 1.5 else 0.0)
    assert np.allclose(w, [N(x) for x in d], rtol=0, atol=1e-15)


def test_kirchhoff_closed_forms():
    P = O.reset_params()
    assert np.all(O.kirchhoff(P, np.eye(2)) == 0)
    for s in (0.9, 0.999, 1.001, 1.2):
        t = O.kirchhoff(P, s * np.eye(2)).reshape(2, 2)
        assert np.allclose(t, 0.5 * P.kappa * (s ** 4 - 1) * np.eye(2), rtol=1e-13, atol=1e-6)
    for g in (1e-4, 0.01, 0.3):
        t = O.kirchhoff(P, np.array([[1, g], [0, 1]])).reshape(2, 2)
        expect = P.mu * np.array([[g * g / 2, g], [g, -g * g / 2]])
        assert np.allclose(t, expect, rtol=1e-12, atol=P.mu * 2e-15)   # dev(b) cancels near 1


def test_water_closed_form():
    P = O.reset_params()
    assert np.all(O.water(P, 1.0) == 0)
    t = O.water(P, 0.9)
    assert math.isclose(t[0], P.kappa * (1 - 0.9 ** -5), rel_tol=1e-14) and t[1] == 0 and t[3] == t[0]


def test_pq_closed_forms():
    P = O.reset_params()
    je, p, q = O.compute_pq(P, 1.1 * np.eye(2))
    assert math.isclose(je, 1.21, rel_tol=1e-15)
    assert math.isclose(p, -P.kappa / 2 * (1.21 ** 2 - 1), rel_tol=1e-14) and q == 0
    g = 0.02
    je, p, q = O.compute_pq(P, np.array([[1, g], [0, 1]]))
    assert je == 1 and p == 0
    assert math.isclose(q, math.sqrt(2) * P.mu * math.sqrt(g ** 4 / 2 + 2 * g * g), rel_tol=1e-13)


def test_grain_modulo_offsets():
    P = O.reset_params()
    # (grain+4)%11 == 10 -> var3 == 1 ; == 0 -> var3 = 1 - GV
    assert O.grain_params(P, 6)["qmax"] == P.IceShearStrength
    assert math.isclose(O.grain_params(P, 7)["qmax"], P.IceShearStrength * (1 - P.GrainVariability))
    assert math.isclose(O.grain_params(P, 18)["qmax"], P.IceShearStrength * (1 - P.GrainVariability))
    assert O.grain_params(P, 17)["qmax"] == P.IceShearStrength
    # (grain+3)%30 == 15 -> var2 == 1
    assert O.grain_params(P, 12)["pmin2"] == -P.IceTensileStrength2
    assert O.grain_params(P, 42)["pmin2"] == -P.IceTensileStrength2
    assert math.isclose(O.grain_params(P, 27)["pmin2"], -P.IceTensileStrength2 * (1 - 15 * 0.033 * P.GrainVariability))
    # the grain is the signed low 16 bits: 65535 -> -1 -> C remainder keeps the sign
    assert O.grain_params(P, 65535 - 65536 + 65536)["qmax"] == O.grain_params(P, -1)["qmax"]
    for g in range(30):
        gp = O.grain_params(P, g)
        assert gp["pmax"] == P.IceCompressiveStrength and gp["pmin"] == -P.IceTensileStrength
        assert gp["beta"] == P.NACC_beta


def test_failure_surface_continuity_and_sides():
    P = O.reset_params()
    for g in (0, 6, 13, 29):
        gp = O.grain_params(P, g)
        # the linear (p<0) and elliptic (p>=0) branches meet at p=0 with q = q0
        q0_ellipse = math.sqrt(gp["mSq"] * gp["beta"] * gp["pmax"] ** 2 / (1 + 2 * gp["beta"]))
        assert O.failure_check(P, g, 1e-9, q0_ellipse * (1 + 1e-9)) == 1
        assert O.failure_check(P, g, 1e-9, q0_ellipse * (1 - 1e-9)) == 0
        assert O.failure_check(P, g, -1e-9, q0_ellipse * (1 + 1e-8)) == 1
        assert O.failure_check(P, g, -1e-9, q0_ellipse * (1 - 1e-8)) == 0
        # tension cut-off at pmin2, ellipse apex at pmax
        assert O.failure_check(P, g, gp["pmin2"] * 1.0001, 0.0) == 1
        assert O.failure_check(P, g, gp["pmin2"] * 0.9999, 0.0) == 0
        assert O.failure_check(P, g, gp["pmax"] * 1.0001, 0.0) == 1
        assert O.failure_check(P, g, gp["pmax"] * 0.9999, 0.0) == 0


def rot(a):
    return np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])


SVD_CASES = [np.eye(2), np.diag([2.0, 0.5]), np.diag([0.5, 2.0]), rot(0.3), 1.3 * rot(-1.1), np.zeros((2, 2)),
             np.array([[1.0, 0.2], [0.0, -0.8]]), np.array([[0.0, 1.0], [1.0, 0.0]])] + \
            [rng.normal(size=(2, 2)) for _ in range(40)] + [np.eye(2) + 1e-3 * rng.normal(size=(2, 2)) for _ in range(40)]


@pytest.mark.parametrize("F", SVD_CASES)
def test_svd_invariants(F):
    U, S, V = O.svd(F)
    assert np.allclose(U @ np.diag(S) @ V.T, F, atol=1e-13)
    assert np.allclose(U.T @ U, np.eye(2), atol=1e-14) and np.allclose(V.T @ V, np.eye(2), atol=1e-14)
    assert abs(np.linalg.det(U) - 1) < 1e-14 and abs(np.linalg.det(V) - 1) < 1e-14
    assert S[0] >= S[1] and abs(S[0]) >= abs(S[1]) - 1e-15
    ref = np.linalg.svd(F, compute_uv=False)
    assert np.allclose(np.sort(np.abs(S)), np.sort(ref), atol=1e-13)
    assert math.copysign(1, S[1]) == math.copysign(1, np.linalg.det(F)) or abs(np.linalg.det(F)) < 1e-300 or S[1] == 0


def test_polar_rotation_closest():
    for F in SVD_CASES[8:20]:
        R, S = O.polar(F)
        assert np.allclose(S, S.T, atol=1e-14)
        # the reading of fill<2> as a row-major matrix gives R^T of the rotation G with S = G A
        assert np.allclose(R.T @ S, F, atol=1e-13) or np.allclose(R @ S, F, atol=1e-13)


def plast_params():
    P = O.reset_params()
    P.DP_threshold_p = 1e5
    P.DP_tan_phi = math.tan(math.radians(30))
    return P


def test_return_map_tension_branch():
    P = plast_params()
    F = np.array([[1.02, 0.003], [-0.001, 1.015]])          # J > 1: p < -p_t
    Jp = 1.0
    Fn, Jn, ud, p, q = O.point_plasticity(P, F, Jp, STATUS_CRUSHED)
    Je = np.linalg.det(F)
    Je_new = math.sqrt(1 + 2 * P.DP_threshold_p / P.kappa)
    assert math.isclose(np.linalg.det(Fn), Je_new, rel_tol=1e-13)
    assert math.isclose(-P.kappa / 2 * (np.linalg.det(Fn) ** 2 - 1), -P.DP_threshold_p, rel_tol=1e-8)
    assert math.isclose(Jn, Jp * Je_new / Je, rel_tol=1e-14)
    # the new elastic state is isotropic
    b = Fn @ Fn.T
    assert abs(b[0, 0] - b[1, 1]) < 1e-13 and abs(b[0, 1]) < 1e-13


def test_return_map_shear_branch():
    P = plast_params()
    g = 0.004                                               # simple shear, J = 1 (p = 0), large q
    F = np.array([[1.0, g], [0.0, 1.0]]) @ (0.9995 * np.eye(2))   # small compression: p > 0
    Fn, Jn, ud, p, q = O.point_plasticity(P, F, 1.0, STATUS_CRUSHED)
    Je = np.linalg.det(F)
    q_dp = (p + P.DP_threshold_p) * P.DP_tan_phi            # the Drucker-Prager cone
    assert q > q_dp and Jn == 1.0
    bn, b = Fn @ Fn.T, F @ F.T
    assert math.isclose(np.trace(bn), np.trace(b), rel_tol=1e-14)      # tr(b) kept
    dev = bn - np.trace(bn) / 2 * np.eye(2)
    q_new = math.sqrt(2) * P.mu / Je * np.linalg.norm(dev)             # q measured with the trial Je
    assert math.isclose(q_new, q_dp, rel_tol=1e-9)
    assert np.allclose(bn @ b, b @ bn, atol=1e-13)                     # same principal directions


def test_return_map_cap_and_intact():
    P = plast_params()
    F = 0.8 * rot(0.2)                                     # strong compression: p > pmax -> q* = 0
    je, p, q = O.compute_pq(P, F)
    assert p > P.IceCompressiveStrength
    Fn, _, _, _, _ = O.point_plasticity(P, F @ np.array([[1, 1e-3], [0, 1]]), 1.0, STATUS_CRUSHED)
    b = Fn @ Fn.T
    assert abs(b[0, 0] - b[1, 1]) < 1e-12 and abs(b[0, 1]) < 1e-12   # projected to the hydrostatic axis
    # intact point inside the surface: untouched
    F = np.eye(2) + 1e-6 * np.array([[1, 2], [0, -1]])
    Fn, Jn, ud, _, _ = O.point_plasticity(P, F, 1.0, 0)
    assert ud == 0 and np.array_equal(Fn, F) and Jn == 1.0


# --------------------------------------------------------------- phases

def tiny_params(GX=32, GY=24, h=0.1, halo=4, dt=1e-5, **kw):
    P = O.reset_params()
    P.GridXTotal, P.GridY, P.GridXDimension, P.GridHaloSize = GX, GY, GX * h, halo
    P.InitialTimeStep, P.Density, P.YoungsModulus = dt, 980.0, 9e9
    P.ParticleVolume = (h / 2) ** 2
    P.SetupType, P.IndVelocity = 2, 0.0
    for k, v in kw.items():
        setattr(P, k, v)
    return O.derive(P)


def random_points(P, n, x0, x1, y0, y1, seed=1, crushed_frac=0.0):
    r = np.random.default_rng(seed)
    soa = O.empty_soa(n + 7)
    soa[4, :n] = r.uniform(x0, x1, n)
    soa[5, :n] = r.uniform(y0, y1, n)
    soa[6:8, :n] = r.normal(0, 0.5, (2, n))
    F = np.eye(2)[None] + 1e-4 * r.normal(size=(n, 2, 2))
    soa[8:12, :n] = F.reshape(n, 4).T
    soa[12:16, :n] = r.normal(0, 2.0, (4, n))
    soa[3, :n] = 1.0
    ud = (r.integers(0, 1000, n)).astype(np.int64)
    ud[r.uniform(size=n) < crushed_frac] |= STATUS_CRUSHED
    soa[0, :n] = ud.view(np.float64)
    return soa


def brute_p2g(P, soa, n):
    """APIC P2G from the math: w = N((x_p - x_i)/h) with the piecewise B-spline,
    contribution m w (v + C (x_i - x_p)) - dt V (4/h^2) tau (x_i - x_p) w."""
    h, halo, GY = P.cellsize, P.GridHaloSize, P.GridY
    grid = np.zeros((3, O.n_nodes(P)))

    def N(r):
        r = abs(r)
        return 0.75 - r * r if r < 0.5 else (0.5 * (1.5 - r) ** 2 if r < 1.5 else 0.0)

    for p in range(n):
        xp = soa[4:6, p]
        v = soa[6:8, p]
        F = soa[8:12, p].reshape(2, 2)
        C = soa[12:16, p].reshape(2, 2)
        J = np.linalg.det(F)
        b = F @ F.T
        tau = P.mu / J * (b - np.trace(b) / 2 * np.eye(2)) + P.kappa / 2 * (J * J - 1) * np.eye(2)
        m = P.ParticleMass
        c = np.floor(xp / h).astype(int)
        for gx in range(c[0] - 2, c[0] + 3):
            for gy in range(c[1] - 2, c[1] + 3):
                xi = np.array([gx, gy]) * h
                w = N((xp[0] - xi[0]) / h) * N((xp[1] - xi[1]) / h)
                if w == 0:
                    continue
                d = xi - xp
                mom = m * w * (v + C @ d) - P.InitialTimeStep * P.ParticleVolume * (4 / h ** 2) * w * (tau @ d)
                idx = (gx + halo) * GY + gy
                grid[0, idx] += m * w
                grid[1:, idx] += mom
    return grid


def test_p2g_against_brute_force_and_conservation():
    P = tiny_params()
    n = 300
    soa = random_points(P, n, 0.6, 2.6, 0.5, 1.8)
    grid, err = O.p2g(P, soa, n)
    assert err == 0
    ref = brute_p2g(P, soa, n)
    scale = np.abs(ref).max(axis=1, keepdims=True)
    assert np.all(np.abs(grid - ref) <= 1e-12 * scale)
    m = P.ParticleMass
    assert math.isclose(grid[0].sum(), n * m, rel_tol=1e-13)
    mom = m * soa[6:8, :n].sum(axis=1)
    assert np.allclose(grid[1:].sum(axis=1), mom, rtol=1e-10, atol=1e-12 * np.abs(grid[1:]).sum())


def test_p2g_thread_count_independent():
    P = tiny_params(GX=64, GY=32)
    soa = random_points(P, 2000, 0.6, 5.6, 0.5, 2.5, seed=3)
    O.lib().orc_set_threads(1)
    g1, _ = O.p2g(P, soa, 2000)
    O.lib().orc_set_threads(7)
    g7, _ = O.p2g(P, soa, 2000)
    O.lib().orc_set_threads(0)
    assert np.array_equal(g1, g7)


def test_p2g_error_codes():
    P = tiny_params()
    for (x, y), code in [((-0.35 - 0.1 * P.GridHaloSize, 1.0), 70), ((1.0, -0.08), 71),
                         ((P.GridXDimension + 0.1 * P.GridHaloSize - 0.05, 1.0), 72), ((1.0, P.GridY * 0.1 - 0.1), 73)]:
        soa = random_points(P, 4, 1.0, 1.5, 1.0, 1.5)
        soa[4, 2], soa[5, 2] = x, y
        grid, err = O.p2g(P, soa, 4)
        assert err == code, (x, y, err)
    soa = random_points(P, 4, 1.0, 1.5, 1.0, 1.5)
    soa[0, 1] = np.int64(STATUS_DISABLED).view(np.float64) if False else np.array([STATUS_DISABLED], np.int64).view(np.float64)[0]
    g_all, _ = O.p2g(P, soa, 4)
    assert math.isclose(g_all[0].sum(), 3 * P.ParticleMass, rel_tol=1e-14)


def test_update_free_fall_clamp_walls():
    P = tiny_params()
    nn = O.n_nodes(P)
    grid = np.zeros((3, nn))
    halo, GY = P.GridHaloSize, P.GridY
    def idx(gx, gy):
        return (gx + halo) * GY + gy
    m = 2.0
    # interior node, slow: free fall
    grid[:, idx(10, 10)] = [m, m * 0.3, m * -0.2]
    # interior node, too fast: clamped to vmax along its direction
    grid[:, idx(12, 10)] = [m, m * 3 * P.vmax, m * 4 * P.vmax]
    # walls
    grid[:, idx(10, 2)] = [m, m * 1.0, m * 1.0]              # bottom layer -> 0
    grid[:, idx(10, GY - 3)] = [m, m * 1.0, m * 1.0]         # top: vy>0 -> 0
    grid[:, idx(2, 10)] = [m, m * -1.0, 0.0]                 # left: vx<0 -> 0
    grid[:, idx(P.GridXTotal - 3, 10)] = [m, m * 1.0, 0.0]   # right: vx>0 -> 0
    grid[:, idx(2, 12)] = [m, m * 1.0, 0.0]                  # left, moving away: kept
    f = O.update_nodes(P, grid, 0.0, -10.0, -10.0)
    assert np.all(f == 0)
    assert grid[1, idx(10, 10)] == 0.3 and math.isclose(grid[2, idx(10, 10)], -0.2 - P.dt_Gravity, rel_tol=1e-15)
    v = grid[1:, idx(12, 10)]
    assert math.isclose(np.hypot(*v), P.vmax, rel_tol=1e-14) and math.isclose(v[1] / v[0], (4 * P.vmax - P.dt_Gravity) / (3 * P.vmax), rel_tol=1e-12)
    assert np.all(grid[1:, idx(10, 2)] == 0)
    assert grid[2, idx(10, GY - 3)] == 0 and grid[1, idx(10, GY - 3)] == 1.0
    assert grid[1, idx(2, 10)] == 0
    assert grid[1, idx(P.GridXTotal - 3, 10)] == 0
    assert grid[1, idx(2, 12)] == 1.0


def test_update_indenter_contact():
    P = tiny_params(IndDiameter=0.6, IndVelocity=2.0)
    halo, GY = P.GridHaloSize, P.GridY
    nn = O.n_nodes(P)
    grid = np.zeros((3, nn))
    cx, cy = 1.0, 1.2
    m = 1.5
    nodes = [(11, 12), (9, 12), (10, 14), (12, 13)]
    for gx, gy in nodes:
        grid[:, (gx + halo) * GY + gy] = [m, 0.0, 0.0]
    g0 = grid.copy()
    f = O.update_nodes(P, grid, 0.0, cx, cy)
    total = np.zeros(2)
    for gx, gy in nodes:
        i = (gx + halo) * GY + gy
        nrm = np.array([gx * P.cellsize - cx, gy * P.cellsize - cy])
        nh = nrm / np.linalg.norm(nrm)
        v_free = np.array([0.0, -P.dt_Gravity])
        vrel = v_free - [2.0, 0]
        if vrel @ nh < 0:
            # normal velocity now equals the indenter's
            assert math.isclose(grid[1:, i] @ nh, np.array([2.0, 0]) @ nh, rel_tol=1e-12, abs_tol=1e-12)
            total += (v_free - grid[1:, i]) * m / P.InitialTimeStep
        else:
            assert np.allclose(grid[1:, i], v_free)
    assert np.allclose(f.reshape(-1, 2).sum(axis=0), total, rtol=1e-12)
    assert np.count_nonzero(f) > 0


def grid_from_field(P, fn):
    halo, GY = P.GridHaloSize, P.GridY
    nn = O.n_nodes(P)
    grid = np.zeros((3, nn))
    idx = np.arange(nn)
    gx = idx // GY - halo
    gy = idx % GY
    vx, vy = fn(gx * P.cellsize, gy * P.cellsize)
    grid[0] = 1.0
    grid[1], grid[2] = vx, vy
    return grid


def test_g2p_uniform_and_affine():
    P = tiny_params()
    n = 200
    soa = random_points(P, n, 0.6, 2.6, 0.5, 1.8, seed=7)
    soa[8:12, :n] = np.array([1, 0, 0, 1.0])[:, None]
    x0 = soa[4:6, :n].copy()
    u = np.array([0.7, -0.4])
    grid = grid_from_field(P, lambda x, y: (u[0] + 0 * x, u[1] + 0 * y))
    s = soa.copy()
    assert O.g2p(P, s, n, grid) == 0
    assert np.allclose(s[6:8, :n], u[:, None], atol=1e-15)
    assert np.allclose(s[12:16, :n], 0, atol=1e-11)
    assert np.allclose(s[4:6, :n], x0 + u[:, None] * P.InitialTimeStep, atol=1e-15)
    A = np.array([[0.3, -1.2], [0.8, 0.1]])
    b = np.array([0.05, -0.02])
    grid = grid_from_field(P, lambda x, y: (A[0, 0] * x + A[0, 1] * y + b[0], A[1, 0] * x + A[1, 1] * y + b[1]))
    s = soa.copy()
    assert O.g2p(P, s, n, grid) == 0
    assert np.allclose(s[12:16, :n].T.reshape(n, 2, 2), A[None], atol=1e-11)      # APIC reproduces affine C
    assert np.allclose(s[6:8, :n], A @ x0 + b[:, None], atol=1e-13)
    Fn = s[8:12, :n].T.reshape(n, 2, 2)
    assert np.allclose(Fn, (np.eye(2) + P.InitialTimeStep * A)[None], atol=1e-15)


def test_substeps_free_fall_momentum_and_mass():
    P = tiny_params(GX=48, GY=48)
    n = 400
    soa = random_points(P, n, 1.5, 3.0, 2.0, 3.2, seed=9)
    soa[6:8, :n] = 0
    soa[12:16, :n] = 0
    soa[8:12, :n] = np.array([1, 0, 0, 1.0])[:, None]
    t, step, f, err = O.substeps(P, soa, n, 20, 0.0, 0)
    assert err == 0 and step == 20 and math.isclose(t, 20 * P.InitialTimeStep, rel_tol=1e-12)
    vy = soa[7, :n].mean()
    assert math.isclose(vy, -20 * P.dt_Gravity, rel_tol=1e-9)
    assert abs(soa[6, :n].mean()) < 1e-12


@pytest.mark.parametrize("comp", [0, 1], ids=["P", "S"])
def test_elastic_wave_speeds(comp):
    # P and S fronts run at the speeds the linearised stress fixes,
    # sqrt((mu + kappa)/rho) and sqrt(mu/rho), less MPM's ~1.3% grid dispersion;
    # the three-dimensional P modulus lambda + 2 mu would read 4% fast, a
    # mu/kappa swap 47% fast on S, a factor 4 in the P2G force scale 2x off
    import _waves as W
    P, hs, n = W.block()
    O.sort(P, hs, n)
    probes, nsub = W.start(P, hs, n, comp)
    t, step = 0.0, 0
    times, series = [], [[] for _ in probes]
    for _ in range(nsub // W.CHUNK):
        t, step, _f, err = O.substeps(P, hs, n, W.CHUNK, t, step)
        assert err == 0
        times.append(t)
        for s_, m in zip(series, probes):
            s_.append(hs[6 + comp, :n][m].mean())
    ratio = W.measured_speed(times, series) / W.expected_speed(P, comp)
    assert 0.97 <= ratio <= 1.01, ratio
    # behind the front both probes move at V/2: equal impedance on both sides
    for s_ in series:
        assert abs(s_[-1] - W.V / 2) <= 0.1 * W.V / 2, s_[-1]


def test_cell_key_and_stable_sort():
    P = tiny_params(GX=64, GY=40)
    n = 3000
    r = np.random.default_rng(5)
    soa = O.empty_soa(n + 10)
    soa[4, :n] = r.uniform(0, 6.0, n)
    soa[5, :n] = r.uniform(0, 3.5, n)
    # exact half-cell positions and duplicates (ties)
    soa[4, :50] = (r.integers(1, 60, 50) + 0.5) * P.cellsize
    soa[4, 50:100] = soa[4, :50]
    soa[5, 50:100] = soa[5, :50]
    soa[8, :n] = np.arange(n)                      # payload: original index
    ud = np.zeros(n, np.int64)
    ud[r.uniform(size=n) < 0.05] = STATUS_DISABLED
    soa[0, :n] = ud.view(np.float64)
    keys_bf = np.array([int(soa[4, i] * P.cellsize_inv + 0.5) * P.GridY + int(soa[5, i] * P.cellsize_inv + 0.5)
                        for i in range(n)])
    for i in range(0, n, 97):
        assert O.cell_key(P, soa[4, i], soa[5, i]) == keys_bf[i]
    active = np.nonzero(ud == 0)[0]
    order = active[np.argsort(keys_bf[active], kind="stable")]
    s = soa.copy()
    m, keys, perm = O.sort(P, s, n)
    assert m == len(active)
    assert np.array_equal(perm, order) and np.array_equal(keys, keys_bf[order])
    assert np.array_equal(s[8, :m], order.astype(float))


def test_cell_key_fixed_point_exact():
    from fractions import Fraction
    r = np.random.default_rng(11)
    fb, ox, oy, GY = 19, 10, 2, 1024
    for _ in range(2000):
        xq, yq = int(r.integers(0, 2 ** 32)), int(r.integers(0, 2 ** 32))
        def cell(q, org):
            v = Fraction(q, 2 ** fb) - org + Fraction(1, 2)
            return math.trunc(v)
        assert O.cell_key_fixed(xq, yq, fb, ox, oy, GY) == cell(xq, ox) * GY + cell(yq, oy)
    # exact half positions
    xq = (ox + 5) * 2 ** fb + 2 ** (fb - 1)
    assert O.cell_key_fixed(xq, oy * 2 ** fb, fb, ox, oy, GY) == 6 * GY
