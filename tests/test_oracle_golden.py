"""The oracle is checked here against the reference's own outputs
(tests/golden, written from the reference) and against closed forms and
invariants.  These are the pins that make the oracle trustworthy as the
checker of the HIP path."""
from dataclasses import replace

import numpy as np
import pytest

from oracle import astrea_oracle as O
from tests._golden import load, meta, rel_l1

CASES = {k: v for k, v in meta()["cases"].items() if k != "rp1d_fail"}


def scheme_for(m, eig):
    return O.Scheme(dimension=m["dimension"], subgrid=m["subgrid"], solver=m["solver"],
                    timestep=m["timestep"], boundary=m["boundary"], gamma=m["gamma"],
                    cfl=m["cfl"], dx=m["dx"], magnetic_2d=m["magnetic_2d"], eig=eig)


@pytest.mark.parametrize("key", [f"{bc}_{c}" for bc in ("edge", "wrap") for c in range(3)])
def test_mc_limiter_bitwise(key):
    g = load("limiter_mc.npz")
    wS = g[key + "_wS"]
    bc = key.split("_")[0]
    w2 = O.pad(wS, bc, 2)
    w = w2[1:-1]
    wF = O.C7_12 * (wS + w[2:]) - O.C1_12 * (w[:-2] + w2[4:])
    wFp = O.pad(wF, bc, 2)
    wL, wR = O.mc_extrapolant_limiter(wS, bc, np.copy(wFp[1:-3]), np.copy(wFp[2:-2]))
    assert np.array_equal(wL, g[key + "_wL"]) and np.array_equal(wR, g[key + "_wR"])


SCHEME_KEYS = [f"{p}_{sg}_{bc}" for p in ("hd", "mhd") for sg in ("pcm", "plm", "ppm", "weno3", "weno5", "weno7")
               for bc in ("edge", "wrap")]


@pytest.mark.parametrize("key", SCHEME_KEYS)
def test_interface_states_and_fluxes(key):
    g = load("schemes_1d.npz")
    mhd, sg, bc = key.split("_")
    s = O.Scheme(dimension=1, subgrid=sg, boundary=bc, gamma=5 / 3 if mhd == "mhd" else 1.4, dx=1 / 32)
    d = O.reconstruct(g[key + "_grid"], s, 0)
    for mine, ref in [(d["wFs"][0], "_wplus"), (d["wFs"][1], "_wminus"), (d["qFs"][0], "_qplus"),
                      (d["qFs"][1], "_qminus"), (d["fluxFs"][0], "_fplus"), (d["fluxFs"][1], "_fminus")]:
        assert np.array_equal(mine, g[key + ref]), ref
    # closed-form spectral radius against the reference's LAPACK eigenvalues
    lm, _ = O.local_max_eig(d["jac_state"], s, 0)
    ref = g[key + "_localmax"]
    assert np.max(np.abs(lm - ref) / np.abs(ref)) < 1e-14
    lm_lapack, _ = O.local_max_eig(d["jac_state"], replace(s, eig="lapack"), 0)
    for solver in (["llf", "hllc", "os"] if mhd == "hd" else ["llf", "hlld", "os"]):
        F = O.riemann(replace(s, solver=solver), 0, lm_lapack, None, d)
        assert np.array_equal(F, g[key + f"_flux_{solver}"]), solver


@pytest.mark.parametrize("name", sorted(CASES))
def test_trajectory_bitwise_lapack(name):
    """With the reference's LAPACK eigenvalues the oracle reproduces every
    golden trajectory to the last bit (same step count and dt sequence)."""
    m, g = CASES[name], load("steps.npz")
    r = O.run(g[name + "_grid0"], scheme_for(m, "lapack"), m["t_end"], m["checkpoints"], m["max_steps"])
    assert r.steps == m["steps"]
    assert np.array_equal(np.array(r.dts), g[name + "_dts"])
    assert np.array_equal(r.grid, g[name + "_final"])


@pytest.mark.parametrize("name", sorted(CASES))
def test_trajectory_closed_eigen(name):
    """The closed-form spectral radius (what the HIP path computes) changes
    the reference result by round-off only: relative L1 <= 1e-13 per variable,
    except WENO-7 whose nonlinear weights amplify it (<= 1e-10)."""
    m, g = CASES[name], load("steps.npz")
    r = O.run(g[name + "_grid0"], scheme_for(m, "closed"), m["t_end"], m["checkpoints"], m["max_steps"])
    assert r.steps == m["steps"]
    tol = 1e-10 if "weno7" in name else 1e-13
    assert np.max(rel_l1(r.grid, g[name + "_final"])) <= tol


def test_failing_draw_reproduced():
    """A seeded Riemann problem on which the reference's LAPACK call raises
    (NaN Jacobian) raises in the oracle at the same step, after the state has
    gone to negative pressure -- the condition the device failure flag reports."""
    m = meta()["cases"]["rp1d_fail"]
    g = load("steps.npz")
    N = m["N"]
    s = O.Scheme(dimension=1, subgrid="ppm", solver="hllc", boundary="edge", gamma=1.4, cfl=.5, dx=1 / N,
                 eig="lapack")
    r = O.run(g["rp1d_fail_grid0"], s, m["t_end"], 1, int(m["steps_ok"]))
    assert np.min(O.cons_to_prim(r.grid, 1.4)[:, 4]) < 0
    with pytest.raises(np.linalg.LinAlgError):
        O.run(g["rp1d_fail_grid0"], s, m["t_end"], 1, int(m["steps_ok"]) + 1)


def test_2d_flux_level_both_parities():
    g = load("space_2d.npz")
    for subgrid, solver, bc, mhd in [("ppm", "hllc", "wrap", False), ("plm", "hllc", "edge", False),
                                     ("weno5", "hlld", "wrap", True), ("ppm", "llf", "edge", False)]:
        for parity in (0, 1):
            key = f"{subgrid}_{solver}_{bc}_p{parity}"
            s = O.Scheme(dimension=2, subgrid=subgrid, solver=solver, boundary=bc, gamma=5 / 3 if mhd else 1.4,
                         dx=1 / 12, magnetic_2d=mhd, eig="lapack", parity=parity)
            fl = O.evolve_space(g[key + "_grid"], s)
            for ax in (0, 1):
                assert np.array_equal(fl[ax]["flux"], g[key + f"_flux_ax{ax}"]), (key, ax)
                assert fl[ax]["eigmax"] == float(g[key + f"_eigmax_ax{ax}"])
                if mhd:
                    assert np.array_equal(fl[ax]["mag_corner"], g[key + f"_corner_ax{ax}"])
                    assert np.array_equal(fl[ax]["face_avg"], g[key + f"_faceavg_ax{ax}"])


def test_initial_condition_pipeline_sod():
    g = load("initialise.npz")
    for sg in ("plm", "ppm"):
        N = 64
        x = np.linspace(0 - .5 / N, 1 + .5 / N, N + 2)[1:-1]     # constructor.py:13-16
        w = np.zeros((N, 8))
        w[:] = [.125, 0, 0, 0, .1, 0, 0, 0]
        w[x <= .5] = [1, 0, 0, 0, 1, 0, 0, 0]
        s = O.Scheme(dimension=1, subgrid=sg, boundary="edge")
        assert np.array_equal(O.initial_average(w, s), g[f"sod_{sg}_64"])


# ---- closed forms and invariants ------------------------------------------
def test_flux_closed_form_hydro():
    rng = np.random.default_rng(1)
    w = np.zeros((50, 8))
    w[:, 0] = rng.uniform(.1, 3, 50)
    w[:, 1:4] = rng.uniform(-2, 2, (50, 3))
    w[:, 4] = rng.uniform(.1, 3, 50)
    g = 1.4
    for ax in (0, 1, 2):
        f = O.phys_flux(w, g, ax)
        rho, v, p = w[:, 0], w[:, 1:4], w[:, 4]
        E = p / (g - 1) + .5 * rho * np.sum(v * v, 1)
        np.testing.assert_allclose(f[:, 0], rho * v[:, ax], rtol=1e-14)
        np.testing.assert_allclose(f[:, 1 + ax], rho * v[:, ax] ** 2 + p, rtol=1e-14)
        np.testing.assert_allclose(f[:, 4], v[:, ax] * (E + p), rtol=1e-13)


def test_spectral_radius_matches_dense_eigensolver_mhd():
    rng = np.random.default_rng(2)
    w = np.zeros((200, 8))
    w[:, 0] = rng.uniform(.1, 3, 200)
    w[:, 1:4] = rng.uniform(-2, 2, (200, 3))
    w[:, 4] = rng.uniform(.1, 3, 200)
    w[:, 5:8] = rng.uniform(-2, 2, (200, 3))
    for ax in (0, 1):
        ref = np.max(np.abs(np.linalg.eigvals(O.jacobian(w, 5 / 3, ax))), -1)
        got = O.spectral_radius(w, 5 / 3, ax, True)
        np.testing.assert_allclose(got, ref, rtol=1e-10)


@pytest.mark.parametrize("name,phi", [
    ("vanleer", lambda r: (r + abs(r)) / (1 + abs(r))),
    ("koren", lambda r: max(0, min(2 * r, (2 + r) / 3, 2))),
    ("superbee", lambda r: max(0, min(2 * r, 1), min(r, 2))),
    ("vanalbada1", lambda r: (r * r + r) / (r * r + 1)),
    ("ospre", lambda r: 1.5 * (r * r + r) / (r * r + r + 1)),
])
def test_slope_limiters_textbook(name, phi):
    rng = np.random.default_rng(3)
    x = np.cumsum(rng.uniform(-1, 1, 40))
    wp = x[:, None]
    got = O.slope_limiter(wp, name)[:, 0]
    for i in range(1, 39):
        a, b = x[i] - x[i - 1], x[i + 1] - x[i]
        r = a / b if b != 0 else 0.
        assert abs(got[i - 1] - phi(r) * b) <= 1e-12 * (1 + abs(b))


def test_minmod_known_answers():
    wp = np.array([0., 1., 3., 2., 2., 5.])[:, None]
    # differences: a=(1,2,-1,0), b=(2,-1,0,3)
    assert O.slope_limiter(wp, "minmod")[:, 0].tolist() == [1., 0., 0., 0.]


def test_ppm_face_interpolant_exact_on_cubic_averages():
    N = 20
    xe = np.arange(N + 1) / N
    P = lambda x: x ** 4 / 4 - x ** 3 + .5 * x ** 2 + x    # antiderivative of a cubic
    avg = (P(xe[1:]) - P(xe[:-1])) * N
    wS = avg[:, None]
    w2 = O.pad(wS, "edge", 2)
    w = w2[1:-1]
    wF = O.C7_12 * (wS + w[2:]) - O.C1_12 * (w[:-2] + w2[4:])
    f = lambda x: x ** 3 - 3 * x ** 2 + x + 1
    np.testing.assert_allclose(wF[2:-2, 0], f(xe[3:-2]), rtol=1e-12, atol=1e-12)


def test_weno5_linear_weights_on_linear_data():
    """Equal smoothness indicators reduce WENO-5 to its linear 5th-order
    stencil, exact for linear data."""
    x = np.arange(30) * .1
    wS = (2 * x + 1)[:, None]
    wL, wR = O.weno(wS, "edge", 5)
    np.testing.assert_allclose(wR[3:-3, 0], (2 * (x + .05) + 1)[3:-3], rtol=1e-12)
    np.testing.assert_allclose(wL[3:-3, 0], (2 * (x - .05) + 1)[3:-3], rtol=1e-12)


@pytest.mark.parametrize("key", [k for k in SCHEME_KEYS if k.startswith("hd_")])
def test_lax_wendroff_reference_pick(key):
    """flux_lw, the reference's column pick np.unique(chars, axis=-1)[..., 1]
    on LAPACK eigenvalues, gives the reference's Lax-Wendroff fluxes to the
    last bit.  flux_lw_speed, which the sound-speed form shares, is pinned
    through it."""
    g = load("schemes_1d.npz")
    _, sg, bc = key.split("_")
    s = O.Scheme(dimension=1, subgrid=sg, solver="lw", boundary=bc, gamma=1.4, dx=1 / 32, eig="lapack")
    d = O.reconstruct(g[key + "_grid"], s, 0)
    lm, chars = O.local_max_eig(d["jac_state"], s, 0)
    assert np.array_equal(O.riemann(s, 0, lm, chars, d), g[key + "_flux_lw"])


def test_lax_wendroff_speed_form_reduces_to_llf():
    """With the wave speed set to max|lambda| the coefficient ss^2/max|lambda|
    is max|lambda| itself, and the Lax-Wendroff form is the LLF flux: checks
    the (k, k+1) pairing and the signs against flux_llf."""
    g = load("schemes_1d.npz")
    s = O.Scheme(dimension=1, subgrid="ppm", solver="lw", boundary="edge", gamma=1.4, dx=1 / 32)
    d = O.reconstruct(g["hd_ppm_edge_grid"], s, 0)
    lm, _ = O.local_max_eig(d["jac_state"], s, 0)
    F_lw = O.flux_lw_speed(lm, lm, d)
    F_llf = O.flux_llf(lm, d)
    np.testing.assert_allclose(F_lw, F_llf, rtol=1e-15, atol=1e-15)
    # the sound speed is below max|lambda| = |u| + c, so LW dissipates less than LLF
    ss = O.lw_sound_speed(d["jac_state"], 1.4)
    assert np.all(ss ** 2 / lm <= lm)
    assert not np.allclose(O.riemann(s, 0, lm, None, d), F_llf)


@pytest.mark.parametrize("solver", ["llf", "hllc", "lw"])
def test_riemann_consistency(solver):
    """F(w, w) = f(w) for a constant state."""
    N = 16
    w = np.zeros((N, 8))
    w[:] = [1.3, .4, 0, 0, .9, 0, 0, 0]
    s = O.Scheme(dimension=1, subgrid="plm", solver=solver, boundary="wrap", dx=1 / N)
    q = O.prim_to_cons(w, 1.4)
    d = O.reconstruct(q, s, 0)
    lm, _ = O.local_max_eig(d["jac_state"], s, 0)
    F = O.riemann(s, 0, lm, None, d)
    np.testing.assert_allclose(F, O.phys_flux(w[:1], 1.4, 0).repeat(N + 1, 0), rtol=1e-14, atol=1e-15)


def test_conservation_periodic_2d():
    rng = np.random.default_rng(4)
    N = 16
    s = O.Scheme(dimension=2, subgrid="ppm", solver="hllc", boundary="wrap", dx=1 / N, cfl=.4)
    c = (np.arange(N) + .5) / N
    xx, yy = np.meshgrid(c, c, indexing="ij")
    w = np.zeros((N, N, 8))
    w[..., 0] = 1 + .3 * np.sin(2 * np.pi * xx) * np.cos(2 * np.pi * yy)
    w[..., 1] = .2 * np.cos(2 * np.pi * yy)
    w[..., 2] = .1 * rng.uniform(-1, 1, (N, N))
    w[..., 4] = 1.5
    q0 = O.initial_average(w, s)
    r = O.run(q0, s, 1., max_steps=5)
    tot0, tot1 = q0.sum((0, 1)), r.grid.sum((0, 1))
    assert np.all(np.abs(tot1 - tot0) <= 1e-13 * np.maximum(1, np.abs(tot0)))


def test_ssprk33_matches_butcher_form():
    """The Shu-Osher stage combinations of SSPRK(3,3) (evolvers.py:165-176)
    equal the textbook Butcher form k1 = L(u), k2 = L(u + dt k1),
    k3 = L(u + dt (k1 + k2)/4), u' = u + dt (k1 + k2 + 4 k3)/6."""
    N = 32
    s = O.Scheme(dimension=1, subgrid="ppm", solver="hllc", boundary="wrap", dx=1 / N, cfl=.5)
    x = (np.arange(N) + .5) / N
    w = np.zeros((N, 8))
    w[:, 0] = 1 + .2 * np.sin(2 * np.pi * x)
    w[:, 1] = .5
    w[:, 4] = 1 + .1 * np.cos(2 * np.pi * x)
    q = O.prim_to_cons(w, 1.4)
    L = lambda u: O.compute_L(O.evolve_space(u, s), s)
    dt = .2 / N
    k1 = L(q)
    k2 = L(q + dt * k1)
    k3 = L(q + dt * (k1 + k2) / 4)
    ref = q + dt * (k1 + k2 + 4 * k3) / 6
    out = O.evolve_time(q, O.evolve_space(q, s), dt, s)
    np.testing.assert_allclose(out, ref, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("case", ["sod", "quad_edge", "quad_wrap", "ot", "shear"])
@pytest.mark.parametrize("subgrid", ["plm", "ppm"])
def test_torch_initial_state_equals_oracle(case, subgrid):
    """astrea_amd.ic's torch pipeline (the initial state of the benchmark
    and serving workloads, and the reference for the device IC kernel) is
    the oracle's constructor restatement to the last bit."""
    import torch
    from astrea_amd import ic
    if case == "sod":
        w, dims, bc, gamma = ic.sod(37), 1, "edge", 1.4
    elif case.startswith("quad"):
        w = ic.quadrants(21, [(1, .2, .1, 1), (.5, -.3, .2, .4), (.2, .1, -.5, .2), (.8, 0, 0, .9)])
        dims, bc, gamma = 2, case[5:], 1.4
    elif case == "ot":
        w, dims, bc, gamma = ic.orszag_tang(16), 2, "wrap", 5 / 3
    else:
        w, dims, bc, gamma = ic.shear_layer(24, 5), 2, "wrap", 1.4
    s = O.Scheme(dimension=dims, subgrid=subgrid, boundary=bc, gamma=gamma)
    ref = O.initial_average(w, s)
    got = ic.initial_average(torch.from_numpy(w), gamma, bc, subgrid == "ppm").numpy()
    assert np.array_equal(got, ref)
