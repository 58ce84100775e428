"""The CPU oracle, pinned before it is trusted.

Three kinds of pin, each chosen so a plausible mistake in the restatement fails
one of them:

- the golden vectors the reference itself wrote (tests/golden/, made by
  tests/golden/make_golden.py from the reference's own CPU backend): the oracle
  must reproduce them bit for bit, forward, adjoint, reconstruction with and
  without checkpoints, the bfloat16 strip, superposition and the shot-summed
  misfit gradient;
- the transpose identity <W m A^T p, u> = <W m p, A u> of one adjoint step against
  one forward step in float64 (what the reference's own TransposeTest pins,
  tests/sensitivity_test.py:694-767), which a dropped or misplaced tap breaks;
- central finite differences of the forward cost against the adjoint gradient
  in float64, outside the layer, which any wrong factor breaks.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from helpers import golden, golden_names, oracle_sim

from oracle import driver as D


def test_stencil_rows_match_the_reference():
    z = golden("stencils")
    for R in range(1, 5):
        for kind in ("taylor", "optimized"):
            np.testing.assert_array_equal(D.cell_coefficients(R, kind), z[f"cell_{kind}_{R}"])
            np.testing.assert_array_equal(D.staggered_weights(R, kind), z[f"stag_{kind}_{R}"])
    for order in (2, 4, 6, 8):
        for kind in ("taylor", "optimized"):
            assert D.stable_dt((1e-3, 2e-3), 1500.0, order, kind) == float(z[f"stable_dt_{kind}_{order}"])


def test_stencil_rows_are_consistent():
    # consistency of the flux: sum (2k - 1) c_k = 1 for every row (the symbol's slope)
    for R in range(1, 5):
        for kind in ("taylor", "optimized"):
            c = D.cell_coefficients(R, kind)
            assert abs(float(c @ (2 * np.arange(1, R + 1) - 1.0)) - 1.0) < 1e-12


@pytest.mark.parametrize("name", golden_names("fwd_") + ["disk_298x298_o8_forward"])
def test_forward_matches_the_reference_bit_for_bit(name):
    z = golden(name)
    sim = oracle_sim(z)
    field, traces = D.simulate(sim, z["src_pos"], z["src_sig"], z["indicator"], z["sensors"])
    np.testing.assert_array_equal(field, z["field"])
    np.testing.assert_array_equal(traces, z["traces"])


def _objective(z, sim):
    if "observed" in z:
        return D.l2_misfit(z["observed"])
    return D.l2_misfit(np.zeros((sim.N, z["sensors"].shape[1]), sim.dtype))


@pytest.mark.parametrize("name", golden_names("adj_"))
def test_adjoints_match_the_reference_bit_for_bit(name):
    z = golden(name)
    sim = oracle_sim(z)
    args = (sim, z["src_pos"], z["src_sig"], z["indicator"], z["sensors"], _objective(z, sim))
    cost, g, traces, _ = D.sensitivity(*args)
    assert cost == float(z["cost"])
    np.testing.assert_array_equal(g["mass"], z["g_mass"])
    np.testing.assert_array_equal(g["stiff"], z["g_stiff"])
    np.testing.assert_array_equal(traces, z["traces"])
    _, g, _, info = D.reconstruction_sensitivity(*args)
    np.testing.assert_array_equal(g["mass"], z["r_g_mass"])
    np.testing.assert_array_equal(g["stiff"], z["r_g_stiff"])
    assert info["drift"] == float(z["r_drift"])
    if "c_g_mass" in z:
        _, g, _, info = D.reconstruction_sensitivity(*args, checkpoints=2)
        np.testing.assert_array_equal(g["mass"], z["c_g_mass"])
        np.testing.assert_array_equal(g["stiff"], z["c_g_stiff"])
        assert info["segment"] == int(z["c_segment"])
        assert info["sweeps"] == float(z["c_sweeps"])
    if "b_g_mass" in z:
        _, g, _, _ = D.reconstruction_sensitivity(*args, strip_dtype="bfloat16")
        np.testing.assert_array_equal(g["mass"], z["b_g_mass"])
    if "s_g_mass" in z:
        _, g, _, info = D.superposition_sensitivity(*args, scale=float(z["s_scale"]))
        np.testing.assert_array_equal(g["mass"], z["s_g_mass"])
        np.testing.assert_array_equal(g["stiff"], z["s_g_stiff"])
        assert info["cancellation"] == float(z["s_cancellation"])


@pytest.mark.parametrize("tag", ["f64", "f32"])
def test_misfit_gradient_matches_the_reference(tag):
    z = golden(f"misfit_acoustic_o4_{tag}")
    sim = oracle_sim(z)
    sensors = D.Sensors(sim, z["rec_coords"])
    np.testing.assert_array_equal(sensors.nodes, z["sensor_nodes"])
    np.testing.assert_array_equal(sensors.weights, z["sensor_weights"])
    sources = list(zip(z["src_pos"], z["src_sig"]))
    for adjoint, key in ((D.sensitivity, ""), (D.reconstruction_sensitivity, "r_")):
        cost, gradient = D.misfit_gradient(sim, sources, z["indicator"], sensors,
                                           list(z["observed"]), adjoint=adjoint)
        assert cost == float(z[key + "cost"])
        np.testing.assert_array_equal(gradient, z[key + "gradient"])


# ----------------------------------------------------------- mathematical pins
def _box(ndim, order, faces, kind):
    Nx = {1: (41,), 2: (37, 29), 3: (27, 23, 21)}[ndim]
    dx = (0.1,) * ndim
    boundary = {
        "neumann": ((0, 0),) * ndim,
        "dirichlet": ((1, 1),) * ndim,
        "mixed": tuple(((1, 0) if d % 2 == 0 else (0, 1)) for d in range(ndim)),
    }[faces]
    dt = 0.3 * D.stable_dt(dx, 1.0, order)
    if kind == "scalar":
        return D.Sim("ScalarWave", Nx, dx, 1, dt, "float64", order, boundary=boundary,
                     wavespeed=1.0, density=1.0)
    return D.Sim("AcousticWave", Nx, dx, 1, dt, "float64", order, boundary=boundary,
                 materials=(1.0, 3.0, 1.0, 5.0))


def _transpose_error(sim):
    rng = np.random.default_rng(0)
    indicator = 0.2 + rng.random(sim.Nx_padded)
    mat = sim.build_materials(indicator)
    st = D.Steps(sim, mat)
    mask = np.zeros(sim.Nx_padded)
    mask[tuple(slice(1, n - 1) for n in sim.Nx)] = 1.0
    u, p = (rng.standard_normal(sim.Nx_padded) * mask for _ in "up")
    st.refill(u)
    st.refill(p)
    zero = np.zeros_like(u)
    Au = st.fd(zero, u, np.zeros_like(u)) - 2 * u
    Ap = np.zeros_like(p)
    st.fn("adjoint_step")(ctypes.byref(st.g), D._ptr(zero), D._ptr(p), D._ptr(Ap),
                          D._ptr(mat["stiff"]), D._ptr(st.minv), ctypes.c_int(sim.derive_inertia),
                          None, ctypes.c_double(sim.dt), st.f, ctypes.c_int(st.dirichlet))
    Ap = Ap - 2 * p
    minv = mat["minv"] if mat["minv"] is not None else 1.0 / mat["stiff"]
    Wm = sim.apply_cell_weights(mask.copy()) / minv
    scale = float(np.sqrt(np.sum(Wm * Ap**2) * np.sum(Wm * u**2)))
    return abs(float(np.sum(Wm * Ap * u) - np.sum(Wm * p * Au))) / scale


@pytest.mark.parametrize("ndim,order", [(1, 2), (1, 4), (2, 2), (2, 4), (2, 6), (2, 8), (3, 2), (3, 4)])
@pytest.mark.parametrize("faces", ["neumann", "dirichlet", "mixed"])
@pytest.mark.parametrize("kind", ["scalar", "acoustic"])
def test_one_adjoint_step_is_the_transpose_of_one_forward_step(ndim, order, faces, kind):
    assert _transpose_error(_box(ndim, order, faces, kind)) < 1e-12


def test_a_planted_tap_error_breaks_the_transpose():
    # the identity is a pin: one wrong weight in the adjoint's table must fail it
    sim = _box(2, 8, "mixed", "acoustic")
    lib = D.library()
    rng = np.random.default_rng(0)
    indicator = 0.2 + rng.random(sim.Nx_padded)
    mat = sim.build_materials(indicator)
    st = D.Steps(sim, mat)
    mask = np.zeros(sim.Nx_padded)
    mask[tuple(slice(1, n - 1) for n in sim.Nx)] = 1.0
    u, p = (rng.standard_normal(sim.Nx_padded) * mask for _ in "up")
    zero = np.zeros_like(u)
    Au = st.fd(zero, u, np.zeros_like(u)) - 2 * u
    planted = sim.grid()
    planted.op[4 * 3 + 2] *= 1.001  # radius 4, weight 3
    Ap = np.zeros_like(p)
    lib.cwo_adjoint_step_f64(ctypes.byref(planted), D._ptr(zero), D._ptr(p), D._ptr(Ap),
                             D._ptr(mat["stiff"]), D._ptr(st.minv), ctypes.c_int(0), None,
                             ctypes.c_double(sim.dt), st.f, ctypes.c_int(st.dirichlet))
    Ap -= 2 * p
    Wm = sim.apply_cell_weights(mask.copy()) / mat["minv"]
    scale = float(np.sqrt(np.sum(Wm * Ap**2) * np.sum(Wm * u**2)))
    assert abs(float(np.sum(Wm * Ap * u) - np.sum(Wm * p * Au))) / scale > 1e-6


def _pml_case(order):
    z = golden("adj_pml_o8_dirichlet_f64" if order == 8 else "adj_pml_o4_mixed_f64")
    sim = oracle_sim(z)
    sim.N = 90
    return z, sim


@pytest.mark.parametrize("order", [4, 8])
def test_the_adjoint_gradient_matches_finite_differences(order):
    z, sim = _pml_case(order)
    obs = D.simulate(sim, z["src_pos"], z["src_sig"][: sim.N], z["indicator"], z["sensors"])[1] * 0.7
    objective = D.l2_misfit(obs)
    indicator = np.array(z["indicator"])
    _, grads, _, _ = D.sensitivity(sim, z["src_pos"], z["src_sig"][: sim.N], indicator,
                                   z["sensors"], objective)
    jm, js = sim.parametrization_jacobian()
    gradient = jm * grads["mass"] + js * grads["stiff"]
    for node in [(15, 15), (20, 12), (12, 22), (25, 18)]:
        h = 1e-6
        base = indicator[node]

        def cost():
            return objective(D.simulate(sim, z["src_pos"], z["src_sig"][: sim.N], indicator,
                                        z["sensors"])[1])[0]

        indicator[node] = base + h
        plus = cost()
        indicator[node] = base - h
        minus = cost()
        indicator[node] = base
        fd = (plus - minus) / (2 * h)
        assert abs(gradient[node] - fd) <= 1e-6 * max(abs(fd), 1e-300), (node, gradient[node], fd)


def test_the_profiles_vanish_at_the_front_and_mirror_across_the_wall():
    # the layer's closed form (cuwave/boundary.py:327-361): damped on the wall node,
    # undamped from the front node (width - 1 in) on, even about the wall node
    z = golden("adj_pml_o4_mixed_f64")
    sim = oracle_sim(z)
    width = sim.pml.width
    for d, ((lo, hi), prof) in enumerate(zip(sim.pml_widths(), sim.pml_profiles())):
        if prof is None:
            continue
        a_node, b_node = prof[0], prof[1]
        n = sim.Nx[d]
        for lined, wall, front, ghost, mirror in ((lo, 1, width, 0, 2),
                                                  (hi, n - 2, n - 1 - width, n - 1, n - 3)):
            if not lined:
                continue
            assert a_node[wall] < 0.0 and 0.0 < b_node[wall] < 1.0
            assert a_node[front] == 0.0 and b_node[front] == 1.0
            assert a_node[ghost] == a_node[mirror] and b_node[ghost] == b_node[mirror]


def test_the_illumination_after_one_step_is_the_source_injection_squared():
    # from rest one step leaves only the injected points: u = w s(0) there
    z = golden("fwd_pml_acoustic_o8_f32")
    sim = oracle_sim(z)
    sim.N = 1
    illum = D.illumination(sim, [(z["src_pos"], z["src_sig"])], z["indicator"])
    mat = sim.build_materials(z["indicator"])
    lin = sim.flatten(z["src_pos"])
    w = sim.excitation_weights(mat, lin)
    expect = np.zeros(sim.Nx_padded)
    for k, node in enumerate(lin):
        v = sim.dtype(w[k]) * sim.dtype(z["src_sig"][0, k])
        expect.flat[node] = float(v) ** 2
    np.testing.assert_array_equal(illum, expect)
    # two shots add
    two = D.illumination(sim, [(z["src_pos"], z["src_sig"])] * 2, z["indicator"])
    np.testing.assert_array_equal(two, 2 * expect)


def test_the_illumination_accumulates_every_step():
    # after k steps it is the sum over t < k of the squared final field of a
    # t + 1 step run: an overwrite (the last step only) or a skipped step fails
    z = golden("fwd_pml_acoustic_o8_f32")
    sim = oracle_sim(z)
    shots = [(z["src_pos"], z["src_sig"])]
    inner = tuple(slice(1, n - 1) for n in sim.Nx)
    expect = np.zeros(sim.Nx_padded)
    for k in range(1, 7):
        sim.N = k
        field = D.simulate(sim, z["src_pos"], z["src_sig"], z["indicator"])
        expect[inner] += field[inner].astype(np.float64) ** 2  # the run's logical grid
        got = D.illumination(sim, shots, z["indicator"])
        np.testing.assert_allclose(got, expect, rtol=1e-12, atol=0)
    assert np.count_nonzero(expect) > np.count_nonzero(expect == 0) // 100


# ------------------------------------------------- independent bulk modulus and density
def _bulk_sim(kind, ndim=2, N=70, order=4, precision="float64", **kw):
    shape = (40, 36) if ndim == 2 else (22, 21, 20)
    return D.Sim(kind, shape, (1.0,) * ndim, N, 0.25, precision, order,
                 pml=D.Pml(6, 1.5, 1e-6), **kw)


def _bulk_points(sim):
    mid = [n // 2 for n in sim.Nx]
    src = np.array([[m] for m in mid])
    src[0, 0] = 9
    sensors = np.array([[10, 11, 12], [10, 17, 24]] if sim.ndim == 2
                       else [[10, 11, 12], [9, 10, 11], [8, 10, 12]])
    t = np.arange(sim.N) * sim.dt
    sig = np.exp(-((t - 3.0) ** 2))[:, None]
    return src, sig, sensors


def test_bulk_density_reduces_to_the_acoustic_two_phase_medium():
    # kappa and rho following one binary indicator: kappa in {1, 4}, rho in {1, 2},
    # every reciprocal exact, so both parametrizations hand the kernels the same
    # material; the traces then agree bit for bit and the chained gradients agree
    # through d kappa / d gamma = 0.75 kappa^2 and d rho / d gamma = 0.5 rho^2
    acoustic = _bulk_sim("AcousticWave", materials=(1.0, 2.0, 1.0, 4.0))
    bulk = _bulk_sim("BulkDensityWave")
    gamma = np.zeros(acoustic.Nx_padded)
    a0, a1 = np.meshgrid(*(np.arange(n) for n in acoustic.Nx_padded), indexing="ij")
    gamma[(a0 - 20) ** 2 + (a1 - 18) ** 2 <= 36] = 1.0
    stack = np.stack([np.where(gamma > 0, 4.0, 1.0), np.where(gamma > 0, 2.0, 1.0)])
    src, sig, sensors = _bulk_points(acoustic)
    _, tr_a = D.simulate(acoustic, src, sig, gamma, sensors)
    _, tr_b = D.simulate(bulk, src, sig, stack, sensors)
    np.testing.assert_array_equal(tr_a, tr_b)
    objective = D.l2_misfit(0.7 * tr_a)
    _, ga, _, _ = D.reconstruction_sensitivity(acoustic, src, sig, gamma, sensors, objective)
    _, gb, _, _ = D.reconstruction_sensitivity(bulk, src, sig, stack, sensors, objective)
    jm, js = acoustic.parametrization_jacobian()
    dgamma = jm * ga["mass"] + js * ga["stiff"]
    km, ks = bulk.parametrization_jacobian(stack)
    kappa, rho = stack
    chained = km * gb["mass"] * (0.75 * kappa**2) + ks * gb["stiff"] * (0.5 * rho**2)
    scale = np.abs(dgamma).max()
    assert scale > 0
    assert np.abs(chained - dgamma).max() <= 1e-12 * scale


@pytest.mark.parametrize("ndim", [2, 3])
def test_bulk_density_gradients_match_finite_differences(ndim):
    from scipy.ndimage import gaussian_filter

    sim = _bulk_sim("BulkDensityWave", ndim=ndim, N=60 if ndim == 2 else 40)
    rng = np.random.default_rng(7)
    fields = [1.0 + 0.3 * np.tanh(gaussian_filter(rng.standard_normal(sim.Nx_padded), 3) * 8)
              for _ in range(2)]
    stack = np.stack(fields)
    src, sig, sensors = _bulk_points(sim)
    _, tr = D.simulate(sim, src, sig, stack, sensors)
    objective = D.l2_misfit(0.8 * tr)
    _, grads, _, _ = D.reconstruction_sensitivity(sim, src, sig, stack, sensors, objective,
                                                  checkpoints=2)
    # the stored-history adjoint agrees (round-off of the reverse march only)
    _, ref, _, _ = D.sensitivity(sim, src, sig, stack, sensors, objective)
    for key in ("mass", "stiff"):
        assert np.abs(grads[key] - ref[key]).max() <= 1e-10 * np.abs(ref[key]).max()
    km, ks = sim.parametrization_jacobian(stack)
    _, valid = D.reconstruction_nodes(sim)
    for which, g in ((0, km * grads["mass"]), (1, ks * grads["stiff"])):
        # a smooth direction over the reversible interior (the gradient is zeroed
        # in the layer): the directional derivative against central differences
        direction = gaussian_filter(rng.standard_normal(sim.Nx_padded), 2) * valid
        h = 1e-5

        def cost(step):
            trial = stack.copy()
            trial[which] += step * direction
            return objective(D.simulate(sim, src, sig, trial, sensors)[1])[0]

        fd = (cost(h) - cost(-h)) / (2 * h)
        adj = float(np.sum(g * direction))
        assert abs(adj - fd) <= 1e-6 * abs(fd), (which, adj, fd)


# ------------------------------------------------------------------ damping
# A damping field d (cuwave/scalar.py:140-160, scalar.cu:455-459): the update
# (2 u1 - u0 (1 - beta) + minv L u1) / (1 + beta), beta = dt minv d / 2, the
# source sharing the 1 / (1 + beta); the adjoint transposing it; a reverse
# march stepping reversible damped nodes with the damping negated and
# recording the rest (cuwave/sensitivity.py:139-195, 592-598).
def _damped_sim(beta=0.03, N=90, order=4, sponge=False, pml=None, strong=None):
    sim = D.Sim("AcousticWave", (34, 31), (1.0, 1.0), N, 0.3, "float64", order,
                materials=(1.0, 2.5, 1.0, 4.0), pml=pml)
    rng = np.random.default_rng(4)
    indicator = 0.2 + 0.6 * rng.random(sim.Nx_padded)
    minv = sim.inverse_inertia(indicator)
    if sponge:
        # a quadratic ramp over the six nodes behind every face
        taper = np.zeros(sim.Nx_padded)
        for d in range(2):
            idx = np.arange(sim.Nx_padded[d]).reshape([-1 if k == d else 1 for k in range(2)])
            for wall in (1, sim.Nx[d] - 2):
                taper = np.maximum(taper, np.clip(1.0 - np.abs(idx - wall) / 6.0, 0.0, 1.0))
        field = 2.0 * beta * taper**2 / (sim.dt * minv)
    else:
        field = 2.0 * beta / (sim.dt * minv) * np.ones(sim.Nx_padded)
    if strong is not None:
        field[strong] = 2.0 * 2.0 / (sim.dt * minv[strong])  # beta 2: no way back
    inner = np.zeros(sim.Nx_padded, dtype=bool)
    inner[tuple(slice(1, n - 1) for n in sim.Nx)] = True
    sim.damping = np.where(inner, field, 0.0).astype(sim.dtype)
    return sim, indicator


def test_a_uniformly_damped_closed_box_sums_to_a_damped_particle():
    # with Neumann faces sum_W L u = 0, so S = sum W m u obeys
    # (1 + beta) S+ = 2 S - (1 - beta) S- + dt^2 sum_k W_k s_k: a wrong sign of
    # beta, or a source missing its 1 / (1 + beta), breaks it
    beta = 0.03
    sim, indicator = _damped_sim(beta)
    mat = sim.build_materials(indicator)
    st = D.Steps(sim, mat)
    src = np.array([[12, 20], [9, 17]])
    lin = sim.flatten(src)
    t = np.arange(sim.N) * sim.dt
    sig = np.stack([np.exp(-((t - 3.0) ** 2)), -0.5 * np.exp(-((t - 4.0) ** 2))], axis=1)
    w = sim.excitation_weights(mat, lin)
    W = sim.apply_cell_weights(np.ones(sim.Nx_padded))
    inner = tuple(slice(1, n - 1) for n in sim.Nx)
    m = 1.0 / sim.inverse_inertia(indicator)
    wsrc = W.ravel()[lin]
    U = np.zeros((2, *sim.Nx_padded))
    S = [0.0, 0.0]  # u^-2, u^-1 at rest
    for k in range(sim.N):
        u = st.fd(U[k % 2], U[1 - k % 2], U[k % 2])
        st.excite(u, sig[k], lin, w)
        st.refill(u)
        S.append(float(np.sum((W * m * u)[inner])))
        load = sim.dt**2 * float(np.sum(wsrc * sig[k]))
        lhs = (1 + beta) * S[-1] - 2 * S[-2] + (1 - beta) * S[-3]
        assert abs(lhs - load) <= 1e-10 * max(abs(S[-1]), 1.0), (k, lhs, load)
    steps = np.diff(S)  # no restoring force: the velocity decays, S settles
    assert abs(steps[-1]) < 0.1 * np.abs(steps).max()


def _damped_fd(sim, indicator, adjoint, nodes, **kw):
    src = np.array([[12], [15]])
    t = np.arange(sim.N) * sim.dt
    sig = np.exp(-((t - 3.0) ** 2))[:, None]
    sensors = np.array([[8, 20, 25], [9, 16, 24]])
    obs = D.simulate(sim, src, sig, indicator, sensors)[1] * 0.7
    objective = D.l2_misfit(obs)
    _, grads, _, info = adjoint(sim, src, sig, indicator, sensors, objective, **kw)
    jm, js = sim.parametrization_jacobian()
    gradient = jm * grads["mass"] + js * grads["stiff"]
    for node in nodes:
        h = 1e-6
        base = indicator[node]

        def cost():
            return objective(D.simulate(sim, src, sig, indicator, sensors)[1])[0]

        indicator[node] = base + h
        plus = cost()
        indicator[node] = base - h
        minus = cost()
        indicator[node] = base
        fd = (plus - minus) / (2 * h)
        assert abs(gradient[node] - fd) <= 1e-6 * max(abs(fd), 1e-300), (node, gradient[node], fd)
    return gradient, info


def test_damped_gradients_match_finite_differences():
    # the damping is fixed (not scaled by the perturbed indicator), so the
    # central difference sees only the material the gradient is taken for
    nodes = [(11, 9), (17, 14), (13, 11), (22, 15), (3, 4)]
    sim, indicator = _damped_sim(0.02)
    _damped_fd(sim, indicator, D.sensitivity, nodes)
    sim, indicator = _damped_sim(0.05, sponge=True)
    g_full, _ = _damped_fd(sim, indicator, D.sensitivity, nodes)
    for checkpoints in (None, 2):
        g, info = _damped_fd(sim, indicator, D.reconstruction_sensitivity, nodes[:4],
                             checkpoints=checkpoints)
        valid = info["valid"]
        assert valid[11, 9] and not valid[1, 1] is None
        np.testing.assert_allclose(g[valid], g_full[valid], rtol=1e-7,
                                   atol=1e-10 * np.abs(g_full).max())


def test_damping_too_lossy_to_reverse_goes_to_the_strip():
    # a patch damped at beta = 2 cannot be marched back at all: its
    # nodes drop out of the reversible set, the ones a reversible node reads are
    # recorded, and the rebuilt gradient still matches where it is exact
    patch = (slice(10, 17), slice(10, 17))
    sim, indicator = _damped_sim(0.05, sponge=True, strong=patch)
    strip, valid = D.reconstruction_nodes(sim, indicator)
    assert not valid[patch].any() and strip.shape[1] > 0
    lin = set(sim.flatten(strip).tolist())
    assert sim.flatten(np.array([[16], [13]]))[0] in lin  # the patch's last row, read across
    assert sim.flatten(np.array([[13], [13]]))[0] not in lin  # its centre: read by no one
    src = np.array([[12], [15]])
    t = np.arange(sim.N) * sim.dt
    sig = np.exp(-((t - 3.0) ** 2))[:, None]
    sensors = np.array([[8, 20, 25], [9, 16, 24]])
    objective = D.l2_misfit(D.simulate(sim, src, sig, indicator, sensors)[1] * 0.7)
    _, gs, _, _ = D.sensitivity(sim, src, sig, indicator, sensors, objective)
    _, gr, _, info = D.reconstruction_sensitivity(sim, src, sig, indicator, sensors, objective)
    assert info["strip"] == strip.shape[1] and info["drift"] < 1e-10
    for k in ("mass", "stiff"):
        np.testing.assert_allclose(gr[k][valid], gs[k][valid], rtol=1e-7,
                                   atol=1e-10 * np.abs(gs[k]).max())
