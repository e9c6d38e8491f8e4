"""The CPU oracle against what the mathematics fixes, independent of the
reference: kernel normalisation and derivative, the 3x3 inverse, brute-force
neighbour sets, momentum conservation, rigid-motion and linear-field
exactness of the strain rate, the radial return landing on the yield
surface, the Gruneisen sound speed at rest and the CFL timestep."""
import ctypes
import os

import numpy as np
import pytest

import cases
from mcmtop_amd import deck
from oracle import Oracle, lib

DATA = os.path.join(os.path.dirname(__file__), "golden", "data")
PI = np.pi


def _w(xi, xj, h):
    a = (ctypes.c_double * 3)(*xi)
    b = (ctypes.c_double * 3)(*xj)
    return lib().orc_kernel_w(a, b, h, PI)


def _gw(xi, xj, h):
    a = (ctypes.c_double * 3)(*xi)
    b = (ctypes.c_double * 3)(*xj)
    out = (ctypes.c_double * 3)()
    lib().orc_kernel_gradw(a, b, h, PI, out)
    return np.array(out[:])


def test_kernel_value_at_origin_and_support():
    h = 4.8e-4
    assert _w((0, 0, 0), (0, 0, 0), h) == pytest.approx(1.0 / (PI * h ** 3), rel=1e-15)
    assert _w((0, 0, 0), (2.0 * h, 0, 0), h) == 0.0
    assert _w((0, 0, 0), (2.5 * h, 0, 0), h) == 0.0
    # continuity at z = 1
    a = _w((0, 0, 0), (h * (1 - 1e-12), 0, 0), h)
    b = _w((0, 0, 0), (h * (1 + 1e-12), 0, 0), h)
    assert a == pytest.approx(b, rel=1e-9)


def test_kernel_integrates_to_one():
    h = 1.3
    r = np.linspace(0.0, 2.0 * h, 20001)
    w = np.array([_w((0, 0, 0), (x, 0, 0), h) for x in r])
    integral = np.trapezoid(4.0 * PI * r ** 2 * w, r)
    assert integral == pytest.approx(1.0, abs=1e-7)


def test_gradient_is_the_derivative_of_the_kernel():
    rng = np.random.default_rng(3)
    h = 0.7
    for _ in range(200):
        xi = rng.uniform(-1.3, 1.3, 3)
        xj = rng.uniform(-0.1, 0.1, 3)
        r = np.linalg.norm(xi - xj)
        if abs(r / h - 1.0) < 1e-3 or r / h > 1.999:
            continue
        g = _gw(xi, xj, h)
        eps = 1e-6
        fd = np.array([(_w(xi + eps * e, xj, h) - _w(xi - eps * e, xj, h)) / (2 * eps) for e in np.eye(3)])
        np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9 * np.abs(fd).max())
        # antisymmetry in the pair
        np.testing.assert_array_equal(g, -_gw(xj, xi, h))


def test_inverse3_against_numpy_and_singular_fallback():
    rng = np.random.default_rng(5)
    for _ in range(50):
        m = rng.normal(size=(3, 3)) + 3 * np.eye(3)
        src = (ctypes.c_double * 9)(*m.ravel(order="F"))
        out = (ctypes.c_double * 9)()
        lib().orc_inverse3(src, out)
        inv = np.array(out[:]).reshape(3, 3, order="F")
        np.testing.assert_allclose(inv, np.linalg.inv(m), rtol=1e-12, atol=1e-14)
    sing = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [0.0, 1.0, 1.0]])
    src = (ctypes.c_double * 9)(*sing.ravel(order="F"))
    out = (ctypes.c_double * 9)()
    lib().orc_inverse3(src, out)
    np.testing.assert_array_equal(np.array(out[:]).reshape(3, 3), np.eye(3))


def _cell_index(d, x, k):
    return np.trunc(d.gridsize[k] * (x - d.gridmin[k])).astype(np.int64) + 1


@pytest.mark.parametrize("name", ["plain_f3dm3", "two_materials_contact", "quarter_f3dm1_xyz_sym_tcrit2"])
def test_neighbour_sets_equal_brute_force(name):
    """Real neighbours: every pair inside the sort domain with
    ((dx^2 + dy^2) + dz^2) < (h_i + h_j)^2.  Ghost neighbours of a single
    z-plane deck: mirror images of the particles in the second cell layer."""
    o = Oracle(os.path.join(DATA, name + ".bin"))
    o.initial()
    o.run(5)
    o.stage("neighbours")
    d = o.state(5, 1)
    x, h = d.f("x"), d.f("h")
    inside = np.ones(d.np, bool)
    for k in range(3):
        b = _cell_index(d, x[:, k], k)
        inside &= (b >= 1) & (b <= d.gridlim[k])
    dd = x[:, None, :] - x[None, :, :]
    dist = (dd[..., 0] ** 2 + dd[..., 1] ** 2) + dd[..., 2] ** 2
    hs = (h[:, None] + h[None, :]) ** 2
    near = (dist < hs) & inside[None, :] & inside[:, None]
    np.fill_diagonal(near, False)
    for i in range(d.np):
        want = set(np.nonzero(near[i])[0] + 1)
        assert set(d.neighbours(i)) == want, i
    if name == "plain_f3dm3":
        layer = _cell_index(d, x[:, 2], 2)
        src = np.nonzero(layer == 2)[0]
        gx = x[src] * np.array([1.0, 1.0, -1.0])
        assert sorted(d.g_par - 1) == sorted(src)
        for i in range(d.np):
            if not inside[i]:
                continue
            dg = gx - x[i]
            dist = (dg[:, 0] ** 2 + dg[:, 1] ** 2) + dg[:, 2] ** 2
            want = set(src[dist < (h[src] + h[i]) ** 2])
            got = set(d.g_par[d.ghost_neighbours(i) - 1] - 1)
            assert got == want, i


@pytest.mark.parametrize("name", ["two_materials_contact", "gas_f3dm10_eos13"])
def test_plain_momentum_is_conserved(name):
    """Sum_i m_i a_i = 0 for the symmetric plain-SPH pair sum without planes."""
    o = Oracle(os.path.join(DATA, name + ".bin"))
    o.initial()
    o.run(3)
    for st in ("neighbours", "strain", "update_h", "constitutive", "dt_update", "momentum"):
        o.stage(st)
    d = o.state(4, 6)
    m = d.f("mass")[:, None]
    a = d.f("a")
    total = (m * a).sum(axis=0)
    scale = (m * np.abs(a)).sum(axis=0).max()
    assert np.abs(total).max() <= 1e-12 * scale


def _block_deck(v_of_x, disctype, jitter=True, n=6):
    dx = 4.0e-4
    pts = deck.block(dx, n, n, n)
    if jitter:
        pts = pts + np.random.default_rng(2).uniform(-0.01 * dx, 0.01 * dx, pts.shape)
    m = deck.material(1, 7850.0, 7850.0 * dx ** 3 * len(pts), 1.2 * dx, deck.f3dm1(200e9, 0.3))
    d = deck.make_deck(pts, np.zeros(3), [m], disctype=disctype)
    d.v = v_of_x(d.x)   # exact values: the binary deck carries them as they are
    return d


@pytest.mark.parametrize("disctype", [0, 1])
def test_rigid_translation_gives_zero_strain_rate(tmp_path, disctype):
    d = _block_deck(lambda x: np.tile([12.0, -7.0, 3.0], (len(x), 1)), disctype)
    p = str(tmp_path / "t.bin")
    d.write_bin(p)
    o = Oracle(p)
    o.initial()
    o.stage("neighbours")
    o.stage("strain")
    s = o.state(1, 2)
    tol = 0.0 if disctype == 0 else 1e-8
    assert np.abs(s.f("rod")).max() <= tol
    assert np.abs(s.f("spin")).max() <= tol


def test_corrected_strain_is_exact_for_a_linear_field(tmp_path):
    """The corrected gradient contracts B^-1 so that sum V (x_j - x_i) (x) g =
    B B^-T (mcm_corr_strain.f90:93,126): a linear field is reproduced exactly
    where B is symmetric, which on an unjittered cubic lattice holds for every
    particle whose support is complete (cubic symmetry makes B a multiple of I).
    Near the surface B is not symmetric and the reference's gradient is off by
    B B^-T -- checked here to be a visible deviation, as a guard that the
    transposition is kept."""
    G = np.array([[100.0, 30.0, -20.0], [-10.0, 50.0, 40.0], [25.0, -35.0, -80.0]])
    d = _block_deck(lambda x: x @ G.T, 1, jitter=False, n=8)
    d.tssf = 1.0e-5   # tiny dt: the half-step shear of the lattice stays below 1e-10
    p = str(tmp_path / "l.bin")
    d.write_bin(p)
    o = Oracle(p)
    o.initial()
    o.stage("neighbours")
    o.stage("strain")
    s = o.state(1, 2)
    dt = s.dt
    # the strain rate is taken at x - v dt/2 = (I - dt/2 G) x
    L = G @ np.linalg.inv(np.eye(3) - 0.5 * dt * G)
    rod = s.f("rod").reshape(-1, 3, 3).transpose(0, 2, 1)    # stored column-major
    spin = s.f("spin").reshape(-1, 3, 3).transpose(0, 2, 1)
    idx = np.round(s.f("x") / 4.0e-4 - 0.5)
    interior = np.all((idx >= 2) & (idx <= 5), axis=1)
    assert interior.sum() == 64
    scale = np.abs(L).max()
    np.testing.assert_allclose(rod[interior], np.broadcast_to(0.5 * (L + L.T), rod[interior].shape), rtol=0,
                               atol=1e-8 * scale)
    np.testing.assert_allclose(spin[interior], np.broadcast_to(0.5 * (L - L.T), spin[interior].shape), rtol=0,
                               atol=1e-8 * scale)
    # surface particles: the B B^-T factor shows
    off = np.abs(rod[~interior] - 0.5 * (L + L.T)).max()
    assert off > 1e-4 * scale


def test_radial_return_lands_on_the_yield_surface():
    """f3dm3 with beta = 0: after a plastic step sqrt(3 J2(sigma + p I - alpha)) = sigma_y."""
    o = Oracle(os.path.join(DATA, "plain_f3dm3.bin"))
    o.initial()
    o.run(9)
    before = o.state(9, 10).f("efps").copy()
    for st in ("neighbours", "strain", "update_h", "constitutive"):
        o.stage(st)
    d = o.state(10, 4)
    yielded = d.f("efps") > before
    assert yielded.sum() > 5
    sig = d.f("sigma").reshape(-1, 3, 3)
    alfa = d.f("alfa").reshape(-1, 3, 3)
    xi = sig - alfa + d.f("p")[:, None, None] * np.eye(3)
    xi = xi - np.trace(xi, axis1=1, axis2=2)[:, None, None] / 3 * np.eye(3)
    seq = np.sqrt(1.5 * np.einsum("nij,nij->n", xi, xi))
    np.testing.assert_allclose(seq[yielded], 400e6, rtol=1e-9)


def test_gruneisen_sound_speed_at_rest_and_cfl_timestep():
    """eos4 at mu = 0, e = 0: c^2 = 4G/(3 rho0) + C0^2; tcrit_opt 1: dt = tssf min h/(c+|v|)."""
    o = Oracle(os.path.join(DATA, "corr_f3dm10_eos4_hopt1.bin"))
    o.initial()
    d = o.state()
    dk = deck.Deck.read_bin(os.path.join(DATA, "corr_f3dm10_eos4_hopt1.bin"))
    m = dk.materials[0]
    G, C0, rho0 = m.card[0], m.eoscard[0], m.rho
    np.testing.assert_allclose(d.f("c"), np.sqrt(4.0 * G / (3.0 * rho0) + C0 ** 2), rtol=1e-14)
    crit = d.f("h") / (d.f("c") + d.f("vabs"))
    assert d.dt == pytest.approx(dk.tssf * crit.min(), rel=1e-15)


def test_cap_model_expf_equals_the_c_library_on_every_float():
    """The cap model's single-precision exp on the host (cap_expf,
    mcmtop_amd/csrc/mcm_cap.h, which the GPU path evaluates) against the C
    library's expf, which the reference's flang build and the oracle call:
    equal bits (or both NaN) for all 2^32 float bit patterns."""
    path = os.path.join(os.path.dirname(__file__), "gpu_checks", "libexpf_check.so")
    if not os.path.exists(path):
        pytest.fail("tests/gpu_checks/libexpf_check.so not built (__graft_entry__.build())")
    chk = ctypes.CDLL(path)
    chk.expf_check_host.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_ulonglong)]
    out = (ctypes.c_ulonglong * 2)()
    assert chk.expf_check_host(0, 1 << 32, min(8, os.cpu_count() or 1), out) == 0
    assert out[0] == 0, "%d mismatches, first pattern 0x%08x" % (out[0], out[1])
