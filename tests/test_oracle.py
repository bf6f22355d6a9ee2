"""Pins the CPU oracle (oracle/lbm_oracle.c) with checks that do not restate its
code: lattice moment identities, exact fixed points, conservation laws, the
quadratic that defines the Smagorinsky relaxation time, the regularized
stress identity, the symmetry-plane / mirrored-domain equivalence, closed-form
momentum exchange, the guard reset and closed-form heat cases."""
import math

import numpy as np
import pytest

from oracle import oracle as orc

F32 = orc.LBMOracle("f32")
F64 = orc.LBMOracle("f64")
CX, CY, CZ, OPP, ZMIR, WT = F64.tables()
C = np.stack([CX, CY, CZ], 1).astype(np.float64)


def test_lattice_identities():
    assert WT.sum() == pytest.approx(1.0, abs=1e-15)
    assert np.abs(WT @ C).max() < 1e-16
    second = np.einsum("i,ia,ib->ab", WT, C, C)
    assert np.abs(second - np.eye(3) / 3).max() < 1e-15
    fourth = np.einsum("i,ia,ib,ic,id->abcd", WT, C, C, C, C)
    d = np.eye(3)
    iso = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d)) / 9
    assert np.abs(fourth - iso).max() < 1e-15
    for i in range(19):
        assert OPP[OPP[i]] == i and (C[OPP[i]] == -C[i]).all()
        assert ZMIR[ZMIR[i]] == i
        assert (C[ZMIR[i]] == C[i] * np.array([1, 1, -1])).all()
    # f32 weights are the WGSL literals
    _, _, _, _, _, wt32 = F32.tables()
    assert wt32[0] == np.float32(0.33333334) and wt32[1] == np.float32(0.055555556) and wt32[7] == np.float32(0.027777778)


@pytest.mark.parametrize("u", [(0.05, 0.0, 0.0), (0.03, -0.02, 0.011), (0.0, 0.0, 0.0)])
def test_feq_moments(u):
    rho = 1.07
    f = np.array([F64.feq(i, rho, *u) for i in range(19)])
    uv = np.array(u)
    assert f.sum() == pytest.approx(rho, abs=1e-15)
    assert np.abs(f @ C - rho * uv).max() < 1e-15
    pi = np.einsum("i,ia,ib->ab", f, C, C)
    assert np.abs(pi - rho * (np.outer(uv, uv) + np.eye(3) / 3)).max() < 1e-15


def empty_case(dims, u=0.05):
    g = orc.grid(dims)
    sim = orc.OSim(u, 0.51, 0.1, 0.0)
    n = g.n
    return g, sim, np.zeros(n, np.uint32), np.zeros(n, np.float32)


@pytest.mark.parametrize("lib,tol", [(F64, 1e-15), (F32, 3e-7)])
def test_uniform_flow_is_fixed_point(lib, tol):
    """Inlet, outlet, moving floor, far field and symmetry plane all return feq at
    uniform flow u_in x: the empty tunnel must stay at its initial state."""
    g, sim, solid, emit = empty_case((12, 7, 6))
    f0, fields0 = lib.init(g, sim)
    f, fields, st = lib.run(g, sim, f0, solid, fields0, emit, 7)
    assert np.abs(f - f0).max() <= tol * np.abs(f0).max()
    assert np.abs(fields[:, 0] - 0.05).max() < 1e-6 and np.abs(fields[:, 3] - 1).max() < 1e-6
    assert st.guard_hits == 0 and all(v == 0 for v in st.forces_q) and all(v == 0 for v in st.forces_d)


def random_state(g, rng, amp=0.01, u=(0.04, 0.0, 0.0)):
    f = np.array([[F64.feq(i, 1.0, *u)] for i in range(19)]) * np.ones((19, g.n))
    f *= 1 + amp * rng.standard_normal((19, g.n))
    fields = np.zeros((g.n, 4))
    fields[:, 0] = u[0]
    fields[:, 3] = 1
    return f, fields


def interior_pull(f, dims, x, y, z):
    nx, ny, nz = dims
    out = np.empty(19)
    for i in range(19):
        sx, sy, sz = x - CX[i], y - CY[i], z - CZ[i]
        out[i] = f[i, sx + nx * (sy + ny * sz)]
    return out


def test_collision_conserves_and_regularizes():
    """Interior cell: mass and momentum of the pulled populations are conserved;
    the outgoing non-equilibrium stress is (1 - 1/tau) times the incoming one, and
    tau solves tau^2 - tau0*tau = (9 sqrt2 / 2) Cs^2 Q / rho (Smagorinsky, Hou et al.)."""
    rng = np.random.default_rng(1)
    dims = (30, 8, 8)
    g, sim, solid, emit = empty_case(dims)
    f, fields = random_state(g, rng, amp=0.05)
    out, _, _ = F64.run(g, sim, f, solid, fields, emit, 1)
    for (x, y, z) in [(12, 3, 4), (20, 5, 2), (26, 4, 5), (3, 2, 3)]:
        fin = interior_pull(f, dims, x, y, z)
        idx = x + dims[0] * (y + dims[1] * z)
        fo = out[:, idx]
        rho = fin.sum()
        m = fin @ C
        assert fo.sum() == pytest.approx(rho, rel=1e-14)
        assert np.abs(fo @ C - m).max() < 1e-14
        u = m / rho
        feq = np.array([F64.feq(i, rho, *u) for i in range(19)])
        pin = np.einsum("i,ia,ib->ab", fin - feq, C, C)
        pout = np.einsum("i,ia,ib->ab", fo - feq, C, C)
        q = math.sqrt((pin * pin).sum())
        nx = dims[0]

        def sst(e0, e1, xx):
            t = min(max((xx - e0) / (e1 - e0), 0.0), 1.0)
            return t * t * (3 - 2 * t)

        tau0 = float(np.float32(0.51)) + 0.4 * sst(nx - 18, nx - 1, x) + 0.25 * (1 - sst(0, 10, x))
        k = (pout * pin).sum() / (pin * pin).sum()  # 1 - 1/tau (negative for tau < 1)
        tau = 1 / (1 - k)
        assert tau * tau - tau0 * tau == pytest.approx(4.5 * math.sqrt(2) * float(np.float32(0.1)) ** 2 * q / rho, rel=1e-8)
        assert np.abs(pout - (1 - 1 / tau) * pin).max() < 1e-14
        # non-stress moments of the non-equilibrium part vanish (regularization)
        ghost = (fo - feq) - np.array([4.5 * WT[i] * np.einsum("ab,ab", np.outer(C[i], C[i]) - np.eye(3) / 3, pout)
                                        for i in range(19)])
        assert np.abs(ghost).max() < 1e-15


def test_symmetry_plane_equals_mirrored_domain():
    """A half tunnel with the z = 0 mirror equals the symmetric full domain inside the
    light cone of the outer z boundaries (populations mirrored with ZMIR)."""
    rng = np.random.default_rng(2)
    K, nx, ny, steps = 8, 24, 10, 3
    gh = orc.grid((nx, ny, K))
    gf = orc.grid((nx, ny, 2 * K))
    sim = orc.OSim(0.05, 0.51, 0.1, 0.0)
    fh, fieldsh = random_state(gh, rng, amp=0.05, u=(0.04, 0.01, 0.0))
    solid_h = np.zeros(gh.n, np.uint32)
    box = np.zeros((K, ny, nx), bool)
    box[0:3, 2:5, 8:12] = True
    solid_h[box.reshape(-1)] = 1
    # full domain: upper half z = K + zh, lower half z = K - 1 - zh (mirrored)
    fh3 = fh.reshape(19, K, ny, nx)
    ff3 = np.empty((19, 2 * K, ny, nx))
    ff3[:, K:] = fh3
    ff3[:, :K] = fh3[ZMIR][:, ::-1]
    sf3 = np.zeros((2 * K, ny, nx), np.uint32)
    sf3[K:] = solid_h.reshape(K, ny, nx)
    sf3[:K] = solid_h.reshape(K, ny, nx)[::-1]
    fieldsf = np.zeros((gf.n, 4))
    fieldsf[:, 0] = 0.04
    fieldsf[:, 3] = 1
    fieldsh[:, 0] = 0.04
    oh, _, _ = F64.run(gh, sim, fh, solid_h, fieldsh, np.zeros(gh.n, np.float32), steps)
    of, _, _ = F64.run(gf, sim, ff3.reshape(19, -1), sf3.reshape(-1), fieldsf, np.zeros(gf.n, np.float32), steps)
    oh3 = oh.reshape(19, K, ny, nx)
    of3 = of.reshape(19, 2 * K, ny, nx)
    zmax = K - 1 - steps
    assert np.abs(of3[:, K:K + zmax] - oh3[:, :zmax]).max() < 1e-15
    assert np.abs(of3[:, K - zmax:K][ZMIR][:, ::-1] - oh3[:, :zmax]).max() < 1e-15


@pytest.mark.parametrize("lib,tol", [(F64, 1e-14), (F32, 2e-6)])
def test_single_cell_momentum_exchange(lib, tol):
    """An isolated solid cell in uniform flow u receives F = 2u (sum over the 18 links
    of -2(f_opp - w) c with f = feq(1, u)), in the group of the solid value."""
    g, sim, solid, emit = empty_case((16, 9, 9), u=0.05)
    idx = 8 + 16 * (4 + 9 * 4)
    solid[idx] = 2  # front-wing group
    f0, fields0 = lib.init(g, sim)
    _, _, st = lib.run(g, sim, f0, solid, fields0, emit, 1)
    d = np.array(st.forces_d)
    u = float(np.float32(0.05))  # the Sim uniform is f32
    assert d[3] == pytest.approx(2 * u, rel=tol)
    assert abs(d[4]) < 1e-7 and abs(d[5]) < 1e-7
    assert np.abs(d[[0, 1, 2, 6, 7, 8]]).max() == 0.0
    if lib is F32:
        assert st.forces_q[3] == pytest.approx(1000, abs=9)  # 18 links rounded at 1e-4 each
        assert st.forces_q[2] == 0 and st.forces_q[5] == 0  # z is never written


def test_mirrored_and_duct_links_carry_no_force():
    g, sim, solid, emit = empty_case((16, 9, 6), u=0.05)
    idx = 8 + 16 * (4 + 9 * 0)  # solid cell on the symmetry plane
    solid[idx] = 1
    f0, fields0 = F64.init(g, sim)
    _, _, st = F64.run(g, sim, f0, solid, fields0, emit, 1)
    d_half = np.array(st.forces_d)
    # links that reach the solid through the mirror (c_z = +1 at z = 0) are skipped:
    # only the 13 links with c_z in {0, +1} count, giving F = 6 u sum(w c_x^2) = 5u/3
    u = float(np.float32(0.05))
    assert d_half[0] == pytest.approx(5 * u / 3, rel=1e-13)
    assert abs(d_half[1]) < 1e-15 and abs(d_half[2]) < 1e-15
    emit2 = emit.copy()
    emit2[:] = -1.0
    emit2[idx] = 0.0
    _, _, st2 = F64.run(g, sim, f0, solid, fields0, emit2, 1)
    assert np.abs(np.array(st2.forces_d)).max() == 0.0


@pytest.mark.parametrize("inject", ["nan", "rho5", "fast"])
def test_guard_resets_cell(inject):
    g, sim, solid, emit = empty_case((14, 8, 8), u=0.06)
    f0, fields0 = F32.init(g, sim)
    x, y, z = 7, 4, 4
    idx = x + 14 * (y + 8 * z)
    f = f0.copy()
    if inject == "nan":
        f[0, idx] = np.nan
    elif inject == "rho5":
        f[0, idx] = 4.0
    else:  # |u| > 0.35: populations moving +x, pulled by this cell
        for i in range(19):
            if CX[i] == 1:
                src = (x - 1) + 14 * ((y - CY[i]) + 8 * (z - CZ[i]))
                f[i, src] = 0.25
    solid[idx + 2] = 1  # a solid neighbour, so the cell would carry force
    out, fields, st = F32.run(g, sim, f, solid, fields0, emit, 1)
    want = np.array([F32.feq(i, 1.0, np.float32(0.06) * np.float32(0.5), 0.0, 0.0) for i in range(19)], np.float32)
    assert st.guard_hits == 1
    assert np.array_equal(out[:, idx], want)
    assert fields[idx, 0] == np.float32(0.03) and fields[idx, 3] == 1.0


def heat_case(dims, u=(0.0, 0.0, 0.0), rng=None):
    g = orc.grid(dims)
    fields = np.zeros((g.n, 4), np.float32)
    fields[:, :3] = u
    fields[:, 3] = 1
    return g, fields


def interior_mask(dims, m=1):
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ((x >= m) & (x < nx - m) & (y >= m) & (y < ny - m) & (z >= m) & (z < nz - m)).reshape(-1)


def test_heat_uniform_cools():
    dims = (10, 8, 6)
    g, fields = heat_case(dims)
    t = np.full(g.n, 50.0, np.float32)
    hp = orc.OHeat(3.0, 0.004, 0.04, 0.0)
    out = F64.heat(g, hp, fields, np.zeros(g.n, np.float32), t)
    m = interior_mask(dims)
    assert np.abs(out[m] - 50.0 * (1 - float(np.float32(0.004)))).max() < 1e-12  # Heat uniform is f32


def test_heat_integer_shift_is_exact():
    rng = np.random.default_rng(3)
    dims = (12, 7, 6)
    g, fields = heat_case(dims, u=(0.5, 0.0, 0.0))
    t = rng.uniform(0, 100, g.n)
    hp = orc.OHeat(2.0, 0.0, 0.0, 0.0)  # u*dt = 1 cell, no smoothing, no cooling
    out = F64.heat(g, hp, fields, np.zeros(g.n, np.float32), t).reshape(dims[::-1])
    t3 = t.reshape(dims[::-1])
    assert np.array_equal(out[:, :, 1:], t3[:, :, :-1])
    assert np.all(out[:, :, 0] == 0.0)  # x < 0 is cold


def test_heat_solid_copies_emit_and_sources_relax():
    dims = (8, 6, 5)
    g, fields = heat_case(dims)
    emit = np.zeros(g.n, np.float32)
    sidx = 4 + 8 * (3 + 6 * 2)
    fields[sidx] = 0
    emit[sidx] = 300.0
    t = np.zeros(g.n)
    hp = orc.OHeat(1.0, 0.0, 0.0, 0.0)
    out = F64.heat(g, hp, fields, emit, t)
    assert out[sidx] == 300.0
    nb = sidx + 1
    assert out[nb] == pytest.approx(0.35 * 300.0, rel=1e-12)
    assert out[sidx + 2] == 0.0
