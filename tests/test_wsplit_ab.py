"""The implicit vertical split (dynamics%use_wsplit, lib/oce_adv_tra_driver.F90:221-231 with
adv_tra_vert_impl, lib/oce_adv_tra_ver.F90:90-244) and the caller's valuesAB time levels (frozen
as the reference's drivers leave it, or FESOM2's AB2 interpolation) in the CPU restatement.

Pinned three ways: bit for bit against fixtures written by the compiled reference
(tests/golden/wsplit_ab.npz, tests/golden/make_golden.py) and against the live reference build
where present; the tridiagonal solve against a dense solve of the implicit upwind system written
from its flux form (not from the reference's coefficients), with the column content it conserves.
"""
import ctypes
import hashlib
from pathlib import Path

import numpy as np
import pytest

import bundle
from fesom2_tracer_dwarf_amd import configs
from fesom2_tracer_dwarf_amd.configs import Scheme

GOLDEN = Path(__file__).resolve().parent / "golden"

import sys  # noqa: E402
sys.path.insert(0, str(GOLDEN))
import make_golden  # noqa: E402

needs_ref = pytest.mark.skipif(not bundle.ref_available(),
                               reason="oracle/_ref not built (reference sources absent)")


def _split_case(ver="QR4C", cap=0.4, seed=41):
    c = configs.bathymetry_case("wsplit", 10, 12, seed=seed, nsteps=3, lev_lo=4,
                                scheme=Scheme("MFCT", ver, "FCT", 0.8, 0.7), w_cfl=1.5)
    return configs.split_vertical_velocity(c, cap)


def test_oracle_equals_reference_fixture_wsplit_ab(oracle_mod):
    fx = np.load(GOLDEN / "wsplit_ab.npz")
    hashes = []
    for name, (case, ab) in make_golden.wsplit_ab_cases().items():
        hashes.append(make_golden.case_hash(case) + str(ab))
        out, _ = oracle_mod.Oracle(case).run(save_steps=[3], ab_mode=ab)
        assert np.array_equal(out[3], fx[name]), name
    assert str(fx["input_sha256"]) == hashlib.sha256("".join(hashes).encode()).hexdigest()


@needs_ref
@pytest.mark.parametrize("ver", ["UPW1", "QR4C", "CDIFF", "PPM"])
def test_oracle_equals_live_reference_wsplit(oracle_mod, tmp_path, ver):
    case = _split_case(ver, cap=0.3, seed=43)
    p = bundle.write_bundle(tmp_path / "ws", case, save_steps=[3], save_work=True)
    bundle.run_ref(p)
    out, _ = oracle_mod.Oracle(case).run(save_steps=[3])
    assert np.array_equal(out[3], bundle.read_values(p, case, 3))


@needs_ref
@pytest.mark.parametrize("ab", [1, 2])
def test_oracle_equals_live_reference_ab_modes(oracle_mod, tmp_path, ab):
    case = configs.bathymetry_case("ab", 9, 11, seed=44, nsteps=4, lev_lo=4,
                                   scheme=Scheme("MFCT", "PPM", "FCT", 0.9, 0.8), w_cfl=0.3)
    p = bundle.write_bundle(tmp_path / "ab", case, save_steps=[4], ab_mode=ab)
    bundle.run_ref(p)
    out, _ = oracle_mod.Oracle(case).run(save_steps=[4], ab_mode=ab)
    assert np.array_equal(out[4], bundle.read_values(p, case, 4))


@needs_ref
@pytest.mark.parametrize("ver", ["QR4C", "PPM"])
def test_oracle_equals_live_reference_ale_thickness(oracle_mod, tmp_path, ver):
    """hnode_new != hnode and areasvol != area: the LO solution (lib/oce_adv_tra_driver.F90:208),
    the flux divergence (:339, :360, :411, :424), the limiter's P+/P- (lib/oce_adv_tra_fct.F90:368-370)
    and the caller's update (src/fesom_analytic.F90:435) use the new thickness and the volume areas."""
    case = make_golden.ale_case(seed=61, ver=ver)
    assert not np.array_equal(case.mesh.hnode_new, case.mesh.hnode)
    p = bundle.write_bundle(tmp_path / "ale", case, save_steps=[3], save_work=True)
    bundle.run_ref(p)
    out, _ = oracle_mod.Oracle(case).run(save_steps=[3])
    assert np.array_equal(out[3], bundle.read_values(p, case, 3))


def _vert_impl(oracle_mod, case, lo):
    """fo_adv_tra_vert_impl on a (N, nl-1) field with the case's w_i."""
    o = oracle_mod.Oracle(case)
    t = np.ascontiguousarray(lo, dtype=np.float64).copy()
    wi = np.ascontiguousarray(case.wi, dtype=np.float64)
    rc = oracle_mod.lib().fo_adv_tra_vert_impl(ctypes.byref(o.mesh), case.dt,
                                               wi.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                               t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0
    return t


def _implicit_upwind_dense(case, T):
    """Implicit first-order upwind advection by the vertical velocity w_i on each column, from
    the flux form: h_k T'_k = h_k T_k + dt/V_k (F_k(T') - F_{k+1}(T')), with the interface flux
    F_1 = -w_1 T_1 A_1 at the surface, F_k = -(max(w_k,0) T_k + min(w_k,0) T_{k-1}) A_k inside,
    and none through the bottom; one dense solve per column."""
    m = case.mesh
    out = T.copy()
    for n in range(m.nod2D):
        L = int(m.nlevels_nod2D[n]) - 1
        w, A, V = case.wi[n], m.area[n], m.areasvol[n]
        h = m.hnode_new[n]
        M = np.diag(h[:L].copy())
        for k in range(L):  # layer k (0-based) between interfaces k (top) and k+1 (bottom)
            c = case.dt / V[k]
            # + c * F_top(T')
            if k == 0:
                M[k, k] += c * w[0] * A[0]
            else:
                M[k, k] += c * max(w[k], 0.0) * A[k]
                M[k, k - 1] += c * min(w[k], 0.0) * A[k]
            # - c * F_bottom(T')
            if k + 1 < L:
                M[k, k + 1] -= c * max(w[k + 1], 0.0) * A[k + 1]
                M[k, k] -= c * min(w[k + 1], 0.0) * A[k + 1]
        out[n, :L] = np.linalg.solve(M, h[:L] * T[n, :L])
    return out


def test_vert_impl_equals_dense_implicit_upwind_and_conserves(oracle_mod):
    case = _split_case(cap=0.2, seed=45)
    m = case.mesh
    rng = np.random.default_rng(3)
    lo = 10.0 + rng.standard_normal((m.nod2D, m.nl - 1))
    got = _vert_impl(oracle_mod, case, lo)
    ref = _implicit_upwind_dense(case, lo)
    wet = np.arange(m.nl - 1)[None, :] < (m.nlevels_nod2D[:, None] - 1)
    assert np.max(np.abs(got - ref)[wet]) <= 1e-12 * np.max(np.abs(ref))
    assert np.array_equal(got[~wet], lo[~wet])  # dry layers untouched
    assert np.abs(case.wi).max() > 0 and np.abs(got - lo)[wet].max() > 1e-3  # the solve acts
    # surface w_i = 0: the column content sum(V h T) is conserved
    vol = m.areasvol[:, : m.nl - 1] * m.hnode_new
    before, after = (vol * lo)[wet].sum(), (vol * got)[wet].sum()
    assert abs(after - before) <= 1e-13 * abs(before)


def test_wsplit_with_zero_implicit_part_equals_unsplit(oracle_mod):
    """we = w, wi = 0: the implicit solve is the identity and the step is the unsplit one."""
    c = configs.bathymetry_case("ws0", 10, 12, seed=46, nsteps=2, lev_lo=4,
                                scheme=Scheme("MFCT", "QR4C", "FCT", 0.8, 0.7), w_cfl=0.5)
    split = configs.split_vertical_velocity(c, 1.0)
    assert split.use_wsplit and not np.any(split.wi)
    a, _ = oracle_mod.Oracle(c).run(save_steps=[2])
    b, _ = oracle_mod.Oracle(split).run(save_steps=[2])
    assert np.array_equal(a[2], b[2])


def test_ab2_first_step_and_recurrence(oracle_mod):
    """AB2 with valuesold = values at the start: step 1 advects (1.5+e)T - (0.5+e)T; afterwards
    valuesAB extrapolates linearly from the last two time levels."""
    case = configs.bathymetry_case("ab2", 9, 11, seed=47, nsteps=2, lev_lo=4, w_cfl=0.3)
    o = oracle_mod.Oracle(case)
    e = 0.1
    v = [np.ascontiguousarray(case.tracers[i]).copy() for i in range(case.ntracers)]
    old = [x.copy() for x in v]  # valuesold starts at the values
    for s in range(2):
        for i in range(case.ntracers):
            ab = -(0.5 + e) * old[i] + (1.5 + e) * v[i]
            old[i] = v[i].copy()
            o.step_ab(v[i], ab, case.schemes[i], case.dt)
    out, _ = oracle_mod.Oracle(case).run(save_steps=[2], ab_mode=2, ab_epsilon=e)
    assert np.array_equal(out[2], np.stack(v))
    frozen, _ = oracle_mod.Oracle(case).run(save_steps=[2], ab_mode=1)
    plain, _ = oracle_mod.Oracle(case).run(save_steps=[2])
    assert not np.array_equal(out[2], plain[2]) and not np.array_equal(frozen[2], plain[2])


def test_ab2_interpolation_pinned_by_exact_cases(oracle_mod):
    """The AB2 time level is pinned by facts the formula must satisfy, not by restating it:
    * a field linear in time, T^n = a + b n, gives valuesAB = T at time n + 1/2 + epsilon (the
      linear extrapolation), for several epsilon;
    * epsilon = 0.5 makes the coefficients -1 and 2, exact in binary, so valuesold = values
      returns values bit for bit, and the first AB2 step equals the plain step bit for bit;
    * epsilon = 0 with a field constant in time returns it (coefficients summing to 1).
    A swapped pair of coefficients, a wrong sign or a missing epsilon fails one of these."""
    ab2 = oracle_mod.ab2_interpolate
    rng = np.random.default_rng(3)
    a, b = rng.uniform(-5, 5, (50, 7)), rng.uniform(-1, 1, (50, 7))
    n = 12.0
    for eps in (0.0, 0.1, 0.25, 0.5):
        got = ab2(a + b * (n - 1.0), a + b * n, eps)
        want = a + b * (n + 0.5 + eps)
        assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want)), eps
    t = rng.uniform(0.0, 30.0, (50, 7))
    assert np.array_equal(ab2(t, t, 0.5), t)
    assert np.max(np.abs(ab2(t, t, 0.0) - t)) <= 2e-15 * np.max(t)
    case = configs.bathymetry_case("ab2x", 9, 11, seed=48, nsteps=1, lev_lo=4, w_cfl=0.3)
    two, _ = oracle_mod.Oracle(case).run(save_steps=[1], ab_mode=2, ab_epsilon=0.5)
    plain, _ = oracle_mod.Oracle(case).run(save_steps=[1])
    assert np.array_equal(two[1], plain[1])
