"""Host-side logic of the drop-in API (no GPU): the reference's own
known-answer tests (qspectra tests/test_operator_tools.py,
test_liouville_space.py, test_polarization.py, test_decorators.py,
test_simulate_utils.py, test_utils.py) run against qspectra_amd, plus
agreement with the oracle on randomised inputs."""
from textwrap import dedent

import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_equal

from qspectra_amd import operator_tools as ot
from qspectra_amd import polarization as pol
from qspectra_amd import utils
from qspectra_amd.dynamics import liouville_space as ls
from qspectra_amd.simulate import decorators, utils as su
from qspectra_amd import engine
import qspectra_amd as qa
from oracle import system as S
from oracle import liouville as Lv
from oracle import propagate as P


def test_operator_tools():
    M = np.array([[1., 2 - 2j], [2 + 2j, 3]])
    assert ot.all_states(2) == [[], [0], [1], [0, 1]]
    assert ot.all_states(2, 'ge') == [[], [0], [1]]
    assert_allclose(ot.operator_1_to_2(M), [[4]])
    assert_allclose(ot.operator_1_to_2(np.diag([1, 10, 100])), np.diag([11, 101, 110]))
    assert_allclose(ot.operator_extend(M, 'f'), [[4]])
    assert_allclose(ot.operator_extend(M), [[0, 0, 0, 0], [0, 1, 2 - 2j, 0],
                                            [0, 2 + 2j, 3, 0], [0, 0, 0, 4]])
    assert_allclose(ot.transition_operator(0, 2, 'ge'), [[0, 1, 0], [1, 0, 0], [0, 0, 0]])
    assert_allclose(ot.transition_operator(0, 2, 'gef', ''), np.zeros((4, 4)))
    assert_allclose(ot.transition_operator(0, 2, 'gef', '-'),
                    [[0, 1, 0, 0], [0, 0, 0, 0], [0, 0, 0, 1], [0, 0, 0, 0]])
    assert_allclose(ot.transition_operator(0, 2, 'gef', '+'),
                    [[0, 0, 0, 0], [1, 0, 0, 0], [0, 0, 0, 0], [0, 0, 1, 0]])
    assert_equal(ot.n_excitations(3, 2), [2, 6, 6])
    assert ot.extract_subspace('gg,eg->gg') == ['g', 'e']
    assert ot.full_liouville_subspace('ge,ee,ff') == 'gg,ge,gf,eg,ee,ef,fg,fe,ff'
    assert ot.hilbert_subspace_index('e', 'gef', 3) == slice(1, 4)
    with pytest.raises(ot.SubspaceError):
        ot.hilbert_subspace_index('f', 'ge', 3)
    assert_allclose(ot.vib_create(3) @ ot.vib_annihilate(3), np.diag([0, 1, 2]))
    U = np.array([[1, 1], [-1, 1]]) / np.sqrt(2)
    assert_allclose(ot.basis_transform_vector([1, 0], U), np.sqrt([0.5, 0.5]))
    with pytest.raises(ValueError):
        ot.basis_transform_operator(np.random.randn(4, 4), np.eye(3))
    # randomised agreement with the oracle's literal loops
    rs = np.random.RandomState(0)
    for n in (3, 5, 7):
        A = rs.randn(n, n)
        assert_allclose(ot.operator_1_to_2(A), S.operator_1_to_2(A))
        for k in range(n):
            for tr in ('+', '-', '-+'):
                assert_equal(ot.transition_operator(k, n, 'gef', tr),
                             S.transition_operator(k, n, 'gef', tr))


def test_liouville_index_maps_and_superoperators():
    assert_equal(ls.liouville_subspace_index('eg,ge', 'ge', 2), [1, 2, 3, 6])
    assert_equal(ls.liouville_subspace_index('eg,fe', 'gef', 2), [1, 2, 7, 11])
    assert_equal(ls.liouville_subspace_index('gg,ee,ff', 'gef', 2), [0, 5, 6, 9, 10, 15])
    assert_equal(ls.liouville_subspace_index('gg', 'ge', 1, 2), [0, 1, 4, 5])
    assert_equal(ls.liouville_subspace_index('eg', 'ge', 1, 2), [2, 3, 6, 7])
    with pytest.raises(ot.SubspaceError):
        ls.liouville_subspace_index('ef', 'ge', 2)
    assert ls.all_liouville_subspaces('ge') == 'gg,ge,eg,ee'
    R = np.random.rand(3, 3, 3, 3)
    sup = ls.tensor_to_super(R)
    for i, a, j, b in np.ndindex(3, 3, 3, 3):
        assert sup[i + 3 * a, j + 3 * b] == R[i, a, j, b]
    rho_v = 0.25 * np.array([1., 2 + 2j, 2 - 2j, 3])
    rho_d = 0.25 * np.array([[1., 2 - 2j], [2 + 2j, 3]])
    assert_allclose(ls.matrix_to_ket_vec(rho_d), rho_v)
    assert_allclose(ls.ket_vec_to_matrix(rho_v), rho_d)
    assert_allclose(ls.matrix_to_bra_vec(rho_d), 0.25 * np.array([1, 2 - 2j, 2 + 2j, 3]))
    op = np.random.rand(2, 2) + 1j * np.random.rand(2, 2)
    assert_allclose(ls.matrix_to_ket_vec(op @ rho_d - rho_d @ op),
                    ls.super_commutator_matrix(op) @ rho_v)
    # layout of a multi-block subspace: blocks cover every index once
    layout, idx = ls.block_layout('gg,ge,eg,ee', 'gef', 3)
    covered = np.sort(np.concatenate([sel for _, _, sel, _ in layout]))
    assert_equal(covered, np.arange(idx.size))


class _Model(object):
    def __init__(self, sub, n):
        self.sub, self.n = sub, n

    def liouville_subspace_index(self, s):
        return ls.liouville_subspace_index(s, self.sub, self.n)


def test_liouville_space_operator_known_answers_and_oracle():
    X = np.array([[1, 2], [3, 4]])
    m = _Model('ge', 1)
    L = ls.LiouvilleSpaceOperator(X, 'gg,eg,ge,ee->gg', m)
    v = [1, 10, 100, 1000]
    assert_allclose(L.left_multiply(v), [21])
    assert_allclose(L.right_multiply(v), [301])
    assert_allclose(L.commutator(v), [-280])
    assert_allclose(L.expectation_value(v), [21])
    L = ls.LiouvilleSpaceOperator(X, 'ee->gg,ee', m)
    assert_allclose(L.left_multiply([1]), [0, 4])
    assert_allclose(L.expectation_value([1]), 4)
    L = ls.LiouvilleSpaceOperator(X, 'ee->gg,ge,eg,ee', m)
    assert_allclose(L.expectation_value([1]), 4)
    # restricted products equal the reference's sliced kron construction
    rs = np.random.RandomState(2)
    O = rs.randn(7, 7) + 1j * rs.randn(7, 7)
    m3 = _Model('gef', 3)
    for mp in ('ee->fe', 'gg,ge,eg,ee->ee', 'fe->ee', 'eg->gg'):
        a = ls.LiouvilleSpaceOperator(O, mp, m3)
        b = Lv.LiouvilleSpaceOperator(O, mp, m3)
        assert_allclose(a.commutator.matrix, b.commutator_matrix())
        assert_allclose(a.expectation_value.matrix, b.expectation_row())
        assert_allclose(a.bra_vector, b.bra_vector)


def test_polarization_and_decorators():
    assert_allclose(pol.invariant_weights_4th_order('xxyy'), np.array([4, -1, -1]) / 30.0)
    assert_allclose(pol.invariant_weights_4th_order([np.pi / 4, -np.pi / 4, np.pi / 2, 0]),
                    [0, 1 / 12.0, -1 / 12.0], atol=1e-10)
    with pytest.raises(ValueError):
        pol.polarization_vector('X')
    with pytest.raises(ValueError):
        pol.invariant_polarizations(((0, 1), (0, 2)))
    assert pol.random_rotation_matrix().shape == (3, 3)
    # qspectra tests/test_decorators.py:8-50
    assert decorators._get_call_args(lambda a, **b: None, a=1, c=2) == {'a': 1, 'c': 2}
    with pytest.raises(NotImplementedError):
        decorators._get_call_args(lambda *a: None, 1, 2, 3)
    binary = {'xx': 1, 'yy': 2, 'zz': 4}
    f = decorators.optional_2nd_order_isotropic_average(
        lambda polarization: (0, binary[polarization]))
    assert_allclose(f('xx', exact_isotropic_average=True), (0, 7 / 3.0))
    assert_allclose(f('xy', exact_isotropic_average=True), (0, 0))
    with pytest.raises(ValueError):
        f('xyz', exact_isotropic_average=True)
    f = decorators.optional_4th_order_isotropic_average(
        lambda polarization: (0, binary[polarization[:2]] + 10 * binary[polarization[2:]]))
    ma = pol.MAGIC_ANGLE
    assert_allclose(f([0, 0, ma, ma], exact_isotropic_average=True),
                    (0, (11 + 12 + 14 + 21 + 22 + 24 + 41 + 42 + 44) / 9.0))


def test_utils_and_symmetrize():
    class Example(object):
        def __init__(self, a=0, b=1, c=None, d=None):
            self.a, self.b = a, b
            self.c = self if c is None else c
            self.d = Example(1, 2, 3, 4) if d is None else d

        def __repr__(self):
            return utils.inspect_repr(self)
    assert repr(Example()) == dedent("""
        Example(
            a=0,
            b=1,
            c=Example(...),
            d=Example(
                  a=1,
                  b=2,
                  c=3,
                  d=4))
        """).strip()
    z = utils.ZeroArray()
    z += 10
    assert z == 10
    assert_allclose(su._symmetrize([0, 1, 2], [1, 1, 1]), ([-2, -1, 0, 1, 2], [0, 0, 1, 1, 1]))
    assert_allclose(su._symmetrize([-2, -1], [1, 1]), ([-2, -1, 0, 1, 2], [1, 1, 0, 0, 0]))
    assert_allclose(su._symmetrize([-0.75, 0.25, 1.25], [1, 1, 1]),
                    ([-0.75, 0.25, 1.25], [1, 1, 1]))
    assert su.is_constant([1, 1, 1, 1], positive=True)
    assert not su.is_constant([1, 1, 1, 0])


def test_rk4_schedule_matches_oracle():
    rs = np.random.RandomState(4)
    cases = [(np.arange(0, 10.0, 1.0), None, None), (np.array([0., 100., 200., 500.]), 0.0, 2.0),
             (np.sort(rs.uniform(0, 50, 20)), 0.0, 0.7), (np.arange(0, 254.5, 2.0), None, 0.25)]
    for t, t0, dt in cases:
        n1, h1 = engine.rk4_schedule(t, t0, dt)
        n2, h2 = P.rk4_schedule(t, t0, dt)
        assert_equal(n1, n2)
        assert_allclose(h1, h2, rtol=1e-15)
    n, h = engine.rk4_schedule([0., 100., 200., 500.], 0.0, 2.0)
    assert_equal(n, [0, 50, 50, 150])


def test_pulses_and_field_samples():
    """qspectra tests/test_pulse.py:9-38 against qspectra_amd.pulse, and the
    host side of field-driven propagation: stage times follow the RK4
    schedule (the oracle's stepping), field samples are -i E (conj for '+')."""
    from qspectra_amd import pulse
    from qspectra_amd.simulate import eom
    with pytest.raises(TypeError, match="Can't instantiate abstract class"):
        pulse.Pulse()
    pump = pulse.CustomPulse(0, 1, lambda x, r: x + r)
    assert (pump.t_init, pump.t_final, pump(1, 2)) == (0, 1, 3)
    pump = pulse.GaussianPulse(carrier_freq=10, fwhm=1, t_peak=5, scale=100, freq_convert=1,
                               t_limits_multiple=2)
    assert_allclose(pump(np.array([5, 5.5]), 10), [100, 50])
    assert_allclose(pump(np.array([5, 5.5]), 5), [100, 50 * np.exp(2.5j)])
    assert pump.t_init == 5 - 2 * S.GAUSSIAN_SD_FWHM and pump.t_final == 5 + 2 * S.GAUSSIAN_SD_FWHM
    pump = pulse.GaussianPulse(carrier_freq=10, fwhm=2, t_peak=0, scale=100, freq_convert=2)
    assert_allclose(pump(np.array([0, 1]), 10.0), [100, 50])
    assert_allclose(pump(1, 0.0), 50 * np.exp(20.0j))
    # a scalar-only pulse function is sampled one time at a time
    scalar = pulse.CustomPulse(0, 2, lambda t, rw: complex(np.cos(float(t))) * (1 + rw))
    ts = np.linspace(0, 2, 7)
    assert_allclose(pulse.sample_field(scalar, ts, 1.0), 2 * np.cos(ts))
    amp, rate = pulse.field_rate(pulse.GaussianPulse(0., 1., scale=3.), 0.)
    sig = S.GAUSSIAN_SD_FWHM
    assert abs(amp - 3) < 1e-12 and abs(rate - np.exp(-0.5) / sig) < 1e-3 / sig
    # stage times: t0 = -1, outputs at -1 (no step), 0.5 (3 steps), 1.0 (1 step)
    t = np.array([-1.0, 0.5, 1.0])
    n, h = engine.rk4_schedule(t, -1.0, 0.5)
    tg, hg = eom.stage_times(t, -1.0, n, h)
    assert_allclose(tg, [-1.0, -0.5, 0.0, 0.5])
    assert_allclose(hg, [0.5, 0.5, 0.5, 0.5])
    no, ho = P.rk4_schedule(t, -1.0, 0.5)
    assert_equal(n, no) and assert_allclose(h, ho)
    g1 = pulse.GaussianPulse(3., 2., scale=0.5 + 0.2j)
    F = eom.field_samples([(g1, '-'), (g1, '+')], 1.0, tg, hg)
    for k, off in enumerate([0.0, 0.25, 0.5]):
        assert_allclose(F[:, k, 0], -1j * g1(tg + off, 1.0))
        assert_allclose(F[:, k, 1], -1j * np.conj(g1(tg + off, 1.0)))


def test_ell_pack_of_field_operators():
    """ELL form of the field operators reproduces every C_i (union pattern
    over terms and batch entries, zero padding)."""
    from qspectra_amd import _native
    rs = np.random.RandomState(2)
    C = np.zeros((3, 2, 9, 9), complex)
    mask = rs.rand(2, 9, 9) < 0.3
    C[:, mask] = rs.randn(3, int(mask.sum())) + 1j * rs.randn(3, int(mask.sum()))
    C[1, 0, 4, 7] = 2.0                       # pattern entry of one batch entry only
    col, coef = _native.ell_pack(C)
    assert col.shape[0] == 1 and coef.shape[:3] == (3, 9, col.shape[2]) and coef.shape[3] == 2
    D = np.zeros_like(C)
    for b in range(3):
        for i in range(9):
            for k in range(col.shape[2]):
                D[b, :, i, col[0, i, k]] += coef[b, i, k]
    assert_allclose(D, C)
    union = np.any(C != 0, axis=(0, 1)).sum(axis=1).max()
    assert col.shape[2] == union
    col0, coef0 = _native.ell_pack(np.zeros((1, 1, 3, 3)))
    assert col0.shape == (1, 3, 1) and not coef0.any()


def test_hamiltonian_variants_commute_and_keep_frames():
    """Rotating-frame and disorder-sample variants (hamiltonian.py:96-274
    semantics): they commute, a sample of a sample replaces the draw, member
    n draws its disorder from RandomState(seed + [n]) with sigma = FWHM *
    GAUSSIAN_SD_FWHM, samples keep rw_freq, rotating again re-targets."""
    import warnings
    from qspectra_amd.constants import GAUSSIAN_SD_FWHM
    from qspectra_amd.hamiltonian import check_hermitian
    H1 = np.array([[12400., 80., 10.], [80., 12500., -30.], [10., -30., 12600.]])
    H = qa.ElectronicHamiltonian(H1, disorder=100.0, random_seed=5)
    a = H.in_rotating_frame(12500.0).sample(3)
    b = H.sample(3).in_rotating_frame(12500.0)
    assert a is b and a.rw_freq == 12500.0 and a.base is H
    d = np.diag(100.0 * GAUSSIAN_SD_FWHM * np.random.RandomState([5, 3]).randn(3))
    assert np.array_equal(H.sample(3).H_1exc, H1 + d)
    assert np.array_equal(a.H_1exc, H1 + d - 12500.0 * np.eye(3))
    assert a.lab_frame is H.sample(3) and a.undisordered is H.in_rotating_frame(12500.0)
    assert H.sample(3).sample(4) is H.sample(4)
    assert H.in_rotating_frame(100.0).in_rotating_frame(200.0) is H.in_rotating_frame(200.0)
    assert np.array_equal(a.H1_nonrotating, H1 + d)
    twin = qa.ElectronicHamiltonian(H1, disorder=100.0, random_seed=5)
    assert twin.sample(3) == H.sample(3) and twin.sample(3) != H.sample(2)
    assert twin.in_rotating_frame(12500.0) != twin
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter('always')
        qa.ElectronicHamiltonian(H1).sample(0)
    assert any(issubclass(x.category, RuntimeWarning) for x in w)
    for bad in ([[1., 2.], [0., 1.]], np.ones((2, 3))):
        with pytest.raises(ValueError):
            check_hermitian(bad)


def test_arbitrary_bath_host():
    """ArbitraryBath (bath.py:50-56) is constructible (the reference's own
    constructor trips over a read-only property) and its real correlation
    function with the Debye J is the Debye one; it has no complex one."""
    from oracle import system as S
    T, lam, gam = S.CM_K * 77, 35.0, 106.0
    b = qa.ArbitraryBath(T, lambda x: 2 * lam * gam * x / (gam ** 2 + x ** 2), 2 * lam / gam)
    ref = S.DebyeBath(T, lam, gam)
    for x in (0.0, -250.0, 3.5, 900.0):
        assert abs(b.corr_func_real(x) - ref.corr_func_real(x)) <= 1e-14 * abs(ref.corr_func_real(x))
    assert not hasattr(b, 'corr_func_complex')
    assert b.spectral_density_limit_at_zero == 2 * lam / gam
    assert not qa.bath.is_debye_form(b)
