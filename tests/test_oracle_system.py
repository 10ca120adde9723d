"""Pins the oracle's system layer (operators, polarization, bath, H).

Expected values are the reference's own known answers
(tests/test_operator_tools.py, test_polarization.py, test_hamiltonian.py,
test_bath.py of qspectra) and the golden fixtures written by the reference.
"""
import numpy as np
from numpy.testing import assert_allclose, assert_equal

from oracle import system as S
from tests.conftest import golden


def test_operator_tools_known_answers():
    # qspectra tests/test_operator_tools.py:74-119
    M = np.array([[1., 2 - 2j], [2 + 2j, 3]])
    assert S.all_states(2) == [[], [0], [1], [0, 1]]
    assert_allclose(S.operator_1_to_2(M), [[4]])
    assert_allclose(S.operator_1_to_2(np.diag([1, 10, 100])), np.diag([11, 101, 110]))
    assert_allclose(S.operator_extend(M), [[0, 0, 0, 0], [0, 1, 2 - 2j, 0],
                                           [0, 2 + 2j, 3, 0], [0, 0, 0, 4]])
    assert_allclose(S.transition_operator(0, 2, 'ge'), [[0, 1, 0], [1, 0, 0], [0, 0, 0]])
    assert_allclose(S.transition_operator(0, 2, 'gef', '-'),
                    [[0, 1, 0, 0], [0, 0, 0, 0], [0, 0, 0, 1], [0, 0, 0, 0]])
    assert_allclose(S.transition_operator(0, 2, 'gef', '+'),
                    [[0, 0, 0, 0], [1, 0, 0, 0], [0, 0, 0, 0], [0, 0, 1, 0]])
    assert_equal(S.n_excitations(3), [1, 3, 3])
    assert S.hilbert_subspace_index('f', 'gef', 3) == slice(4, 7)
    assert S.hilbert_subspace_index('f', 'ef', 3) == slice(3, 6)
    assert S.full_liouville_subspace('ge,ee') == 'gg,ge,eg,ee'
    U = np.array([[1, 1], [-1, 1]]) / np.sqrt(2)
    assert_allclose(S.basis_transform_operator(np.array([[1, -1], [-1, 1]]) / 2.0, U),
                    [[1, 0], [0, 0]], atol=1e-15)
    assert_allclose(S.basis_transform_vector([0.5, -0.5, -0.5, 0.5], U), [1, 0, 0, 0],
                    atol=1e-15)


def test_polarization_known_answers():
    # qspectra tests/test_polarization.py:22-63
    assert_allclose(S.invariant_weights_4th_order('xxxx'), np.ones(3) / 15)
    assert_allclose(S.invariant_weights_4th_order('xxyy'), np.array([4, -1, -1]) / 30.0)
    ma = S.MAGIC_ANGLE
    assert_allclose(S.invariant_weights_4th_order([0, 0, ma, ma]), [1 / 9.0, 0, 0], atol=1e-10)
    assert_allclose(S.invariant_weights_4th_order([np.pi / 4, -np.pi / 4, np.pi / 2, 0]),
                    [0, 1 / 12.0, -1 / 12.0], atol=1e-10)
    assert S.invariant_polarizations(S.FOURTH_ORDER_INVARIANTS[0]) == [
        'xxxx', 'xxyy', 'xxzz', 'yyxx', 'yyyy', 'yyzz', 'zzxx', 'zzyy', 'zzzz']


def test_hamiltonian_known_answers():
    # qspectra tests/test_hamiltonian.py:82-106, 118-125, 154-173
    M = np.array([[1., 0], [0, 3]])
    H = S.ElectronicHamiltonian(M, dipoles=[[1, 0, 0], [0, 1, 0]], disorder=1,
                                energy_spread_extra=1.0)
    assert H.freq_step == 10.0 and H.time_step == 0.1
    assert_allclose(H.E('gef'), [0, 1, 3, 4])
    assert_allclose(H.dipole_operator('gef', 'x', '-+'),
                    [[0, 1, 0, 0], [1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]])
    assert_allclose(H.number_operator(1, 'gef'), np.diag([0, 0, 1, 1]))
    Hr = H.in_rotating_frame()
    assert_allclose(Hr.H('e'), [[-1, 0], [0, 1]])
    assert_allclose(Hr.E('gef'), [0, -1, 1, 0])
    assert Hr.freq_step == 4.0

    class Bath(object):
        temperature = 2
    H.bath = Bath()
    assert_allclose(H.thermal_state('e'), 1 / (np.exp(0.5) + np.exp(-0.5))
                    * np.array([[np.exp(0.5), 0], [0, np.exp(-0.5)]]))
    h2 = S.ElectronicHamiltonian(np.array([[1., 2], [2, 3]]))
    assert_allclose(S.basis_transform_vector([0.2, -0.4, -0.4, 0.8], h2.U('e')),
                    [0.7236068, 0.4472136, 0.4472136, 0.2763932], rtol=1e-6)


def test_sampling_matches_reference():
    g = golden('redfield_7site')
    H = S.ElectronicHamiltonian(g['in_H1'], disorder=100 / S.GAUSSIAN_SD_FWHM,
                                energy_spread_extra=float(g['extra']))
    for n in range(4):
        assert_allclose(H.sample(n).H_1exc, g['sample_H1_%d' % n], rtol=0, atol=1e-12)
    hs = H.in_rotating_frame().sample(2)
    assert_allclose(hs.E('gef'), g['sample2_rot_E_gef'], atol=1e-9)
    assert hs.time_step == float(g['sample2_time_step'])


def test_bath_against_reference_and_closed_forms():
    g = golden('bath')
    b77 = S.DebyeBath(S.CM_K * 77, 35, 106)
    b300 = S.DebyeBath(S.CM_K * 300, 35, 106)
    x = g['x']
    assert_allclose([b77.corr_func_complex(v) for v in x], g['c77'], rtol=1e-13, atol=1e-13)
    assert_allclose([b300.corr_func_complex(v) for v in x], g['c300'], rtol=1e-13, atol=1e-13)
    assert_allclose([b77.corr_func_real(v) for v in x], g['r77'], rtol=1e-13)
    # x = 0 closed forms (bath.py:28-29, :96-97)
    assert b77.corr_func_complex(0) == 35 * (2 * b77.temperature / 106 - 1j)
    # the truncated Matsubara sum evaluated in extended precision
    T, lam, gam = b77.temperature, 35.0, 106.0
    for xv in (-400.0, 37.5, 810.0):
        nu = 2 * np.pi * np.arange(1000, dtype=np.longdouble) * np.longdouble(T)
        xl = np.longdouble(xv)
        den_r = (nu ** 2 - gam ** 2) * nu
        den_i = -(nu ** 2 - gam ** 2) * xl
        mag = den_r ** 2 + den_i ** 2
        with np.errstate(invalid='ignore', divide='ignore'):
            sr = np.where(nu > 0, nu * den_r / mag, 0).sum()
            si = np.where(nu > 0, -nu * den_i / mag, 0).sum()
        cot = 1 / np.tan(np.longdouble(gam) / (2 * np.longdouble(T)))
        a = complex(float(cot), -1.0) / complex(gam, -xv)
        ref = lam * gam * (a + 4 * T * complex(float(sr), float(si)))
        assert abs(b77.corr_func_complex(xv) - ref) < 1e-12 * abs(ref)


def test_arbitrary_bath_pins():
    """ArbitraryBath (bath.py:50-56) with the Debye spectral density is the
    Debye real correlation function, itself pinned by the golden C(w); for a
    super-ohmic J the detailed-balance identities C(x) - C(-x) = J(x) and
    C(-x) = exp(-x/T) C(x) fix the rest, and x = 0 gives T J'(0)."""
    g = golden('bath')
    T, lam, gam = S.CM_K * 77, 35.0, 106.0
    debye = S.ArbitraryBath(T, lambda x: 2 * lam * gam * x / (gam ** 2 + x ** 2),
                            2 * lam / gam)
    assert_allclose([debye.corr_func_real(v) for v in g['x']], g['r77'], rtol=1e-13)
    assert debye.corr_func_real(0) == T * 2 * lam / gam
    wc = 150.0
    J = lambda x: lam * x ** 3 / wc ** 2 * np.exp(-x / wc)
    sup = S.ArbitraryBath(T, J, 0.0)
    assert sup.corr_func_real(0) == 0.0
    for x in (0.5, 12.0, 80.0, 410.0, 1500.0):
        cp, cm = sup.corr_func_real(x), sup.corr_func_real(-x)
        assert abs((cp - cm) - J(x)) <= 1e-12 * abs(J(x))
        assert abs(cm - np.exp(-x / T) * cp) <= 1e-12 * abs(cp)
    assert not hasattr(sup, 'corr_func_complex')
