"""CPU: the MBPT(2) oracle (oracle/mbpt.py) and the host side of
atomx_amd.mbpt.

The oracle's J-coupled pair sum (methods/mbpt.py:602-790, restated) is
pinned by the m-scheme second-order sum
    E_ab = sum_{i in a, j in b (i < j for a = b)} sum_{r < s}
           |<ij||rs>|^2 / (e_i + e_j - e_r - e_s)
over the spinors of the same orbitals, with the antisymmetrised integrals
of the spinor oracle (3j algebra, itself pinned by the Dirac-Fock energy):
angular-momentum recoupling makes the two equal, so a wrong phase, 6j
argument order, (2J+1) weight, pair norm or channel double count fails it.
The 6j symbols are pinned by closed forms and orthogonality; the product's
channel coefficients are checked by applying them to the oracle's Slater
tables."""

import itertools
import math

import numpy as np
import pytest

import oracle as O
from oracle import mbpt as OM
from oracle.spinor import SignedIntegrals


def _scf(Z, shells, ns, r_max):
    from atomx_amd import bspline as B
    from atomx_amd import dirac as D
    t = B.make_knots(ns, 7, 1e-4, r_max)
    bd = B.prepare_bspline_data(t, ns, 7)
    occ = [s.occupation for s in shells]
    kap = [s.kappa for s in shells]
    banks = O.Banks(bd)
    tables = O.angular_tables(occ, kap)

    class Backend:
        def fock(self, i, P, Q):
            return O.build_hf_matrix_dbsr(i, P, Q, occ, kap, bd, banks, skip_pq_exchange=True,
                                          tables=tables)

        def energy(self, P, Q):
            return O.two_electron_energy(P, Q, occ, kap, bd, banks)

    scf = D.DiracFockSCF(Z, shells, bd, backend=Backend()).bare_nucleus_start()
    scf.run(tol=1e-8)
    return scf, banks


@pytest.fixture(scope="module")
def neon_box():
    """Ne 1s2 2s2 2p6 in a 6 bohr box (30 splines, order 7): compact low
    virtuals, so a few per kappa carry a large second-order energy."""
    from atomx_amd import dirac as D
    return _scf(10, D.neon_subshells(), 30, 6.0)


def _spinors(kl, occ, vir):
    orbs, eps, occ_sp, vir_sp = {}, {}, {}, []
    for kappa in kl:
        mult = list(range(-2 * abs(kappa) + 1, 2 * abs(kappa), 2))
        for i, (P, Q, e, lab) in enumerate(occ[kappa]):
            key = (100 + i, kappa)
            orbs[key], eps[key] = (P, Q), e
            occ_sp[lab] = [key + (m,) for m in mult]
        for j, (P, Q, e) in enumerate(vir[kappa]):
            key = (200 + j, kappa)
            orbs[key], eps[key] = (P, Q), e
            vir_sp += [key + (m,) for m in mult]
    return orbs, eps, occ_sp, vir_sp


def test_reference_angular_unit_values():
    """The reference's own unit values (tests/test_mbpt2.py:20-46) on the
    product functions."""
    from atomx_amd.mbpt import _allowed_k_range, reduced_C_jj
    assert math.isclose(reduced_C_jj(-1, -1, 0), -math.sqrt(2.0), abs_tol=1e-12)
    assert math.isclose(reduced_C_jj(-1, -2, 1), 2.0 / math.sqrt(3.0), abs_tol=1e-12)
    assert reduced_C_jj(-1, -1, 1) == 0.0
    assert _allowed_k_range(-1, -2) == [1]
    assert _allowed_k_range(-3, 2) == [2, 4]
    assert OM.allowed_k(-3, 2) == [2, 4] and OM.allowed_k(-1, -2) == [1]


def test_6j_closed_forms_orthogonality_and_product_agreement():
    from atomx_amd.angular import wigner_6j
    for f in (O.wigner_6j_exact, wigner_6j):
        # {a b c; 0 c b} = (-1)^(a+b+c) / sqrt((2b+1)(2c+1))
        for ta, tb, tc in itertools.product(range(7), repeat=3):
            if (ta + tb + tc) % 2 or not abs(tb - tc) <= ta <= tb + tc:
                continue
            want = (-1) ** ((ta + tb + tc) // 2) / math.sqrt((tb + 1) * (tc + 1))
            assert f(ta, tb, tc, 0, tc, tb) == pytest.approx(want, abs=1e-14)
        # sum_x (2x+1)(2f+1) {a b x; c d f}{a b x; c d g} = delta_fg
        for ta, tb, tc, td in ((3, 1, 5, 3), (5, 3, 5, 7), (7, 7, 5, 3)):
            fs = [t for t in range(0, 16) if (ta + td + t) % 2 == 0]
            xs = range(0, 16)
            for tf, tg in itertools.product(fs, repeat=2):
                s = sum((tx + 1) * (tf + 1) * f(ta, tb, tx, tc, td, tf) * f(ta, tb, tx, tc, td, tg)
                        for tx in xs)
                if any(f(ta, tb, tx, tc, td, tf) for tx in xs):
                    assert s == pytest.approx(1.0 if tf == tg else 0.0, abs=1e-13)
    rng = np.random.default_rng(1)
    for a in rng.integers(0, 10, size=(3000, 6)):
        assert wigner_6j(*map(int, a)) == pytest.approx(O.wigner_6j_exact(*map(int, a)), abs=1e-14)


def test_oracle_pair_energies_equal_the_m_scheme_sums(neon_box):
    scf, banks = neon_box
    E2, pairs = OM.mbpt2_energy(scf, kappa_max=2, n_virt_max=3, banks=banks)
    assert E2 == pytest.approx(sum(pairs.values()), rel=1e-14) and E2 < -1e-2
    kl, occ, vir = OM.bases(scf, 2, 3, banks=banks)
    orbs, eps, occ_sp, vir_sp = _spinors(kl, occ, vir)
    g = SignedIntegrals(orbs, scf.bd, banks).g
    vpairs = list(itertools.combinations(vir_sp, 2))
    for la, lb in (("2s1/2", "2p1/2"), ("2p3/2", "2p3/2"), ("1s1/2", "2p3/2"), ("2s1/2", "2s1/2")):
        ij = (list(itertools.combinations(occ_sp[la], 2)) if la == lb
              else [(i, j) for i in occ_sp[la] for j in occ_sp[lb]])
        em = 0.0
        for i, j in ij:
            for r, s in vpairs:
                if i[2] + j[2] != r[2] + s[2]:
                    continue
                v = g(i, j, r, s)
                if v:
                    em += v * v / (eps[i[:2]] + eps[j[:2]] - eps[r[:2]] - eps[s[:2]])
        assert em < 0.0
        assert pairs[la, lb] == pytest.approx(em, rel=1e-13), (la, lb)


def test_oracle_beryllium_total_equals_the_m_scheme_sum():
    from atomx_amd import dirac as D
    scf, banks = _scf(4, [D.Subshell("1s", 1, -1, 2.0), D.Subshell("2s", 2, -1, 2.0)], 30, 30.0)
    E2, _ = OM.mbpt2_energy(scf, kappa_max=2, n_virt_max=4, banks=banks)
    kl, occ, vir = OM.bases(scf, 2, 4, banks=banks)
    orbs, eps, occ_sp, vir_sp = _spinors(kl, occ, vir)
    g = SignedIntegrals(orbs, scf.bd, banks).g
    occ_all = occ_sp["1s"] + occ_sp["2s"]
    em = 0.0
    for i, j in itertools.combinations(occ_all, 2):
        for r, s in itertools.combinations(vir_sp, 2):
            if i[2] + j[2] == r[2] + s[2]:
                v = g(i, j, r, s)
                em += v * v / (eps[i[:2]] + eps[j[:2]] - eps[r[:2]] - eps[s[:2]])
    assert E2 == pytest.approx(em, rel=1e-13) and E2 < 0.0


def test_product_channel_coefficients_reproduce_the_oracle(neon_box):
    """The product's per-channel J coefficients, applied to the oracle's
    Slater tables in the kernel's r <= s order, give the oracle's pair
    energies (the host half of the device pair sum)."""
    from atomx_amd.mbpt import _channel, _kappa_list
    scf, banks = neon_box
    for kmax, nv, mm in ((2, 3, None), (2, 3, 1), (1, 4, None)):
        _, pairs = OM.mbpt2_energy(scf, kappa_max=kmax, n_virt_max=nv, multipole_max=mm,
                                   banks=banks)
        kl, occ, vir = OM.bases(scf, kmax, nv, banks=banks)
        assert kl == _kappa_list(kmax)
        stack = [(ci, e, (P, Q)) for ci, kappa in enumerate(kl) for (P, Q, e) in vir[kappa]]
        cols = [x[2] for x in stack]
        okl = [k for k in kl if occ[k]]
        for ia, ka in enumerate(okl):
            for kb in okl[ia:]:
                for i, (Pa, Qa, ea, la) in enumerate(occ[ka]):
                    for j in range(i if ka == kb else 0, len(occ[kb])):
                        Pb, Qb, eb, lb = occ[kb][j]
                        T = [OM.rk_table(((Pa, Qa), (Pb, Qb)), (cols, cols), scf.bd, k, banks)
                             for k in range(8)]
                        e = 0.0
                        for r, s in itertools.combinations_with_replacement(range(len(stack)), 2):
                            c = _channel(ka, kb, kl[stack[r][0]], kl[stack[s][0]], mm)
                            de = stack[r][1] + stack[s][1] - ea - eb
                            if c is None or abs(de) < 1e-10:
                                continue
                            lo, nj, cd, ce, _ = c
                            amp = np.zeros(nj)
                            for (jj, k), v in cd.items():
                                amp[jj] += v * T[k][r, s]
                            for (jj, k), v in ce.items():
                                amp[jj] -= v * T[k][s, r]
                            js = sum((lo + 2 * x + 1) * amp[x] ** 2 for x in range(nj))
                            norm = (2.0 if (ka == kb and i == j) else 1.0) * (2.0 if r == s else 1.0)
                            e += -js / (norm * de)
                        assert e == pytest.approx(pairs[la, lb], rel=1e-13, abs=1e-18), (la, lb)


def test_open_shell_and_unconverged_references_are_rejected():
    from atomx_amd import dirac as D
    from atomx_amd.mbpt import mbpt2_energy

    class State:
        Z, bd, nuclear, A = 3, {"ns": 10, "ks": 5}, "point", None
        shells = [D.Subshell("1s", 1, -1, 2.0), D.Subshell("2s", 2, -1, 1.0)]
        P = Q = np.zeros((2, 10))
        eps = np.zeros(2)

    with pytest.raises(ValueError, match="closed-shell"):
        mbpt2_energy(State(), kappa_max=1, n_virt_max=5)
    state = State()
    state.converged = False
    with pytest.raises(ValueError, match="converged"):
        mbpt2_energy(state)
