"""CPU: the DIIS extrapolation and the per-kappa Lowdin retraction
(atomx_amd/diis.py), and the SCF with DIIS on the oracle backend reaching
the same fixed point as without it."""

import numpy as np
import pytest

from atomx_amd.diis import DiisHistory, extrapolate, lowdin_by_kappa


def test_affine_map_fixed_point_in_one_extrapolation():
    # F(y) = A y + b, contractive; three iterates span the plane, so the
    # extrapolation of the outputs is the fixed point itself
    rng = np.random.default_rng(0)
    A = np.array([[0.6, 0.3], [-0.2, 0.5]])
    b = np.array([1.0, -2.0])
    fixed = np.linalg.solve(np.eye(2) - A, b)
    h = DiisHistory(8)
    y = rng.normal(size=2)
    for _ in range(3):
        x = A @ y + b
        h.push(x, x - y)
        y = x + 0.1 * rng.normal(size=2)  # inputs off the plain iteration
    assert np.allclose(h.extrapolate(), fixed, rtol=0, atol=1e-9)


def test_history_depth_and_rejections():
    h = DiisHistory(3)
    for i in range(5):
        h.push([float(i)], [1.0 / (i + 1)])
    assert len(h) == 3 and [v[0] for v in h.x] == [2.0, 3.0, 4.0]
    assert extrapolate([np.ones(2)], [np.ones(2)]) is None      # one entry
    # cancelling (1, 0) against (1.1, 0) needs c = (11, -10): rejected
    r = [np.array([1.0, 0.0]), np.array([1.1, 0.0])]
    assert extrapolate([np.zeros(2), np.ones(2)], r) is None
    r = [np.array([1.0, 0.0]), np.array([-1.0, 0.0])]  # c = (1/2, 1/2) is fine
    assert np.allclose(extrapolate([np.zeros(2), np.ones(2)], r), [0.5, 0.5])
    # coefficients sum to one: equal residuals give the mean iterate
    x = [np.array([0.0, 2.0]), np.array([2.0, 0.0])]
    same = [np.array([1.0, 1.0]), np.array([1.0, 1.0])]
    assert np.allclose(extrapolate(x, same), [1.0, 1.0])


def test_lowdin_orthonormal_in_each_kappa_metric():
    rng = np.random.default_rng(1)
    ns = 6
    M = rng.normal(size=(2 * ns, 2 * ns))
    S = M @ M.T + 2 * ns * np.eye(2 * ns)
    kappas = [-1, -1, 1, -1]
    P, Q = rng.normal(size=(4, ns)), rng.normal(size=(4, ns))
    P2, Q2 = lowdin_by_kappa(P, Q, kappas, {-1: S, 1: S})
    for kap in (-1, 1):
        rows = [i for i, k in enumerate(kappas) if k == kap]
        V = np.concatenate((P2[rows], Q2[rows]), axis=1).T
        assert np.allclose(V.T @ S @ V, np.eye(len(rows)), atol=1e-12)
    # the symmetric retraction is the closest orthonormal set: already
    # orthonormal input is returned unchanged
    P3, Q3 = lowdin_by_kappa(P2, Q2, kappas, {-1: S, 1: S})
    assert np.allclose(P3, P2, atol=1e-12) and np.allclose(Q3, Q2, atol=1e-12)
    with pytest.raises(ValueError):
        lowdin_by_kappa(np.vstack([P[:1], P[:1]]), np.vstack([Q[:1], Q[:1]]), [-1, -1],
                        {-1: S})


def test_scf_with_diis_reaches_the_same_energy():
    import oracle as O
    from atomx_amd import dirac as D
    from atomx_amd.batch_solve import fe_ion_subshells
    from atomx_amd.workloads import fe_ion_workload
    ion = fe_ion_workload(seed=0)[2]        # 9 orbitals, ns = 41, 17 plain sweeps
    bd = ion["bsp_data"]
    shells = fe_ion_subshells(2)
    occ, kap = [s.occupation for s in shells], [s.kappa for s in shells]
    banks = O.Banks(bd)
    tables = O.angular_tables(occ, kap)

    class Backend:
        def fock(self, i, P, Q):
            return O.build_hf_matrix_dbsr(i, P, Q, occ, kap, bd, banks, skip_pq_exchange=True,
                                          tables=tables)

        def energy(self, P, Q):
            return O.two_electron_energy(P, Q, occ, kap, bd, banks)

    runs = {}
    for diis in (False, True):
        scf = D.DiracFockSCF(26, shells, bd, backend=Backend()).bare_nucleus_start()
        applied = []
        step = scf.diis_step
        scf.diis_step = lambda x, c: applied.append(step(x, c)) or applied[-1]
        runs[diis] = (scf.run(tol=1e-9, diis=diis), scf.iterations, sum(applied))
    assert runs[True][2] >= 1                       # an extrapolation was taken
    assert runs[True][1] <= runs[False][1]
    assert abs(runs[True][0] - runs[False][0]) < 1e-10
