"""CPU restatement of the DF solver's "orbital" mode Fock pieces — TEST
INFRASTRUCTURE ONLY (imported by tests/ only, as the checker).

For occupation-product input (no mean-field seam):

* ``direct_potential``: methods/dirac_fock.py:840-921 — weighted banded
  densities (occ_b, occ_b - 1 for the target), four case-1 contractions at
  k = 0 on the full banks;
* ``exchange_matrix_hf``: :969-1068 — per source b and exchange rank k,
  c_k = occ_b Lambda / ((2j_a+1)(2j_b+1)) times the gather case-4
  contraction of the source's full densities, then symmetrised;
* ``intra_shell_potential``: :1469-1551 — -(q_a - 1) Lambda(k) /
  (2j (2j+1)) times the case-1 contractions of the target's own density for
  even k = 2..2j;
* ``orbital_fock_blocks``: PP = V_J - K + V_i, QQ likewise, PQ = -K_PQ,
  QP = -K_QP (the assembly of :2214-2227).

Pinned by tests/test_orbital_mode_cpu.py: on the DBSR_HF neon spinors of the
reference fixture (closed shells) the orbital-mode Fock expectation
<a|h0 + blocks|a> of every orbital equals the fixture's Fock expectation at
2e-10 Ha, and the direct potential contracts with a density to the scalar
cell F^0 sum.
"""

from __future__ import annotations

import numpy as np

from . import (Banks, banded_to_full_add, convol_case1, convol_case4_gather, density_full,
               density_sym, exchange_k_range, exchange_lambda)


def _sizes(bd):
    ks = bd["ks"]
    return (bd.get("ksp", ks), bd.get("ksq", ks), bd.get("nsp", bd["ns"]),
            bd.get("nsq", bd["ns"]), bd["ns"])


def direct_potential(P, Q, occ, bd, banks, target=None):
    ksp, ksq, _, _, ns = _sizes(bd)
    ddp, ddq = np.zeros((ns, ksp)), np.zeros((ns, ksq))
    for j in range(len(occ)):
        c = occ[j] - 1.0 if j == target else occ[j]
        if abs(c) < 1e-10:
            continue
        ddp += c * density_sym(P[j][:ns], P[j][:ns], ns, ksp)
        ddq += c * density_sym(Q[j][:ns], Q[j][:ns], ns, ksq)
    vpp, vqq = np.zeros((ns, ns)), np.zeros((ns, ns))
    banded_to_full_add(vpp, convol_case1(banks.get_rkb(0, "pppp")[0], ddp, ns, ksp, ksp), ns,
                       ksp, 1.0)
    banded_to_full_add(vpp, convol_case1(banks.get_rkb(0, "pqpq")[0], ddq, ns, ksp, ksq), ns,
                       ksp, 1.0)
    banded_to_full_add(vqq, convol_case1(banks.get_rkb(0, "qpqp")[0], ddp, ns, ksq, ksp), ns,
                       ksq, 1.0)
    banded_to_full_add(vqq, convol_case1(banks.get_rkb(0, "qqqq")[0], ddq, ns, ksq, ksq), ns,
                       ksq, 1.0)
    return vpp, vqq


def exchange_matrix_hf(target, P, Q, occ, kappas, bd, banks):
    ksp, ksq, nsp, nsq, ns = _sizes(bd)
    ka = kappas[target]
    tja = 2 * abs(ka) - 1
    kpp, kqq, kpq, kqp = (np.zeros((ns, ns)) for _ in range(4))
    for b in range(len(occ)):
        if b == target:
            continue
        kb = kappas[b]
        tjb = 2 * abs(kb) - 1
        for k in exchange_k_range(ka, kb):
            lam = exchange_lambda(ka, kb, k)
            if lam < 1e-15:
                continue
            c = occ[b] * lam / ((tja + 1.0) * (tjb + 1.0))
            rpp = banks.get_rkb(k, "pppp")[0][:nsp, :nsp]
            rqq = banks.get_rkb(k, "qqqq")[0][:nsq, :nsq]
            kpp[:nsp, :nsp] += c * convol_case4_gather(
                rpp, density_full(P[b][:nsp], P[b][:nsp], nsp), nsp, ksp)
            kqq[:nsq, :nsq] += c * convol_case4_gather(
                rqq, density_full(Q[b][:nsq], Q[b][:nsq], nsq), nsq, ksq)
            kpq += c * convol_case4_gather(banks.get_rkb(k, "pqpq")[0],
                                           density_full(P[b], Q[b], ns), ns, ksp, ksq)
            kqp += c * convol_case4_gather(banks.get_rkb(k, "qpqp")[0],
                                           density_full(Q[b], P[b], ns), ns, ksq, ksp)
    kpp = 0.5 * (kpp + kpp.T)
    kqq = 0.5 * (kqq + kqq.T)
    kpq = 0.5 * (kpq + kqp.T)
    return kpp, kqq, kpq, kpq.T.copy()


def intra_shell_potential(target, P, Q, occ, kappas, bd, banks):
    ksp, ksq, _, _, ns = _sizes(bd)
    vpp, vqq = np.zeros((ns, ns)), np.zeros((ns, ns))
    ka = kappas[target]
    tja = 2 * abs(ka) - 1
    qa = occ[target]
    if tja <= 1 or qa * (qa - 1.0) < 1e-15:
        return vpp, vqq
    la = ka if ka > 0 else -ka - 1
    dp = density_sym(P[target][:ns], P[target][:ns], ns, ksp)
    dq = density_sym(Q[target][:ns], Q[target][:ns], ns, ksq)
    for k in range(2, tja + 1):
        if (k + 2 * la) % 2:
            continue
        lam = exchange_lambda(ka, ka, k)
        if lam < 1e-15:
            continue
        c = -(qa - 1.0) * lam / (tja * (tja + 1.0))
        banded_to_full_add(vpp, convol_case1(banks.get_rkb(k, "pppp")[0], dp, ns, ksp, ksp), ns,
                           ksp, c)
        banded_to_full_add(vqq, convol_case1(banks.get_rkb(k, "qpqp")[0], dp, ns, ksq, ksp), ns,
                           ksq, c)
        banded_to_full_add(vqq, convol_case1(banks.get_rkb(k, "qqqq")[0], dq, ns, ksq, ksq), ns,
                           ksq, c)
        banded_to_full_add(vpp, convol_case1(banks.get_rkb(k, "pqpq")[0], dq, ns, ksp, ksq), ns,
                           ksp, c)
    return vpp, vqq


def orbital_fock_blocks(target, P, Q, occ, kappas, bd, banks=None):
    """(PP, QQ, PQ, QP) two-electron blocks of the orbital-mode Fock matrix."""
    if banks is None:
        banks = Banks(bd)
    vpp, vqq = direct_potential(P, Q, occ, bd, banks, target)
    kpp, kqq, kpq, kqp = exchange_matrix_hf(target, P, Q, occ, kappas, bd, banks)
    ipp, iqq = intra_shell_potential(target, P, Q, occ, kappas, bd, banks)
    return vpp - kpp + ipp, vqq - kqq + iqq, -kpq, -kqp
