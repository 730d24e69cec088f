"""The reference's multi-threaded assembly strategies of
Domain::computeInternalForces (FE/Domain.C:130-738), restated in the oracle
for the CPU baseline: Jobs::mapInitial's centroid-sorted partition
(FE/Jobs.C:104-165), per-thread partial Fint (Vector::gatherFrom's reverse
local-node order for case 6, LA/Vector.C:828-889; forward for 2, 3, 7) and
the fixed-order reduction over threads (FE/Domain.C:309-316).

Pinned by an independent reassembly: the per-element forces the oracle
computed, summed in numpy in case 6's order from the partition it reports,
must give its Fint bit for bit; the partition must be the (y, x) centroid
order cut into the reference's shares; and every strategy agrees with the
element-order sum to round-off.
"""
import numpy as np
import pytest

import oracle_lib as O

COPPER = dict(E=115e9, nu=0.3, rho=8930.0, A=400e6, B=100e6, n=1.0)


def bar():
    import redynela_amd as R
    return R.mesh.ogrid_bar(4, 2, 8, 3.2e-3, 32.4e-3, 0.4, 227.0)


@pytest.mark.parametrize("threads", [1, 3, 4, 7])
def test_case6_partition_and_order_reassembled_independently(threads):
    mesh = bar()
    m = O.Model(mesh, COPPER, dict(rho_b=1.0), mathmode=1)
    m.set_strategy(6, threads)
    assert m.begin(1.0) == 1
    for _ in range(3):
        assert m.step() == 1
    el, off = m.jobs()
    ne, nen = mesh["conn"].shape
    # the partition: every element once, equal shares as FE/Jobs.C:146-151 cuts them, (y, x) centroid order
    assert sorted(el.tolist()) == list(range(ne))
    nppr = ne / threads
    share = [min(int((i + 1) / nppr), threads - 1) for i in range(ne)]
    assert np.array_equal(np.diff(off), np.bincount(share, minlength=threads))
    # centroids as FE/Jobs.C:108-114 forms them (node sum in local order, then / n), initial coordinates
    cen = np.zeros((ne, 3))
    for e in range(ne):
        v = np.zeros(3)
        for i in range(nen):
            v = v + mesh["x"][mesh["conn"][e, i]]
        cen[e] = v / nen
    key = cen[el][:, [1, 0]]
    for a, b in zip(key[:-1], key[1:]):
        assert a[0] < b[0] or (a[0] == b[0] and a[1] <= b[1])
    # the forces of the last call, reassembled in case 6's order
    m.compute_internal_forces()
    Fe = m.get(O.F_FE).reshape(ne, nen, 3)
    part = np.zeros((threads, m.nn * 3))
    for t in range(threads):
        f = part[t]
        for e in el[off[t]:off[t + 1]]:
            for i in range(nen - 1, -1, -1):  # Vector::gatherFrom: I = n - 1 ... 0
                g = 3 * int(mesh["conn"][e, i])
                for j in range(3):
                    f[g + j] += Fe[e, i, j]
    fint = np.zeros(m.nn * 3)
    for i in range(m.nn * 3):
        v = 0.0
        for t in range(threads):
            v += part[t, i]
        fint[i] = v
    assert np.array_equal(fint, m.get(O.F_FINT))


@pytest.mark.parametrize("strategy", [2, 3, 6, 7])
def test_strategies_agree_with_element_order_to_round_off(strategy):
    # the same forces summed in another order: round-off apart until the
    # first Jacobi stop (|C_pq| < 1e-9, LA/Tensor2.C:1180) that the round-off
    # flips, after which a point's strain increment moves by ~1e-9 (on this
    # bar from increment 6); the displacements stay within 1e-9 of each other
    mesh = bar()
    ref = O.Model(mesh, COPPER, dict(rho_b=1.0), mathmode=1)
    m = O.Model(mesh, COPPER, dict(rho_b=1.0), mathmode=1)
    m.set_strategy(strategy, 4)
    assert ref.begin(1.0) == 1 and m.begin(1.0) == 1
    for k in range(20):
        assert ref.step() == 1 and m.step() == 1
        if k < 4:
            a, b = m.get(O.F_FINT), ref.get(O.F_FINT)
            assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
    x0 = mesh["x"].reshape(-1)
    du, dr = m.get(O.F_X) - x0, ref.get(O.F_X) - x0
    assert np.max(np.abs(du - dr)) <= 1e-9 * np.max(np.abs(dr))


def test_case6_is_deterministic_for_a_thread_count():
    mesh = bar()
    out = []
    for _ in range(2):
        m = O.Model(mesh, COPPER, dict(rho_b=1.0), mathmode=1)
        m.set_strategy(6, 5)
        m.begin(1.0)
        for _ in range(10):
            m.step()
        out.append(m.get(O.F_X))
    assert np.array_equal(out[0], out[1])


def test_unknown_strategy_refused():
    m = O.Model(bar(), COPPER, dict(rho_b=1.0), mathmode=1)
    with pytest.raises(ValueError):
        m.set_strategy(1, 4)
