'''Batched radial stepping through uppe_batch_* against standalone handles and the FP64 oracle.

A batch member runs the same kernels as a standalone handle on the same inputs, so with one source launch shape
for both (autotune off, samples per thread pinned) the two agree to rounding of identical operation orders:
1e-12 relative L2 in FP64 and FP32 alike (in practice bit for bit). With autotune on, the source kernel may take another
number of samples per thread, which reorders the FP32 depletion sums: 1e-5 then, the FP32 field tolerance. Each
member is also checked against the oracle's track of its own vol_dict: 1e-10 (FP64) and 1e-5 (FP32) on the
field; 1e-10 / 1e-4 relative on the electron density, or 1e-12 of the gas density where the gas is barely
ionised (a weak field's rate is a high power of the field, which multiplies the field's rounding error).
'''
import sys

import numpy as np
import pytest

from conftest import rel_l2
from searay_amd import batch, configs, uppe_kernel

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _standalone(vol, precision, max_steps=0):
    rs = uppe_kernel.RunSetup(vol)
    st = rs.make_stepper(uppe_kernel.hip_stepper_factory(precision))
    st.set_field(rs.stepper_field())
    zi, zf, L = rs.plane(0)
    n, _ = st.advance(zi, L, 0, rs.subcycles, rs.dzmin, max_steps=max_steps)
    q = st.get_state()
    st.close()
    return n, q


def _mixed_vols(adaptive, n=6, steps=4):
    vols, drawn = configs.batched_radial(n=n, scale=8, steps=steps)
    if adaptive:
        for v in vols:
            v['minimum step'] = v['minimum step'] / 50.0
    return vols, drawn


@pytest.mark.parametrize('spt', ['1', '4'])
@pytest.mark.parametrize('adaptive', [False, True])
@pytest.mark.parametrize('precision', ['fp64', 'fp32', 'bf16x3'])
def test_batch_matches_standalone(precision, adaptive, spt, monkeypatch):
    # the same source launch shape (samples per thread) in the batch and in the standalone handles
    monkeypatch.setenv('SEARAY_UPPE_SOURCE_SPT', spt)
    vols, drawn = _mixed_vols(adaptive)
    _say('batch vs standalone %s adaptive=%s grids %s' % (precision, adaptive, [d[0] for d in drawn]))
    with batch.Batch(vols, precision=precision, autotune=False) as b:
        b.set_fields()
        steps = b.advance()
        qs = [m.get_state() for m in b.members]
        infos = [m.info() for m in b.members]
    for i, vol in enumerate(vols):
        n, q = _standalone(vol, precision)
        assert steps[i] == n, (i, steps[i], n)
        assert infos[i]['steps'] == n
        assert np.all(np.isfinite(q))
        assert rel_l2(qs[i], q) < 1e-12, (i, rel_l2(qs[i], q))


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_batch_autotuned_matches_standalone(precision):
    vols, drawn = _mixed_vols(False)
    _say('autotuned batch vs standalone %s' % precision)
    with batch.Batch(vols, precision=precision, autotune=True) as b:
        b.set_fields()
        b.advance()
        qs = [m.get_state() for m in b.members]
        plans = [b.plan(i) for i in range(len(vols))]
        # a second advance reuses the tuned shapes
        b.set_fields()
        b.advance()
        qs2 = [m.get_state() for m in b.members]
        assert [b.plan(i) for i in range(len(vols))] == plans
    for i, vol in enumerate(vols):
        _, q = _standalone(vol, precision)
        assert rel_l2(qs[i], q) < 1e-5, (i, rel_l2(qs[i], q))
        assert rel_l2(qs2[i], qs[i]) == 0.0
    # members of one shape and one table set share a run
    by_grid = {}
    for i, d in enumerate(drawn):
        by_grid.setdefault(d[0], set()).add(plans[i]['group'])
    assert all(len(g) == 1 for g in by_grid.values())


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_batch_members_against_oracle(oracle_mod, precision):
    vols, drawn = _mixed_vols(True, n=4, steps=3)
    _say('batch members vs oracle %s' % precision)
    with batch.Batch(vols, precision=precision) as b:
        b.set_fields()
        b.advance()
        outs = []
        for rs, m in zip(b.setups, b.members):
            zi, zf, _ = rs.plane(0)
            a, j, chi, ne = m.finish(zf - zi)
            outs.append((a, ne))
    tol = 1e-10 if precision == 'fp64' else 1e-5
    ne_tol = 1e-10 if precision == 'fp64' else 1e-4
    for i, vol in enumerate(vols):
        ref = oracle_mod.oracle_track(vol)
        a, ne = outs[i]
        assert rel_l2(a, ref[0][..., 1]) < tol, (i, rel_l2(a, ref[0][..., 1]))
        # the electron density to ne_tol relative, or, where the gas is barely ionised, to 1e-12 of the gas
        # density (the rate's steep field dependence multiplies the field's rounding error there)
        ne_ref = np.fft.irfft(ref[3][..., 1], axis=0)
        ngas = vol['density reference']  # uniform gas: n_g = n_ref
        err, dev = rel_l2(ne, ne_ref), np.max(np.abs(ne - ne_ref)) / ngas
        _say('member %d %s: field %.2e, ne %.2e relative, %.2e of the gas density (peak ne %.2e)'
             % (i, drawn[i], rel_l2(a, ref[0][..., 1]), err, dev, np.max(np.abs(ne_ref)) / ngas))
        assert err < ne_tol or dev < 1e-12, (i, err, dev)


def test_batch_full_size_instances():
    '''The 64-instance workload (Nw = 1025, Nr in {64, ..., 512}, O2 / Ar, 0.1-5 P_cr), two steps: every member
    stays finite, and a member of each grid size agrees with its standalone run (FP32, autotuned).'''
    vols, drawn = configs.batched_radial(n=64, steps=200)
    _say('full-size batch: 64 instances, 2 steps')
    with batch.Batch(vols, precision='fp32') as b:
        b.set_fields()
        steps = b.advance(max_steps=2)
        assert np.all(steps == 2)
        qs = {}
        for i, m in enumerate(b.members):
            q = m.get_state()
            assert np.all(np.isfinite(q)), i
            if drawn[i][0] not in qs:
                qs[drawn[i][0]] = (i, q)
    for nr, (i, q) in sorted(qs.items()):
        _, qref = _standalone(vols[i], 'fp32', max_steps=2)
        assert rel_l2(q, qref) < 1e-5, (nr, i, rel_l2(q, qref))


def test_member_handles_belong_to_batch():
    vols, _ = _mixed_vols(False, n=2)
    with batch.Batch(vols, precision='fp32') as b:
        m = b.members[0]
        assert b.lib.uppe_destroy(m.h) != 0  # a member is released by its batch only
        b.set_fields()
        assert np.all(b.advance(max_steps=1) == 1)


@pytest.mark.parametrize('graph_steps', [1, 3])
def test_batch_graph_replays_match_kernel_launches(graph_steps):
    '''Graph replays (graph_steps steps each, several per advance, a new capture per advance, host reads in
    between) leave every member exactly where kernel-by-kernel launches leave it.'''
    vols, _ = _mixed_vols(True, n=5, steps=4)
    states = {}
    for gs in (0, graph_steps):
        with batch.Batch(vols, precision='fp32', autotune=False, graph_steps=gs) as b:
            b.set_fields()
            seq = []
            for max_steps in (2, 3, 0):
                steps = b.advance(max_steps=max_steps)
                seq.append((steps.copy(), [m.get_state() for m in b.members]))
            states[gs] = seq
    for (s0, q0), (s1, q1) in zip(states[0], states[graph_steps]):
        assert np.array_equal(s0, s1)
        for a, c in zip(q0, q1):
            assert np.array_equal(a, c)


@pytest.mark.parametrize('width', [16, 32])
@pytest.mark.parametrize('precision', ['fp64', 'fp32', 'bf16x3'])
def test_batch_tile_widths_match_standalone(precision, width, monkeypatch):
    '''Contraction tiles of 16 and 32 complex columns (every operator fragment feeding two or four MFMA column
    tiles) give the standalone handles' results exactly: the column tiling does not reorder any sum.'''
    monkeypatch.setenv('SEARAY_UPPE_TILE_WIDTH', str(width))
    monkeypatch.setenv('SEARAY_UPPE_SOURCE_SPT', '4')
    vols, drawn = _mixed_vols(True)
    _say('batch with %d-column tiles vs standalone %s' % (width, precision))
    with batch.Batch(vols, precision=precision, autotune=True) as b:
        b.set_fields()
        steps = b.advance()
        qs = [m.get_state() for m in b.members]
        for i, v in enumerate(vols):
            assert b.plan(i)['k2r']['tiles'] == -(-v['wave grid'][0] // width)
    for i, vol in enumerate(vols):
        n, q = _standalone(vol, precision)
        assert steps[i] == n
        assert rel_l2(qs[i], q) < 1e-12, (i, rel_l2(qs[i], q))


def test_batch_members_with_different_bands_split_runs(monkeypatch):
    '''Members of one grid whose nonlinear bands differ cannot share a launch: the batch splits them into runs
    (consecutive members with equal tables) and every member still equals its standalone run.'''
    monkeypatch.setenv('SEARAY_UPPE_SOURCE_SPT', '4')
    vols, drawn = configs.batched_radial(n=6, scale=8, steps=3)
    for v in vols:
        v['wave grid'] = (v['wave grid'][0], 32, 1, 2)
        v['radial modes'] = 32
    band = vols[0]['frequency band']
    for i in (2, 3):
        vols[i]['nonlinear band'] = (0.0, 0.6 * band[1])
    _say('batch with split runs (nonlinear bands differ)')
    with batch.Batch(vols, precision='fp32', autotune=False) as b:
        b.set_fields()
        steps = b.advance()
        qs = [m.get_state() for m in b.members]
        plans = [b.plan(i) for i in range(len(vols))]
    assert len({p['group'] for p in plans}) == 1
    assert [(p['first_slot'], p['run_length']) for p in plans] == [(0, 2), (0, 2), (2, 2), (2, 2), (4, 2), (4, 2)]
    for i, vol in enumerate(vols):
        n, q = _standalone(vol, 'fp32')
        assert steps[i] == n
        assert rel_l2(qs[i], q) < 1e-12, (i, rel_l2(qs[i], q))


@pytest.mark.parametrize('precision', ['fp64', 'fp32', 'bf16x3'])
def test_batch_with_rotator_matches_standalone(precision):
    '''Members 2 (alone in its grid) and 4 (the middle one of three Nr = 64 members) carry the 32-state N2 rotator:
    a member with a rotator splits its group into runs of its own, and each instance's rotator rows come from the
    per-instance table of the batched source launch. Same source launch shape in the batch and standalone
    (autotune off), so 1e-12 as above.'''
    vols, drawn = _mixed_vols(False)
    assert [d[0] for d in drawn[3:6]] == [64, 64, 64]
    for i in (2, 4):
        vols[i]['rotator'] = configs.n2_rotator(32)
    _say('batch with rotator vs standalone %s grids %s' % (precision, [d[0] for d in drawn]))
    with batch.Batch(vols, precision=precision, autotune=False) as b:
        b.set_fields()
        steps = b.advance()
        qs = [m.get_state() for m in b.members]
    for i, vol in enumerate(vols):
        n, q = _standalone(vol, precision)
        assert steps[i] == n
        assert np.all(np.isfinite(q))
        assert rel_l2(qs[i], q) < 1e-12, (i, rel_l2(qs[i], q))
    # the rotator changes the member's field: it is not lost in the comparison
    i = 4
    vol = dict(vols[i])
    del vol['rotator']
    _, q0 = _standalone(vol, precision)
    assert rel_l2(qs[i], q0) > 1e-6


@pytest.mark.parametrize('precision', ['fp64', 'fp32', 'bf16x3'])
def test_batch_members_share_only_equal_tables(precision, monkeypatch):
    '''A run reads one copy of the operator or of the kz tables only when every member's inputs for it are equal.
    Four members of one grid: a reference, another pulse power (same operator and tables), a denser gas (other
    kz) and a wider box (other kz; the same operator, whose eigenvectors and volume ratios do not depend on the
    cell size). The run shares the operator but not the tables; each member equals its standalone handle.'''
    monkeypatch.setenv('SEARAY_UPPE_SOURCE_SPT', '1')
    vols = [configs.radial_o2_ionization(scale=8, nr=32, steps=4, length='1 mm', power=p)
            for p in (3.0, 1.0, 3.0, 3.0)]
    vols[2]['density reference'] = 1.5 * vols[2]['density reference']
    vols[3]['size'] = (2 * vols[3]['size'][0], 2 * vols[3]['size'][1], vols[3]['size'][2])
    vols[3]['object'] = configs._box(vols[3])
    _say('batch sharing check %s' % precision)
    with batch.Batch(vols, precision=precision, autotune=False) as b:
        b.set_fields()
        steps = b.advance()
        qs = [m.get_state() for m in b.members]
    for i, vol in enumerate(vols):
        n, q = _standalone(vol, precision)
        assert steps[i] == n
        assert rel_l2(qs[i], q) < 1e-12, i
