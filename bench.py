'''Benchmark of the MI355X UPPE stepper.

Headline workload (the one the >= 20x target over the plain FP64 stepper is quoted on): radial UPPE, Kerr in dry air
(SimpleAir), 800 nm 50 fs pulse, w0 = 1 mm, 0.5 P_cr with 1 % seeded spectral-phase noise; Nw = 513 frequencies
over a 1 ps window, Nr = 128 Hankel modes; 100 fixed RK4 steps over 1 m. A "step" is one RK4 z-step (four
right-hand-side evaluations) over the whole grid; the timed region runs the first K steps of that run from the
initial field, on inputs resident in HBM.

Printed JSON line (one):
  value         grid-point z-steps per second = Nw * Nr * K / T, in BF16x3 (FP32 with the Hankel contraction as
                three BF16 matrix-core products, FP32 accumulation) by default
  roofline      the dominant kernel (largest share of kernel time in a repeat of the timed steps with
                dispatch-stamped HIP events on every launch) against HBM bandwidth: its algorithmic bytes per
                launch over its average launch duration. Per grid point in FP32: k->r contraction 28 B, source 16 B,
                r->k contraction 52 B over the nonlinear band's columns (96 B per right-hand side); doubled in FP64.
                A contraction is also reported against the matrix-core peak of the type it issues. `traffic` (HBM
                bytes per launch, rocprofv3 FETCH_SIZE and WRITE_SIZE, copy-calibrated) and `mfma_busy_frac` come
                from the committed counter summary of the same kernel and workload (profiles/counters_*.json) when
                there is one. `per_kernel` lists every step kernel's average duration and share
  cpu_baseline  the FP64 oracle (the reference's algorithm and pass structure, OpenMP) over the whole run
  extras        the fused FP64 and FP32 steppers and the plain FP64 stepper on the same steps, the speedup over
                the plain stepper, the field error against fused FP64 after the 100 steps; the same block for the
                radial O2 photoionisation workload (Nw = 2049, Nr = 256, 400 adaptive subcycles over 20 cm) with its
                own roofline; the 3-D Cartesian workload (Nw = 513, 256 x 256, 10 steps, FP32); the adjoint gradient
                of an output-spectrum objective (wall and device time against a forward run, checkpoints every 10
                and every stage kept, errors against the FP64 adjoint and against central finite differences of the
                FP64 forward), also on the photoionisation workload; the 64-instance batched workload against the
                same instances run one after another.

``--workload kerr`` runs the headline alone; ``--workload o2`` the photoionisation workload as the headline
(``uppe`` is the same); ``--workload cartesian``, ``batch``, ``adjoint`` and ``paraxial`` one workload each (the
paraxial one: the focused air filament of the reference's paraxial example, Nw = 256 envelope frequencies,
Nr = 256 with 128 modes, 400 subcycles in the first 2 cm plane).
'''
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from searay_amd import configs, paraxial_kernel, uppe_kernel  # noqa: E402

METRIC = 'UPPE grid-pt z-steps/s (1 GPU) vs FP64 plain at 1e-5 rel err; adjoint/fwd cost'
METRIC_PARAXIAL = 'paraxial envelope grid-pt z-steps/s (1 GPU)'
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {'fp64': 78.6, 'fp32': 157.3, 'bf16x3': 2500.0}
DTYPE = {'fp64': 'f64', 'fp32': 'f32', 'bf16x3': 'f32 (Hankel: bf16x3 MFMA, f32 accumulate)', 'fp64-plain': 'f64'}


def log(msg):
    print('[bench] ' + msg, file=sys.stderr, flush=True)


def run_steps(rs, st, K, schedule=None, record=False):
    zi, zf, L = rs.plane(0)
    st.set_field(rs.stepper_field())
    st.synchronize()
    t0 = time.perf_counter()
    n, rec = st.advance(zi, L, 0, rs.subcycles, rs.dzmin, max_steps=K, dz_schedule=schedule,
                        record_cap=K if record else 0)
    st.synchronize()
    t1 = time.perf_counter()
    return n, t1 - t0, rec


def nl_columns(rs):
    '''Complex columns the r->k contraction covers: up to one past the nonlinear filter's last nonzero bin, and at
    least the conditioner's four bins.'''
    if getattr(rs, 'nl_filter', None) is None:
        return int(rs.N[0])  # the envelope model filters nothing
    nz = np.nonzero(np.asarray(rs.nl_filter) != 0.0)[0]
    return int(min(rs.N[0], max((nz[-1] + 1) if nz.size else 0, 4)))


def kernel_work(rs, kernel, el, prec):
    '''Algorithmic bytes and MFMA flops (as issued) of one launch of a radial step kernel. Per grid point in FP32:
    k->r 28 B (the stage input in, E_r out: 16 B, plus 12 B of kz tables), source 16 B (E_r in, J_r out), r->k 52 B
    (J_r in and the RK stage: 40 B, plus the tables) over the columns of the nonlinear band only; 96 B per RHS in
    all. FP64 doubles them. The contractions issue 4 N_cols Nk Nr flops, three times that in BF16x3.'''
    T = rs.field_tool.GetTransverseTool()
    nw, nr, nk = rs.N[0], rs.N[1], getattr(T, 'Nk', rs.N[1])
    scale = el / 4.0
    issued = 3.0 if prec == 'bf16x3' else 1.0
    if kernel == 'hankel':
        return 28.0 * scale * nw * nr, 4.0 * nw * nk * nr * issued
    if kernel == 'hankel_r2k':
        f1 = nl_columns(rs)
        return 52.0 * scale * f1 * nk, 4.0 * f1 * nk * nr * issued
    if kernel == 'source':
        return 16.0 * scale * nw * nr, 0.0
    return None, None


KERNEL_NAMES = {'hankel': 'f_hankel k->r (MFMA contraction with the fused RK stage prologue)',
                'hankel_r2k': 'f_hankel r->k (MFMA contraction with the fused RK stage epilogue)',
                'source': 'f_source (strip transforms, rate, density scan, Kerr and plasma current)'}


def step_bytes(rs, el):
    '''Algorithmic bytes of one RK4 step: 4 right-hand sides x 96 B per grid point in FP32.'''
    return 384.0 * (el / 4) * rs.N[0] * rs.N[1]


def measured_counters(name):
    '''A committed rocprofv3 counter summary (profiles/<name>.json), if present.'''
    try:
        with open(os.path.join(ROOT, 'profiles', name + '.json')) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def roofline_of(rs, prec, classes, ms_per_step, counters=None):
    '''The roofline object of the dominant kernel (largest share of kernel time) of one radial step from the
    per-class event totals, with the counted HBM traffic and MFMA busy fraction from the committed counter summary
    of the same kernel and workload when there is one.'''
    el = 8 if prec in ('fp64', 'fp64-plain') else 4
    timed = {k: v for k, v in classes.items() if v[1] > 0 and k in KERNEL_NAMES}
    dom = max(timed, key=lambda k: timed[k][0])
    ms, cnt = timed[dom]
    avg_s = ms / cnt * 1e-3
    B, F = kernel_work(rs, dom, el, prec)
    achieved = B / avg_s / 1e9
    traffic = None
    extra = {}
    if counters and prec in counters and dom in counters[prec]:
        c = counters[prec][dom]
        traffic = c.get('bytes_per_launch')
        if 'mfma_busy_cycles_per_launch' in c:
            extra['mfma_busy_frac'] = c['mfma_busy_cycles_per_launch'] / (counters['simds'] * avg_s *
                                                                           counters['clock_mhz'] * 1e6)
            extra['mfma_flop_per_launch_counted'] = c.get('mfma_flop_per_launch')
    if F:
        pk = PEAK_TFLOPS['fp64' if el == 8 else ('bf16x3' if prec == 'bf16x3' else 'fp32')]
        extra.update(mfma_achieved_tflops=F / avg_s / 1e12, mfma_peak_tflops=pk, mfma_frac=F / avg_s / 1e12 / pk,
                     flop_per_launch=F)
    total = sum(v[0] for v in classes.values())
    per_kernel = {}
    for k, (kms, kn) in classes.items():
        if kn == 0:
            continue
        b, _ = kernel_work(rs, k, el, prec)
        a = kms / kn * 1e-3
        per_kernel[k] = dict(avg_launch_us=a * 1e6, launches=kn, share=round(kms / total, 4),
                             achieved_gbs=(b / a / 1e9) if b else None)
    return dict(bound='hbm', achieved=achieved, peak=HBM_PEAK_GBS, unit='GB/s', frac=achieved / HBM_PEAK_GBS,
                traffic=traffic, kernel=KERNEL_NAMES[dom], kernel_class=dom, avg_launch_us=avg_s * 1e6,
                launches=cnt, bytes_per_launch=B,
                step_hbm_frac=step_bytes(rs, el) / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                per_kernel=per_kernel, **extra)


def pcie_inclusive(rs, prec, K):
    '''Grid-pt z-steps/s with the host field upload (set_field) and the finish outputs (host arrays) inside the
    timed region.'''
    st = rs.make_stepper(uppe_kernel.hip_stepper_factory(prec))
    zi, zf, L = rs.plane(0)
    run_steps(rs, st, 2)
    st.synchronize()
    t0 = time.perf_counter()
    st.set_field(rs.stepper_field())
    n, _ = st.advance(zi, L, 0, rs.subcycles, rs.dzmin, max_steps=K)
    st.finish(zf - zi)
    t1 = time.perf_counter()
    st.close()
    return rs.N[0] * rs.N[1] * n / (t1 - t0)


CLASSES = ('hankel', 'hankel_r2k', 'source', 'point', 'reduce', 'transverse_fft')


def radial_measure(vol, prec, K, W, plain_steps, compare=True, counters=None, error_steps=None, module=uppe_kernel):
    '''One radial workload on one handle: K timed steps after W warm-up steps (inputs resident in HBM), the same K
    steps again with dispatch-stamped HIP events on every launch (per-kernel averages and the roofline of the
    dominant kernel), then the same-precision fused FP64 and the FP32 steppers over the same K steps, the plain FP64
    stepper over plain_steps steps, and the field error of prec (and FP32) against fused FP64 on FP64's step schedule
    after error_steps steps (default K). module: uppe_kernel, or paraxial_kernel for the envelope model.'''
    rs = module.RunSetup(vol)
    pts = rs.N[0] * rs.N[1]
    st = rs.make_stepper(uppe_kernel.hip_stepper_factory(prec))
    if W > 0:
        run_steps(rs, st, W)
    n, secs, _ = run_steps(rs, st, K)
    info = st.info()
    st.profile_enable(True)
    n2, secs2, _ = run_steps(rs, st, K)
    classes = {kc: st.profile_read(kc) for kc in CLASSES}
    st.profile_enable(False)
    st.close()
    ms_per_step = 1e3 * secs / n
    out = dict(value=pts * n / secs, steps=n, ms_per_step=ms_per_step, nan_steps=info['nan_steps'],
               ms_per_step_with_events=1e3 * secs2 / n2,
               roofline=roofline_of(rs, prec, classes, ms_per_step, counters))
    if not compare:
        return rs, out
    for p in ('fp64', 'fp32'):
        if p == prec:
            continue
        sx = rs.make_stepper(uppe_kernel.hip_stepper_factory(p))
        run_steps(rs, sx, max(W, 1))
        nx, tx, _ = run_steps(rs, sx, K)
        sx.close()
        out['%s_fused_pt_steps_per_s' % p] = pts * nx / tx
    sp = rs.make_stepper(uppe_kernel.hip_stepper_factory('fp64-plain'))
    run_steps(rs, sp, 1)
    npl, tpl, _ = run_steps(rs, sp, min(plain_steps, K))
    sp.close()
    out['fp64_plain_pt_steps_per_s'] = pts * npl / tpl
    out['speedup_vs_fp64_plain'] = out['value'] / out['fp64_plain_pt_steps_per_s']
    out['fp64_fused_speedup_vs_fp64_plain'] = out['fp64_fused_pt_steps_per_s'] / out['fp64_plain_pt_steps_per_s']
    E = error_steps or K
    s64 = rs.make_stepper(uppe_kernel.hip_stepper_factory('fp64'))
    _, _, sched = run_steps(rs, s64, E, record=True)
    q64 = s64.get_state()
    s64.close()
    for p in sorted({prec, 'fp32'}):
        sx = rs.make_stepper(uppe_kernel.hip_stepper_factory(p))
        run_steps(rs, sx, E, schedule=sched)
        qx = sx.get_state()
        sx.close()
        out['rel_l2_err_vs_fp64' if p == prec else 'fp32_rel_l2_err_vs_fp64'] = float(
            np.linalg.norm(qx - q64) / np.linalg.norm(q64))
    out['error_note'] = 'k-space state after %d steps on the FP64 step schedule' % E
    return rs, out


def ionizing_adjoint_workload():
    '''The radial O2 photoionisation workload (Nw = 2049, Nr = 256, Kerr + PPT rate + plasma current) over its
    first 4 mm in 100 fixed steps of 40 um (the adjoint holds the step sizes constant; the adaptive control starts
    this workload at 46 um, and fixed 100 um steps overflow the field).'''
    vol = configs.radial_o2_ionization(steps=100, length='4 mm')
    vol['minimum step'] = vol['size'][2] / 100
    return vol


def device_ms(st, fn):
    '''Milliseconds of kernel execution on the handle's stream during fn() (dispatch-stamped HIP events on every
    launch, summed), so that host work and transfers (set_field's upload, finish's downloads) are not counted.'''
    st.profile_enable(True)
    fn()
    total = sum(st.profile_read(kc)[0] for kc in CLASSES)
    st.profile_enable(False)
    return total


def adjoint_cost(prec='fp32', reps=3, workload='kerr', fd=False):
    '''The gradient of the on-axis output-spectrum objective with respect to 32 Legendre spectral-phase
    coefficients, by the adjoint with checkpoints every 10 steps (and with every stage kept, 100), against a plain
    forward run of the same objective: wall time and device (kernel) time of both, and the gradient's relative L2
    error against the FP64 adjoint. With fd, also against central finite differences (h = 1e-4) of the FP64 forward
    on all 32 coefficients. Workloads: the radial Kerr workload (Nw = 513, Nr = 128, 100 steps), or the
    photoionisation workload (ionizing_adjoint_workload).'''
    from searay_amd import adjoint
    if workload == 'kerr':
        vol, sig = configs.radial_kerr_air(), configs.kerr_pulse_sigma()
        name = 'radial Kerr Nw=513 Nr=128, 100 steps'
    else:
        vol, sig = ionizing_adjoint_workload(), configs.ionizing_pulse_sigma()
        name = 'radial Kerr + O2 photoionisation + plasma current, Nw=2049 Nr=256, 100 fixed steps over 4 mm'
    c = 0.05 * np.random.default_rng(3).standard_normal(32)
    p64 = adjoint.SpectralPhaseProblem(vol, sig, precision='fp64', checkpoint_every=10)
    _, g64 = p64.value_and_grad(c)
    out = dict(workload='gradient of the on-axis output spectrum objective, 32 Legendre phase coefficients, ' + name,
               precision=prec)
    if fd:
        h = 1e-4
        gfd = np.zeros(32)
        for i in range(32):
            e = np.zeros(32)
            e[i] = h
            gfd[i] = (p64.loss(c + e) - p64.loss(c - e)) / (2 * h)
        out['fd'] = dict(step=h, fp64_adjoint_rel_l2_err_vs_fd=float(np.linalg.norm(g64 - gfd) / np.linalg.norm(gfd)))
    p64.close()
    for every in (10, 100):
        p = adjoint.SpectralPhaseProblem(vol, sig, precision=prec, checkpoint_every=every, target=p64.target)
        p.loss(c)
        p.value_and_grad(c)
        t0 = time.perf_counter()
        for _ in range(reps):
            p.loss(c)
        tf = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            _, g = p.value_and_grad(c)
        tg = (time.perf_counter() - t0) / reps
        # device times: the median of reps measurements each (one summed-event measurement scatters by a few %)
        df = float(np.median([device_ms(p.st, lambda: p.loss(c)) for _ in range(reps)]))
        dg = float(np.median([device_ms(p.st, lambda: p.value_and_grad(c)) for _ in range(reps)]))
        p.close()
        rec = dict(forward_ms=tf * 1e3, forward_and_gradient_ms=tg * 1e3, gradient_over_forward=tg / tf,
                   forward_device_ms=df, forward_and_gradient_device_ms=dg, gradient_over_forward_device=dg / df,
                   grad_rel_l2_err_vs_fp64=float(np.linalg.norm(g - g64) / np.linalg.norm(g64)))
        if fd and every == 10:
            out['fd']['grad_rel_l2_err_vs_fd'] = float(np.linalg.norm(g - gfd) / np.linalg.norm(gfd))
        out['checkpoint_every_%d' % every] = rec
    return out


def batch_workload(prec, K, W, seq_steps, graph_steps=8):
    '''The 64-instance batched workload (configs.batched_radial): every instance stepped K fixed steps in one
    batch (one launch per kernel for each run of equally shaped instances, autotuned launch shapes), against the
    same instances stepped one after another as standalone handles (seq_steps steps each).'''
    from searay_amd import batch
    vols, drawn = configs.batched_radial(n=64, steps=200)
    pts = sum(v['wave grid'][0] * v['wave grid'][1] for v in vols)
    b = batch.Batch(vols, precision=prec, graph_steps=graph_steps)
    b.set_fields()
    log('batch warmup %d steps (autotune)' % W)
    b.advance(max_steps=max(W, 1))
    b.set_fields()
    b.synchronize()
    log('batch timed region: %d steps of 64 instances' % K)
    t0 = time.perf_counter()
    steps = b.advance(max_steps=K)
    b.synchronize()
    secs = time.perf_counter() - t0
    plans = {}
    for i, d in enumerate(drawn):
        plans.setdefault(d[0], b.plan(i))
    nan = sum(m.info()['nan_steps'] for m in b.members)
    finite = all(bool(np.all(np.isfinite(m.get_state()))) for m in b.members)
    b.close()
    n = int(steps.min())
    seq_value = None
    if seq_steps > 0:
        log('sequential standalone handles, %d steps each' % seq_steps)
        tseq = 0.0
        for v in vols:
            rs = uppe_kernel.RunSetup(v)
            st = rs.make_stepper(uppe_kernel.hip_stepper_factory(prec))
            run_steps(rs, st, 1)
            _, dt, _ = run_steps(rs, st, seq_steps)
            tseq += dt
            st.close()
        seq_value = pts * seq_steps / tseq
    return dict(value=pts * n / secs, steps=n, secs=secs, pts=pts, nan_steps=nan, finite=finite, graph_steps=graph_steps,
                grids=sorted({d[0] for d in drawn}), plans={str(k): v for k, v in sorted(plans.items())},
                sequential_pt_steps_per_s=seq_value,
                speedup_vs_sequential=(pts * n / secs / seq_value) if seq_value else None,
                instances=[dict(nr=d[0], power_pcr=round(d[1], 4), gas=d[2]) for d in drawn])


def cartesian_workload(prec, K, W, compare=True, cpu=False):
    '''The 3-D Cartesian filamentation workload (Nw = 513, 256 x 256 over 4 mm x 4 mm, 20 P_cr super-Gaussian
    beam with 10 % amplitude noise, Kerr + O2 photoionisation, 200 fixed steps over 10 cm): K timed steps on one
    handle, per-kernel-class averages with HIP events and the roofline of the transverse FFT pass (28 B per grid
    point per launch in FP32: 40 B on the passes with the stage prologue or epilogue, 16 B on the plain ones), the
    plain FP64 stepper and the fused FP64 stepper (whose state after 3 steps gives the error), and with cpu the FP64
    oracle on one step of the host cores.'''
    vol = configs.cartesian_filamentation()
    rs = uppe_kernel.RunSetup(vol)
    pts = rs.N[0] * rs.N[1] * rs.N[2]
    st = rs.make_stepper(uppe_kernel.hip_stepper_factory(prec))
    if W > 0:
        run_steps(rs, st, W)
    log('cartesian timed region: %d steps (%s)' % (K, prec))
    n, secs, _ = run_steps(rs, st, K)
    st.profile_enable(True)
    run_steps(rs, st, min(K, 10))
    classes = {kc: st.profile_read(kc) for kc in ('transverse_fft', 'source', 'reduce')}
    st.profile_enable(False)
    st.close()
    el = 8 if prec.startswith('fp64') else 4
    ms, cnt = classes['transverse_fft']
    avg_s = ms / max(cnt, 1) * 1e-3
    B = 28.0 * (el / 4) * pts
    out = dict(value=pts * n / secs, steps=n, ms_per_step=1e3 * secs / n, precision=prec,
               workload='3-D Cartesian filamentation, Nw=513, 256x256 over 4 mm, Kerr + O2, 200 fixed steps over '
                        '10 cm; first %d steps' % n,
               roofline=dict(bound='hbm', kernel='f_tfft (transverse FFT pass)', achieved=B / avg_s / 1e9,
                             peak=HBM_PEAK_GBS, unit='GB/s', frac=B / avg_s / 1e9 / HBM_PEAK_GBS, traffic=None,
                             avg_launch_us=avg_s * 1e6, launches=cnt, bytes_per_launch=B),
               class_avg_us={k: (v[0] / v[1] * 1e3 if v[1] else None) for k, v in classes.items()})
    if compare and prec != 'fp64-plain':
        for p, k in (('fp64-plain', 3), ('fp64', min(K, 10))):
            sx = rs.make_stepper(uppe_kernel.hip_stepper_factory(p))
            run_steps(rs, sx, 1)
            nx, tx, _ = run_steps(rs, sx, k)
            sx.close()
            out['%s_pt_steps_per_s' % p.replace('-', '_')] = pts * nx / tx
        out['speedup_vs_fp64_plain'] = out['value'] / out['fp64_plain_pt_steps_per_s']
        qs = {}
        for p in ('fp64', prec):
            sx = rs.make_stepper(uppe_kernel.hip_stepper_factory(p))
            run_steps(rs, sx, 3)
            qs[p] = sx.get_state()
            sx.close()
        out['rel_l2_err_vs_fp64'] = float(np.linalg.norm(qs[prec] - qs['fp64']) / np.linalg.norm(qs['fp64']))
        out['error_note'] = 'k-space state after 3 fixed steps'
    if cpu:
        log('CPU baseline: FP64 oracle, 1 step')
        cv, cn, ct_, threads = cpu_baseline(uppe_kernel, vol, 1)
        out['cpu_baseline'] = dict(value=cv, unit='grid-pt z-steps/s', cores=threads, kind='port',
                                   sample='%d RK4 step of this workload (513 x 256 x 256), FP64, %.1f s' % (cn, ct_))
    return out


def cpu_baseline(module, vol, steps):
    from oracle import oracle
    rs = module.RunSetup(vol)
    st = rs.make_stepper(oracle.oracle_factory)
    zi, zf, L = rs.plane(0)
    st.set_field(rs.stepper_field())
    st.advance(zi, L, 0, rs.subcycles, rs.dzmin, max_steps=1)
    st.set_field(rs.stepper_field())
    t0 = time.perf_counter()
    n, _ = st.advance(zi, L, 0, rs.subcycles, rs.dzmin, max_steps=steps)
    dt = time.perf_counter() - t0
    st.close()
    return rs.N[0] * rs.N[1] * rs.N[2] * n / dt, n, dt, oracle.num_threads()


def headline_line(metric, r, K, W, prec, workload, extra_config, cpu, data):
    return {'metric': metric, 'value': r['value'], 'unit': 'grid-pt z-steps/s', 'n_gpus': 1, 'steps': r['steps'],
            'warmup': W, 'ms_per_step': r['ms_per_step'], 'higher_is_better': True, 'scaling': 'weak',
            'vs_baseline': None, 'dtype': DTYPE[prec], 'data': data,
            'config': dict(workload=workload, precision=prec, **extra_config), 'roofline': r['roofline'],
            'cpu_baseline': cpu}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=100)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--precision', default=None, choices=['bf16x3', 'fp32', 'fp64', 'fp64-plain'],
                    help='default: bf16x3 (fp32 for --workload adjoint and cartesian)')
    ap.add_argument('--rate', default='explicit', choices=['explicit', 'series'])
    ap.add_argument('--plain-steps', type=int, default=20)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-compare', action='store_true')
    ap.add_argument('--workload', default='all',
                    choices=['all', 'kerr', 'o2', 'uppe', 'paraxial', 'adjoint', 'batch', 'cartesian'])
    ap.add_argument('--calibrate-copy', type=int, default=0, metavar='MIB',
                    help='only run a plain device-to-device copy of MIB MiB (for scaling HBM byte counters)')
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log('copied %d MiB three times (read %d MiB + write %d MiB per copy)' % (args.calibrate_copy,
                                                                              args.calibrate_copy,
                                                                              args.calibrate_copy))
        return
    if args.gpus != 1:
        print('bench.py runs on one GPU only: this stepper has no multi-GPU path (use --gpus 1)', file=sys.stderr)
        sys.exit(2)
    K, W = args.steps, args.warmup
    compare = not args.no_compare

    if args.workload == 'adjoint':
        prec = args.precision or 'fp32'
        if prec == 'fp64-plain':
            sys.exit('the adjoint runs on the fused paths only')
        log('adjoint gradient cost (%s), radial Kerr 513 x 128 and radial O2 2049 x 256' % prec)
        print(json.dumps(dict(kerr=adjoint_cost(prec, fd=compare), ionizing=adjoint_cost(prec, workload='ionizing'))))
        return

    if args.workload == 'cartesian':
        prec = args.precision or 'fp32'
        r = cartesian_workload(prec, min(K, 40), min(W, 2), compare, cpu=not args.no_cpu)
        line = headline_line('UPPE grid-pt z-steps/s (1 GPU), 3-D Cartesian', r, K, min(W, 2), prec, r['workload'],
                             {}, r.pop('cpu_baseline', None), 'synthetic (seeded super-Gaussian beam, 10% amplitude '
                             'noise)')
        line.update({k: v for k, v in r.items() if k not in line and k not in ('workload', 'precision')})
        print(json.dumps(line))
        return

    if args.workload == 'batch':
        prec = args.precision or 'bf16x3'
        if prec == 'fp64-plain':
            sys.exit('batches run on the fused paths only')
        r = batch_workload(prec, min(K, 200), W, seq_steps=0 if args.no_compare else args.plain_steps)
        log('%d steps in %.3f s: %.4e pt-steps/s' % (r['steps'], r['secs'], r['value']) +
            ('' if r['speedup_vs_sequential'] is None else
             ' (%.2fx the sequential handles)' % r['speedup_vs_sequential']))
        print(json.dumps(dict(metric='batched radial UPPE grid-pt z-steps/s (1 GPU)', value=r['value'],
                              unit='grid-pt z-steps/s', n_gpus=1, steps=r['steps'], warmup=W,
                              ms_per_step=1e3 * r['secs'] / r['steps'], higher_is_better=True, scaling='weak',
                              vs_baseline=None, dtype=DTYPE[prec], data='synthetic (seeded instances)',
                              config=dict(workload='64 radial instances, Nw=1025, Nr in {64,128,256,512}, O2/Ar, '
                                                   '0.1-5 P_cr, 200 fixed steps over 5 mm', precision=prec),
                              batch=r)))
        return

    prec = args.precision or 'bf16x3'
    data = 'synthetic (seeded Gaussian pulse with 1% spectral-phase noise)'
    if args.workload == 'paraxial':
        vol = configs.paraxial_air_filament(steps=400, rate=args.rate)
        log('paraxial workload: %d timed steps (%s)' % (K, prec))
        rs, r = radial_measure(vol, prec, K, W, args.plain_steps, compare, module=paraxial_kernel)
        cpu = None
        if not args.no_cpu:
            cv, cn, ct_, threads = cpu_baseline(paraxial_kernel, vol, 60)
            cpu = dict(value=cv, unit='grid-pt z-steps/s', cores=threads, kind='port',
                       sample='%d RK4 steps of this workload (%d x %d), FP64, %.1f s' % (cn, rs.N[0], rs.N[1], ct_))
        wl = ('radial paraxial envelope, focused air filament: Kerr + O2 photoionisation (cycle-averaged rate) + plasma '
              'current, Nw=256, Nr=256 with 128 Hankel modes, 400 adaptive subcycles in the first 2 cm plane; first %d '
              'RK4 steps' % r['steps'])
        line = headline_line(METRIC_PARAXIAL, r, K, W, prec, wl, {'grid': [rs.N[0], rs.N[1]]}, cpu, data)
        line.update({k: v for k, v in r.items() if k not in line})
        print(json.dumps(line))
        return

    if args.workload in ('o2', 'uppe'):
        vol = configs.radial_o2_ionization(rate=args.rate)
        log('radial O2 workload (Nw=2049, Nr=256): %d timed steps (%s)' % (K, prec))
        rs, r = radial_measure(vol, prec, K, W, args.plain_steps, compare,
                               counters=measured_counters('counters_radial_o2_2049x256'))
        cpu = None
        if not args.no_cpu:
            cv, cn, ct_, threads = cpu_baseline(uppe_kernel, vol, 60)
            cpu = dict(value=cv, unit='grid-pt z-steps/s', cores=threads, kind='port',
                       sample='%d RK4 steps of this workload (2049 x 256), FP64, %.1f s' % (cn, ct_))
        line = headline_line(METRIC, r, K, W, prec, O2_WORKLOAD % r['steps'],
                             {'rate_model': args.rate, 'grid': [rs.N[0], rs.N[1]]}, cpu, data)
        line.update({k: v for k, v in r.items() if k not in line})
        print(json.dumps(line))
        return

    # the headline: the radial Kerr workload the north-star target is quoted on (Nw = 513, Nr = 128, 100 fixed
    # steps over 1 m), then the other workloads as extras
    vol = configs.radial_kerr_air()
    log('radial Kerr workload (Nw=513, Nr=128): %d timed steps (%s)' % (K, prec))
    rs, r = radial_measure(vol, prec, K, W, args.plain_steps, compare,
                           counters=measured_counters('counters_radial_kerr_513x128'), error_steps=100)
    log('%d steps: %.4f ms/step, %.4e pt-steps/s' % (r['steps'], r['ms_per_step'], r['value']))
    cpu = None
    if not args.no_cpu:
        log('CPU baseline: FP64 oracle, 100 steps of the radial Kerr workload')
        cv, cn, ct_, threads = cpu_baseline(uppe_kernel, vol, 100)
        cpu = dict(value=cv, unit='grid-pt z-steps/s', cores=threads, kind='port',
                   sample='%d RK4 steps of this workload (513 x 128), the whole run, FP64, %.1f s' % (cn, ct_))
    line = headline_line(METRIC, r, K, W, prec, KERR_WORKLOAD % r['steps'], {'grid': [rs.N[0], rs.N[1]]}, cpu,
                         data)
    line.update({k: v for k, v in r.items() if k not in line})
    if args.workload == 'all' and compare:
        Ko = min(K, 100)
        log('radial O2 workload (Nw=2049, Nr=256): %d steps (%s)' % (Ko, prec))
        vol2 = configs.radial_o2_ionization(rate=args.rate)
        rs2, r2 = radial_measure(vol2, prec, Ko, W, args.plain_steps, True,
                                 counters=measured_counters('counters_radial_o2_2049x256'))
        r2['workload'] = O2_WORKLOAD % r2['steps']
        log('host-inclusive rate (set_field + %d steps + finish)' % Ko)
        r2['pcie_inclusive_pt_steps_per_s'] = pcie_inclusive(rs2, prec, Ko)
        line['radial_o2_2049x256'] = r2
        log('Cartesian workload (513 x 256 x 256): 10 steps (fp32)')
        line['cartesian_513x256x256'] = cartesian_workload('fp32', 10, 1, True, cpu=False)
        log('adjoint gradient cost, radial Kerr 513 x 128 (with finite differences)')
        line['adjoint'] = adjoint_cost(fd=True)
        log('adjoint gradient cost, radial O2 photoionisation 2049 x 256')
        line['adjoint_ionizing'] = adjoint_cost(workload='ionizing')
        log('batched workload: 64 radial instances (%s)' % prec)
        bw = batch_workload(prec, 200, 2, seq_steps=10)
        bw.pop('instances')
        line['batched_64'] = bw
    print(json.dumps(line))


KERR_WORKLOAD = ('radial UPPE, Kerr in air, 800 nm 50 fs pulse at 0.5 P_cr, Nw=513, Nr=128, 100 fixed steps over '
                 '1 m; first %d RK4 steps')
O2_WORKLOAD = ('radial UPPE, Kerr + O2 PPT photoionisation + plasma current, Nw=2049, Nr=256, 400 adaptive '
               'subcycles over 20 cm; first %d RK4 steps')


if __name__ == '__main__':
    main()
