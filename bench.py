'''Benchmark of the MI355X UPPE stepper on the radial photoionisation workload.

Workload: radial UPPE with Kerr, O2 PPT photoionisation (Ip = 12.07 eV, 5.4e18 cm^-3) and the plasma current;
800 nm, 40 fs, w0 = 100 um, 3 P_cr; Nw = 2049 frequencies over (0, 8), Nr = 256 Hankel modes; 400 adaptive
subcycles over 20 cm. A "step" is one adaptive RK4 z-step (four right-hand-side evaluations) over the whole
grid; the timed region runs the first K steps of that run from the initial field, on inputs resident in HBM.

Printed JSON line (one):
  value         grid-point z-steps per second = Nw * Nr * K / T
  roofline      the dominant kernel class (largest share of kernel time) against HBM bandwidth: algorithmic
                bytes per launch (per grid point: 40 B for a radial transform launch -- the k->r pass, 16 B, and
                the r->k pass with the RK stage, 40 B, plus 24 B of kz tables, averaged over the two -- and 16 B
                for the source pass; doubled in FP64) over the average launch duration. Durations come from HIP
                events recorded on the library's stream around every launch, in a repeat of the timed steps (the
                events lengthen the step, so `value` is taken without them). The transform is also reported
                against the matrix-core peak of the type it issues. `traffic` is the measured HBM bytes per
                launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated) read from the committed profile
                summary of the same kernel and workload, or null. `mfma_busy_frac` is the committed
                rocprofv3 SQ_VALU_MFMA_BUSY_CYCLES per transform launch over (SIMDs x engine clock x this run's
                average launch duration); `mfma_flop_per_launch_counted` the counted MFMA flops per launch
  cpu_baseline  the FP64 oracle (reference algorithm and pass structure, OpenMP) on a bounded sample
  extras        speedup over the plain FP64 GPU stepper, relative L2 error of the state against the fused FP64
                stepper on the same step schedule; the same on the radial Kerr workload (Nw = 513, Nr = 128,
                100 fixed steps); the cost of the adjoint gradient of an output-spectrum objective on that workload
                against a forward run, with checkpoints every 10 steps and with every stage kept, and its error
                against the FP64 adjoint, also on the photoionisation workload over its first 4 mm in 100
                fixed steps (adjoint_ionizing); the 64-instance batched workload (Nw = 1025, Nr in {64, 128, 256, 512},
                O2 / Ar, 0.1-5 P_cr, 200 fixed steps; one launch per kernel for each group of equally shaped
                instances, autotuned launch shapes, HIP graphs) against the same instances run one after another

``--workload batch`` runs only the batched workload (its JSON line carries the drawn instances).
``--workload kerr`` runs only the radial Kerr workload (Nw = 513, Nr = 128, 100 fixed steps) against the plain FP64
stepper, as the bench line's `radial_kerr_513x128`.
``--workload cartesian`` measures the 3-D Cartesian filamentation workload (Nw = 513, 256 x 256, Kerr + O2) the
same way, its roofline on the transverse FFT pass (28 B per grid point per launch in FP32: 40 B on the passes
with the stage prologue or epilogue, 16 B on the plain ones).

``--workload adjoint`` runs only that adjoint measurement (forward and gradient sweeps, FP32 by default), for
kernel traces of the adjoint.

``--workload paraxial`` measures the paraxial envelope stepper the same way on the focused air filament of the
reference's paraxial example (Nw = 256 envelope frequencies, Nr = 256 cells with 128 Hankel modes, Kerr + O2
photoionisation with the cycle-averaged rate) with 400 subcycles in the first 2 cm plane.
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


def hankel_bytes(rs, el):
    '''Algorithmic bytes of one radial transform launch, averaged over k->r and r->k: 40 B per grid point in
    FP32 (16 + 40 + 24 over two launches), twice that in FP64.'''
    return 40.0 * (el / 4) * rs.N[0] * rs.N[1]


def step_bytes(rs, el):
    '''Algorithmic bytes of one RK4 step: 4 right-hand sides x 96 B per grid point in FP32.'''
    return 384.0 * (el / 4) * rs.N[0] * rs.N[1]


def measured_mfma(prec):
    '''MFMA busy cycles and counted MFMA flops per radial-transform launch (the average of a k->r and an r->k
    launch), the SIMD count and the engine clock, from the committed rocprofv3 counter summary of the radial O2
    workload, if present.'''
    path = os.path.join(ROOT, 'profiles', 'mfma_radial_o2_2049x256.json')
    try:
        with open(path) as f:
            tab = json.load(f)
        return tab[prec]['hankel'], tab['simds'], tab['clock_mhz']
    except (OSError, KeyError, ValueError):
        return None


def measured_traffic(prec, kernel, workload='uppe'):
    '''HBM bytes per launch from the committed rocprofv3 counter summary of this workload, if present.'''
    name = 'traffic_radial_o2_2049x256.json' if workload == 'uppe' else 'traffic_paraxial_fil_256x256.json'
    path = os.path.join(ROOT, 'profiles', name)
    try:
        with open(path) as f:
            tab = json.load(f)
        return tab[prec][kernel]['bytes_per_launch']
    except (OSError, KeyError, ValueError):
        return None


def hankel_flops(rs):
    '''Algorithmic flops of one radial transform launch: 4 Nw Nk Nr (real contraction of complex data).'''
    T = rs.field_tool.GetTransverseTool()
    return 4.0 * rs.N[0] * T.Nk * rs.N[1]


def source_bytes(rs, el):
    '''Algorithmic bytes of one source launch: read E_r and write J_r (Nw complex per strip).'''
    return 2.0 * rs.N[0] * rs.N[1] * 2 * el  # 16 B per grid point in FP32


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


def kerr_513x128_speedup(prec, steps_plain):
    '''Radial Kerr workload (Nw = 513, Nr = 128, 100 fixed steps over 1 m): the fused stepper against the plain
    FP64 stepper, and its field error against the fused FP64 stepper (fixed steps: identical schedules).'''
    vol = configs.radial_kerr_air()
    rs = uppe_kernel.RunSetup(vol)
    pts = rs.N[0] * rs.N[1]
    out = {}
    states = {}
    for p, K in ((prec, rs.subcycles), ('fp64', rs.subcycles), ('fp64-plain', steps_plain)):
        st = rs.make_stepper(uppe_kernel.hip_stepper_factory(p))
        run_steps(rs, st, 2)
        n, secs, _ = run_steps(rs, st, K)
        out[p] = pts * n / secs
        if p != 'fp64-plain':
            states[p] = st.get_state()
        st.close()
    err = float(np.linalg.norm(states[prec] - states['fp64']) / np.linalg.norm(states['fp64']))
    return dict(workload='radial Kerr in air, Nw=513, Nr=128, 100 fixed steps over 1 m',
                pt_steps_per_s=out[prec], fp64_plain_pt_steps_per_s=out['fp64-plain'],
                speedup_vs_fp64_plain=out[prec] / out['fp64-plain'], rel_l2_err_vs_fp64=err)


def ionizing_adjoint_workload():
    '''The radial O2 photoionisation workload (Nw = 2049, Nr = 256, Kerr + PPT rate + plasma current) over its
    first 4 mm in 100 fixed steps of 40 um (the adjoint holds the step sizes constant; the adaptive control starts
    this workload at 46 um, and fixed 100 um steps overflow the field).'''
    vol = configs.radial_o2_ionization(steps=100, length='4 mm')
    vol['minimum step'] = vol['size'][2] / 100
    return vol


def adjoint_cost(prec='fp32', reps=3, workload='kerr'):
    '''The gradient of the on-axis output-spectrum objective with respect to 32 Legendre spectral-phase
    coefficients, by the adjoint with checkpoints every 10 steps (and with every stage kept, 100), against a plain
    forward run of the same objective; and the gradient's relative L2 error against the FP64 adjoint. Workloads:
    the radial Kerr workload (Nw = 513, Nr = 128, 100 steps), or the photoionisation workload
    (ionizing_adjoint_workload).'''
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
    p64.close()
    out = dict(workload='gradient of the on-axis output spectrum objective, 32 Legendre phase coefficients, ' + name,
               precision=prec)
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
        p.close()
        out['checkpoint_every_%d' % every] = dict(forward_ms=tf * 1e3, forward_and_gradient_ms=tg * 1e3,
                                                  gradient_over_forward=tg / tf,
                                                  grad_rel_l2_err_vs_fp64=float(np.linalg.norm(g - g64) /
                                                                                np.linalg.norm(g64)))
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


def cartesian_workload(args):
    '''The 3-D Cartesian filamentation workload (Nw = 513, 256 x 256 over 4 mm x 4 mm, 20 P_cr super-Gaussian
    beam with 10 % amplitude noise, Kerr + O2 photoionisation, 200 fixed steps over 10 cm): K timed steps on one
    handle, per-kernel-class averages with HIP events, the plain FP64 stepper and the fused FP64 stepper (whose
    state after 3 steps gives the error), and the FP64 oracle on one step of the host cores.'''
    prec = args.precision or 'fp32'
    vol = configs.cartesian_filamentation()
    rs = uppe_kernel.RunSetup(vol)
    pts = rs.N[0] * rs.N[1] * rs.N[2]
    K, W = min(args.steps, 40), min(args.warmup, 2)
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
    # per grid point: 40 B on the passes with the stage prologue or epilogue, 16 B on the plain passes; averaged
    B = 28.0 * (el / 4) * pts
    out = dict(metric='UPPE grid-pt z-steps/s (1 GPU), 3-D Cartesian', value=pts * n / secs, unit='grid-pt z-steps/s',
               n_gpus=1, steps=n, warmup=W, ms_per_step=1e3 * secs / n, higher_is_better=True, scaling='weak',
               vs_baseline=None, dtype=DTYPE[prec], data='synthetic (seeded super-Gaussian beam, 10% amplitude noise)',
               config=dict(workload='3-D Cartesian filamentation, Nw=513, 256x256 over 4 mm, Kerr + O2, 200 fixed '
                                    'steps over 10 cm; first %d steps' % n, precision=prec),
               roofline=dict(bound='hbm', kernel='transverse FFT pass (f_tfft)', achieved=B / avg_s / 1e9,
                             peak=HBM_PEAK_GBS, unit='GB/s', frac=B / avg_s / 1e9 / HBM_PEAK_GBS, traffic=None,
                             avg_launch_us=avg_s * 1e6, launches=cnt, bytes_per_launch=B),
               class_avg_us={k: (v[0] / v[1] * 1e3 if v[1] else None) for k, v in classes.items()})
    if not args.no_compare and prec != 'fp64-plain':
        for p, k in (('fp64-plain', 3), ('fp64', 10)):
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
    if not args.no_cpu:
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=400)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--precision', default=None, choices=['bf16x3', 'fp32', 'fp64', 'fp64-plain'],
                    help='default: bf16x3 (fp32 for --workload adjoint)')
    ap.add_argument('--rate', default='explicit', choices=['explicit', 'series'])
    ap.add_argument('--plain-steps', type=int, default=20)
    ap.add_argument('--cpu-steps', type=int, default=60)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-compare', action='store_true')
    ap.add_argument('--workload', default='uppe', choices=['uppe', 'paraxial', 'adjoint', 'batch', 'cartesian', 'kerr'])
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

    if args.workload == 'kerr':
        prec = args.precision or 'bf16x3'
        if prec == 'fp64-plain':
            sys.exit('the Kerr workload compares a fused path with the plain FP64 stepper')
        log('radial Kerr workload (Nw=513, Nr=128): %s vs plain FP64' % prec)
        print(json.dumps(kerr_513x128_speedup(prec, args.plain_steps)))
        return

    if args.workload == 'adjoint':
        prec = args.precision or 'fp32'
        if prec == 'fp64-plain':
            sys.exit('the adjoint runs on the fused paths only')
        log('adjoint gradient cost (%s), radial Kerr 513 x 128 and radial O2 2049 x 256' % prec)
        print(json.dumps(dict(kerr=adjoint_cost(prec), ionizing=adjoint_cost(prec, workload='ionizing'))))
        return

    if args.workload == 'cartesian':
        print(json.dumps(cartesian_workload(args)))
        return

    if args.workload == 'batch':
        prec = args.precision or 'bf16x3'
        if prec == 'fp64-plain':
            sys.exit('batches run on the fused paths only')
        K = min(args.steps, 200)
        r = batch_workload(prec, K, args.warmup, seq_steps=0 if args.no_compare else args.plain_steps)
        log('%d steps in %.3f s: %.4e pt-steps/s' % (r['steps'], r['secs'], r['value']) +
            ('' if r['speedup_vs_sequential'] is None else
             ' (%.2fx the sequential handles)' % r['speedup_vs_sequential']))
        print(json.dumps(dict(metric='batched radial UPPE grid-pt z-steps/s (1 GPU)', value=r['value'],
                              unit='grid-pt z-steps/s', n_gpus=1, steps=r['steps'], warmup=args.warmup,
                              ms_per_step=1e3 * r['secs'] / r['steps'], higher_is_better=True, scaling='weak',
                              vs_baseline=None, dtype=DTYPE[prec], data='synthetic (seeded instances)',
                              config=dict(workload='64 radial instances, Nw=1025, Nr in {64,128,256,512}, O2/Ar, '
                                                   '0.1-5 P_cr, 200 fixed steps over 5 mm', precision=prec),
                              batch=r)))
        return

    paraxial = args.workload == 'paraxial'
    if paraxial:
        vol = configs.paraxial_air_filament(steps=400, rate=args.rate)
        module = paraxial_kernel
    else:
        vol = configs.radial_o2_ionization(rate=args.rate)
        module = uppe_kernel
    rs = module.RunSetup(vol)
    pts = rs.N[0] * rs.N[1]
    K, W = args.steps, args.warmup
    prec = args.precision or 'bf16x3'

    st = rs.make_stepper(uppe_kernel.hip_stepper_factory(prec))
    log('warmup %d steps (%s)' % (W, prec))
    if W > 0:
        run_steps(rs, st, W)
    log('timed region: %d steps' % K)
    n, secs, _ = run_steps(rs, st, K)
    value = pts * n / secs
    ms_per_step = 1e3 * secs / n
    info = st.info()
    log('%d steps in %.3f s: %.3f ms/step, %.4e pt-steps/s, NaN steps %d' % (n, secs, ms_per_step, value,
                                                                           info['nan_steps']))
    log('same %d steps again with HIP events around every launch' % K)
    st.profile_enable(True)
    n2, secs2, _ = run_steps(rs, st, K)
    classes = {}
    for kc in ('hankel', 'source', 'point', 'reduce', 'transverse_fft'):
        ms, cnt = st.profile_read(kc)
        classes[kc] = (ms, cnt)
    st.profile_enable(False)
    st.close()

    dom = max(classes, key=lambda k: classes[k][0])
    ms, cnt = classes[dom]
    avg_s = ms / max(cnt, 1) * 1e-3
    el = 8 if prec in ('fp64', 'fp64-plain') else 4
    if dom == 'hankel':
        B = hankel_bytes(rs, el)
        F = hankel_flops(rs)
        pk = PEAK_TFLOPS['fp64' if el == 8 else ('bf16x3' if prec == 'bf16x3' else 'fp32')]
        issued = 3.0 if prec == 'bf16x3' else 1.0
        mfma = F * issued / avg_s / 1e12
        kname = 'hankel (radial transform: MFMA contraction with fused stage prologue / RK epilogue)'
        extra = dict(mfma_achieved_tflops=mfma, mfma_peak_tflops=pk, mfma_frac=mfma / pk,
                     flop_per_launch=F * issued)
        mm = measured_mfma(prec) if args.workload == 'uppe' else None
        if mm:
            # counter figures over this run's launch time: the r->k launch contracts only the nonlinear band,
            # so the counted flops fall below the algorithmic 4 Nw Nk Nr per launch
            h, simds, mhz = mm
            extra['mfma_flop_per_launch_counted'] = h['flop_per_launch']
            extra['mfma_busy_frac'] = h['busy_cycles_per_launch'] / (simds * avg_s * mhz * 1e6)
    else:
        B = source_bytes(rs, el)
        kname = dom
        extra = {}
    achieved = B / avg_s / 1e9
    roofline = dict(bound='hbm', achieved=achieved, peak=HBM_PEAK_GBS, unit='GB/s', frac=achieved / HBM_PEAK_GBS,
                    traffic=measured_traffic(prec, dom, args.workload), kernel=kname, avg_launch_us=avg_s * 1e6, launches=cnt,
                    bytes_per_launch=B, **extra)
    roofline['step_hbm_frac'] = step_bytes(rs, el) / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS
    share = {k: round(v[0] / max(1e-12, sum(c[0] for c in classes.values())), 4) for k, v in classes.items()
             if v[1] > 0}
    roofline['class_share'] = share

    extras = {}
    if not args.no_compare and prec != 'fp64-plain':
        log('plain FP64 stepper, %d steps' % args.plain_steps)
        sp = rs.make_stepper(uppe_kernel.hip_stepper_factory('fp64-plain'))
        npl, tpl, _ = run_steps(rs, sp, min(args.plain_steps, K))
        sp.close()
        plain_value = pts * npl / tpl
        extras['fp64_plain_pt_steps_per_s'] = plain_value
        extras['speedup_vs_fp64_plain'] = value / plain_value
        log('fused FP64 reference for the error, %d steps' % K)
        s64 = rs.make_stepper(uppe_kernel.hip_stepper_factory('fp64'))
        n64, t64, sched = run_steps(rs, s64, K, record=True)
        q64 = s64.get_state()
        s64.close()
        extras['fp64_fused_pt_steps_per_s'] = pts * n64 / t64
        sx = rs.make_stepper(uppe_kernel.hip_stepper_factory(prec))
        nx, tx, _ = run_steps(rs, sx, K, schedule=sched)
        qx = sx.get_state()
        sx.close()
        extras['rel_l2_err_vs_fp64'] = float(np.linalg.norm(qx - q64) / np.linalg.norm(q64))
        extras['error_note'] = 'k-space state after %d steps on the FP64 step schedule' % K
        log('host-inclusive rate (set_field + %d steps + finish)' % K)
        extras['pcie_inclusive_pt_steps_per_s'] = pcie_inclusive(rs, prec, K)
        if not paraxial:
            log('radial Kerr workload (Nw=513, Nr=128): %s vs plain FP64' % prec)
            extras['radial_kerr_513x128'] = kerr_513x128_speedup(prec, args.plain_steps)
            log('adjoint gradient cost, radial Kerr 513 x 128')
            extras['adjoint'] = adjoint_cost()
            log('adjoint gradient cost, radial O2 photoionisation 2049 x 256')
            extras['adjoint_ionizing'] = adjoint_cost(workload='ionizing')
            log('batched workload: 64 radial instances (%s)' % prec)
            bw = batch_workload(prec, 200, 2, seq_steps=10)
            bw.pop('instances')
            extras['batched_64'] = bw

    cpu = None
    if not args.no_cpu:
        log('CPU baseline: FP64 oracle, %d steps' % args.cpu_steps)
        cv, cn, ct_, threads = cpu_baseline(module, vol, args.cpu_steps)
        cpu = dict(value=cv, unit='grid-pt z-steps/s', cores=threads, kind='port',
                   sample='%d RK4 steps of this workload (%d x %d) from the initial field, FP64, %.1f s'
                   % (cn, rs.N[0], rs.N[1], ct_))

    if paraxial:
        workload = ('radial paraxial envelope, focused air filament: Kerr + O2 photoionisation (cycle-averaged rate) '
                    '+ plasma current, Nw=256, Nr=256 with 128 Hankel modes, 400 adaptive subcycles in the first '
                    '2 cm plane; first %d RK4 steps' % K)
    else:
        workload = ('radial UPPE, Kerr + O2 PPT photoionisation + plasma current, Nw=2049, Nr=256, '
                    '400 adaptive subcycles over 20 cm; first %d RK4 steps' % K)
    out = {
        'metric': METRIC_PARAXIAL if paraxial else METRIC, 'value': value, 'unit': 'grid-pt z-steps/s', 'n_gpus': 1, 'steps': n, 'warmup': W,
        'ms_per_step': ms_per_step, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': DTYPE[prec], 'data': 'synthetic (seeded Gaussian pulse with 1% spectral-phase noise)',
        'config': {'workload': workload,
                   'precision': prec, 'rate_model': args.rate, 'grid': [rs.N[0], rs.N[1]]},
        'roofline': roofline,
        'cpu_baseline': cpu,
        'ms_per_step_with_events': 1e3 * secs2 / n2,
        'nan_steps': info['nan_steps'],
    }
    out.update(extras)
    print(json.dumps(out))


if __name__ == '__main__':
    main()
