'''The FP64 oracle pinned against library routines, closed forms and the reference's own test.

Each check is chosen so that a plausible restatement mistake fails it:
* strip FFTs (fft.cl:66-280) vs numpy.fft - packing, sign and normalisation;
* linear propagation (uppe.cl:54-108 + grid_tools.py:218-325) vs an independent numpy evaluation of
  a(z) = T^-1 exp(i(kz+kg)z) T a(0), with an absorbing band so the decay sign is exercised;
* test/gaussian_test.py:110-155 restated (Cartesian and radial);
* RK4 stage weights, stage positions and conditioner (uppe_kernel.py:305-322) vs the textbook polynomial for
  dq/dz = lambda q, and the step-size rule (paraxial.cl:40-58, uppe_kernel.py:218-227) vs its closed form;
* the nonlinear current (uppe_kernel.py:139-186) in the time domain: Kerr J = -d/dtau (chi E) and plasma
  0.1 J - dJ/dtau = n_e E, with tau the (reversed) time index, checked by finite differences;
* the density scan (ionization.cl:127-155) vs its closed form for a constant rate, including the zeroed last
  sample and the soft time window, and the device rate formulas vs the host rate models;
* the UPPE right-hand side (E = i kz A, the Kerr current, S = (i/2) exp(+Im kz z) exp(-i(Re kz + kg) z) J / Re kz;
  uppe.cl:73-108) vs the closed-form plane-wave solutions of tests/plane_wave.py in BK7: planting E = i w A, a
  dropped 1/2, 1/w for 1/Re kz, a flipped exponent sign or exp(-Im kz z) in a copy of the oracle fails them.
'''
import numpy as np
import pytest

import plane_wave as pw
from conftest import rel_l2
from searay_amd import _abi, configs, grid_tools, ionization, uppe_kernel


class TestStripFFT:
    @pytest.mark.parametrize('n', [8, 16, 128, 1024, 4096])
    def test_against_numpy(self, oracle_mod, n):
        rng = np.random.default_rng(n)
        x = rng.standard_normal(n)
        X = np.fft.rfft(x)
        got = oracle_mod.rfft(x)
        assert rel_l2(got, X) < 1e-12
        assert got[0].imag == 0.0 and got[-1].imag == 0.0
        Y = X * np.exp(1j * rng.standard_normal(X.shape))
        assert rel_l2(oracle_mod.irfft(Y), np.fft.irfft(Y)) < 1e-12
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        assert rel_l2(oracle_mod.complex_fft(z, 1.0, -1.0), np.fft.fft(z)) < 1e-12
        assert rel_l2(oracle_mod.complex_fft(z, -1.0, -1.0), np.fft.ifft(z)) < 1e-12


def _stepper_for(vol):
    from oracle import oracle
    rs = uppe_kernel.RunSetup(vol)
    st = rs.make_stepper(oracle.oracle_factory)
    return rs, st


class TestLinearPropagation:
    @pytest.mark.parametrize('coords', ['cylindrical', 'cartesian'])
    def test_exact_linear_solution(self, oracle_mod, coords):
        if coords == 'cartesian':
            vol = configs.cartesian_filamentation(scale=16, steps=3, length='1 cm')
        else:
            vol = configs.radial_kerr_air(scale=8, steps=3)
        vol['chi3'] = 0.0
        vol.pop('ionizer', None)
        # absorbing band inside the grid so Im kz > 0 somewhere; the attenuation length is chosen so that
        # exp(+Im kz z) in the RHS stays finite (stronger absorption overflows to inf * 0 = NaN, as in the
        # reference's CurrentToODERHS)
        air = vol['dispersion inside']
        att = 1e6 if coords == 'cylindrical' else 1e5
        air.add_opacity_region(att, 2 * np.pi / 1.2, 2 * np.pi / 0.9)
        rec = []
        A, J, chi, ne, dom4d, evals = oracle_mod.oracle_track(vol, record=rec)
        rs = uppe_kernel.RunSetup(vol)
        zi, zf, L = rs.plane(0)
        rs.mat.UpdateMaterial(zi + 0.5 * L)
        assert np.any(np.imag(rs.mat.kz) > 0)
        phase = np.exp(1j * (rs.mat.kz + rs.mat.kg[:, None, None]) * (zf - zi))
        a0 = rs.A0
        if coords == 'cartesian':
            ak = np.fft.fft(np.fft.fft(a0, axis=1), axis=2)
            expect = np.fft.ifft(np.fft.ifft(phase * ak, axis=2), axis=1)
        else:
            T = rs.field_tool.GetTransverseTool()
            expect = T.rspace_host(phase * T.kspace_host(a0))
        # the conditioner zeroes the four lowest bins after every stage
        expect[:4] = 0.0
        assert rel_l2(A[..., 1], expect) < 1e-12
        # zero source: alpha = kappa = 0, dz = dphi/(0.5 dphi/L) = 2L, clamped to L -> one step per subcycle
        assert evals == [4 * vol['subcycles']]
        assert np.allclose(rec, L, rtol=1e-15)
        assert np.all(J[..., 1] == 0) and np.all(ne[..., 1] == 0)

    @pytest.mark.parametrize('coords', ['cartesian', 'cylindrical'])
    def test_reference_gaussian_beam(self, oracle_mod, coords):
        '''test/gaussian_test.py:110-155: |A| at 2 zR and 4 zR is the Gaussian beam 1/sqrt(5), 1/sqrt(17).'''
        vol, p = configs.gaussian_vacuum(coords)
        A, J, chi, ne, dom4d, evals = oracle_mod.oracle_track(vol)
        assert A.shape == vol['wave grid']
        dw = (dom4d[1] - dom4d[0]) / A.shape[0]
        dx = (dom4d[3] - dom4d[2]) / A.shape[1]
        w = np.linspace(dom4d[0] + 0.5 * dw, dom4d[1] - 0.5 * dw, A.shape[0])
        x = np.linspace(dom4d[2] + 0.5 * dx, dom4d[3] - 0.5 * dx, A.shape[1])
        r00 = p['r00']
        for plane, fac in ((1, 5), (2, 17)):
            wfac = np.exp(-(w[32] - 1.0) ** 2 / .1 ** 2)
            if coords == 'cartesian':
                dy = (dom4d[5] - dom4d[4]) / A.shape[2]
                y = np.linspace(dom4d[4] + 0.5 * dy, dom4d[5] - 0.5 * dy, A.shape[2])
                line = np.abs(A[32, :, 64, plane])
                yfac = np.exp(-y[64] ** 2 / r00 ** 2 / fac)
            else:
                line = np.abs(A[32, :, 0, plane])
                yfac = 1.0
            expect = wfac * yfac * np.exp(-x ** 2 / r00 ** 2 / fac) / fac ** 0.5
            assert np.allclose(line, expect, .02, .01)


class TestStepper:
    def _hooked(self, lam, dzmin_frac, steps=4):
        from oracle import oracle
        vol = configs.radial_kerr_air(scale=16, steps=steps)
        rs = uppe_kernel.RunSetup(vol)
        st = rs.make_stepper(oracle.oracle_factory)
        st.set_linear_rhs(lam)
        zi, zf, L = rs.plane(0)
        return rs, st, zi, zf, L

    @pytest.mark.parametrize('lam', [-3e-5 + 2e-5j, 1e-5 - 4e-5j])
    def test_rk4_polynomial(self, lam):
        rs, st, zi, zf, L = self._hooked(lam, 1.0)
        a0 = rs.A0
        st.set_field(a0)
        q0 = st.get_state()
        n, rec = st.advance(zi, L, 0, 1, dzmin=L, record_cap=8)
        assert n == 1 and rec[0] == L
        x = L * lam
        R = 1 + x + x ** 2 / 2 + x ** 3 / 6 + x ** 4 / 24
        expect = R * q0
        expect[:4] = 0.0
        got = st.get_state()
        assert rel_l2(got, expect) < 1e-14
        assert np.all(got[:4] == 0)

    @pytest.mark.parametrize('lam', [-3e-6 + 2e-6j, 4e-7 - 6e-6j])
    def test_step_size_rule(self, lam):
        '''dz = dphi / (0.5 dphi/L + |kappa| - alpha [alpha < 0]) with alpha, kappa weighted by
        |q|^2/(|q|^2 + amin^2), amin = 1e-2 max|q|; minimum over points, then dzmin, then L - z.'''
        rs, st, zi, zf, L = self._hooked(lam, 0.0, steps=1)
        a0 = rs.A0
        st.set_field(a0)
        dphi = 0.4
        c = abs(lam.imag) - (lam.real if lam.real < 0 else 0.0)
        rho = 1.0 / (1.0 + 1e-4)
        dz_expect = dphi / (0.5 * dphi / L + c * rho)
        n, rec = st.advance(zi, L, 0, 1, dzmin=1.0, record_cap=10000)
        assert n == len(rec) > 2
        assert np.allclose(rec[:-1], dz_expect, rtol=1e-12)
        assert rec[-1] <= dz_expect * (1 + 1e-12)
        assert np.sum(rec) == pytest.approx(L, rel=1e-12)

    def test_dzmin_floor(self):
        lam = -3e-6 + 2e-6j
        rs, st, zi, zf, L = self._hooked(lam, 0.0, steps=1)
        st.set_field(rs.A0)
        dzmin = L / 3.5
        n, rec = st.advance(zi, L, 0, 1, dzmin=dzmin, record_cap=100)
        # estimate < dzmin, so three steps of dzmin and a last step of L - 3 dzmin (uppe_kernel.py:224-227)
        assert n == 4
        assert np.allclose(rec[:3], dzmin, rtol=1e-15)
        assert rec[3] == pytest.approx(L - 3 * dzmin, rel=1e-12)


class TestSource:
    def _source_stepper(self, nw=257, redchi3=1.0, ion=None, ngas=2.0):
        from oracle import oracle
        st = oracle.OracleStepper(_abi.UPPE_RADIAL, nw, 2, 1, 2)
        st.set_hankel(np.eye(2)[:, :, None], np.ones(2))
        w = np.linspace(0, 8.0 * (nw - 1) / nw, nw)
        dt = np.pi / w[-1]
        kz = np.ones((nw, 2, 1), dtype=complex)
        st.set_medium(w, kz, np.zeros(nw), np.full((2, 1), ngas), np.ones(nw), np.ones(nw), redchi3, dt,
                      ion or ionization.no_ionization_params())
        return st, w, dt

    @staticmethod
    def _spectrum(e_t):
        return np.fft.rfft(e_t, axis=0)

    def test_kerr_current_is_minus_index_derivative(self):
        st, w, dt = self._source_stepper()
        nt = st.nt
        tau = np.arange(nt) * dt
        tc, T, w0 = tau[nt // 2], 40.0, 1.0
        g = np.exp(-((tau - tc) / T) ** 2)
        e = 0.3 * np.cos(w0 * tau) * g
        de = 0.3 * (-w0 * np.sin(w0 * tau) * g + np.cos(w0 * tau) * g * (-2 * (tau - tc) / T ** 2))
        E = np.repeat(self._spectrum(e)[:, None, None], 2, axis=1)
        J, chi, ne = st.source(E)
        j_t = np.fft.irfft(J[:, 0, 0])
        expect = -2.0 * 1.0 * 3 * e ** 2 * de          # -d/dtau (ng chi3 E^3)
        assert rel_l2(j_t, expect) < 1e-9
        assert np.allclose(chi[:, 0, 0], 2.0 * e ** 2, rtol=1e-12, atol=1e-13 * np.max(2.0 * e ** 2))
        assert np.all(ne == 0)

    def _constant_rate_params(self, log_rate):
        p = ionization.no_ionization_params()
        p.update(kind=ionization.ION_SERIES, e_conv=1.0, rate_conv=1.0, e_cutoff=1e-3,
                 c=[log_rate, 0, 0, 0, 0, 0])
        return p

    @staticmethod
    def _window(n, nt):
        '''SoftTimeWindow(4, 64) as a function of the time index (fft.cl:331-355), written from its definition.'''
        s = np.ones(nt)
        idx = np.arange(nt)
        s[(idx < 4) | (idx > nt - 5)] = 0.0
        q = (idx - 4) / 64.0
        ramp = 10 * q ** 3 - 15 * q ** 4 + 6 * q ** 5
        lo = (idx >= 4) & (idx < 68)
        s[lo] *= ramp[lo]
        q2 = (nt - 5 - idx) / 64.0
        ramp2 = 10 * q2 ** 3 - 15 * q2 ** 4 + 6 * q2 ** 5
        hi = (idx <= nt - 5) & (idx > nt - 69)
        s[hi] *= ramp2[hi]
        return s

    def test_density_closed_form_constant_rate(self):
        r0 = 1e-3
        st, w, dt = self._source_stepper(redchi3=0.0, ion=self._constant_rate_params(np.log(r0)), ngas=0.7)
        nt = st.nt
        e = 0.02 * np.sin(np.arange(nt) * dt * 1.0)
        E = np.repeat(self._spectrum(e)[:, None, None], 2, axis=1)
        J, chi, ne = st.source(E)
        k = np.arange(nt)
        raw = r0 * dt * (nt - 1.5 - k)                 # rate at index nt-1 zeroed, trapezoid from the top
        raw[nt - 1] = 0.0
        expect = 0.7 * (1 - np.exp(-raw)) * self._window(k, nt)
        assert np.allclose(ne[:, 0, 0], expect, rtol=1e-13, atol=1e-17)

    def test_scan_matches_bruteforce(self, oracle_mod):
        rng = np.random.default_rng(5)
        rate = rng.random((64, 3))
        ng = np.array([1.0, 2.0, 0.5])
        dt = 0.37
        got = oracle_mod.plasma_density(rate, ng, dt)
        expect = np.zeros_like(rate)
        for j in range(3):
            raw, prev = 0.0, 0.0
            for k in range(63, -1, -1):
                cur = 0.0 if k == 63 else rate[k, j]
                raw += 0.5 * (cur + prev) * dt
                expect[k, j] = ng[j] * (1 - np.exp(-raw))
                prev = cur
        assert np.allclose(got, expect, rtol=1e-14, atol=0)

    def test_plasma_current_ode(self):
        '''0.1 J - dJ/dtau = n_e E in the index time tau (physical t = -tau): dJ/dt = n_e E - 0.1 J.'''
        r0 = 2e-3
        st, w, dt = self._source_stepper(nw=1025, redchi3=0.0, ion=self._constant_rate_params(np.log(r0)),
                                         ngas=0.5)
        nt = st.nt
        tau = np.arange(nt) * dt
        tc = tau[nt // 2]
        e = 0.05 * np.cos(1.0 * tau) * np.exp(-((tau - tc) / 60.0) ** 2)
        E = np.repeat(self._spectrum(e)[:, None, None], 2, axis=1)
        J, chi, ne = st.source(E)
        j_t = np.fft.irfft(J[:, 0, 0])
        # 4th-order centred derivative
        d = (-np.roll(j_t, -2) + 8 * np.roll(j_t, -1) - 8 * np.roll(j_t, 1) + np.roll(j_t, 2)) / (12 * dt)
        src = ne[:, 0, 0] * e
        resid = 0.1 * j_t - d - src
        inner = slice(200, nt - 200)
        # the 4th-order difference has a truncation error of about (w dt)^4/30 ~ 1e-3 at this sampling
        assert np.max(np.abs(resid[inner])) < 3e-3 * np.max(np.abs(src))
        # with the opposite sign convention the residual would be of the order of the source itself
        bad = 0.1 * j_t + d - src
        assert np.max(np.abs(bad[inner])) > 0.5 * np.max(np.abs(src))

    @pytest.mark.parametrize('kind', ['explicit', 'series'])
    def test_rate_kernels_match_host_model(self, kind):
        ionizer = ionization.Ionizer(configs.o2_ionizer(kind))
        p = ionizer.kernel_params()
        st, w, dt = self._source_stepper(redchi3=0.0, ion=p, ngas=1.0)
        nt = st.nt
        e0 = 0.015    # simulation units; constant field -> constant rate
        E = np.zeros((st.shape[0], 2, 1), dtype=complex)
        E[0] = e0 * nt
        J, chi, ne = st.source(E)
        r0 = ionizer.Rate(np.array([e0]))[0]
        k = np.arange(nt)
        raw = r0 * dt * (nt - 1.5 - k)
        raw[nt - 1] = 0.0
        expect = 1.0 * (1 - np.exp(-raw)) * self._window(k, nt)
        assert np.allclose(ne[:, 0, 0], expect, rtol=1e-10, atol=1e-300)
        assert r0 > 0


class TestOracleABI:
    def test_bad_grids_rejected(self, oracle_mod):
        with pytest.raises(_abi.UppeError):
            oracle_mod.OracleStepper(_abi.UPPE_RADIAL, 64, 16, 1, 16)
        with pytest.raises(_abi.UppeError):
            oracle_mod.OracleStepper(_abi.UPPE_RADIAL, 65, 12, 1, 12)
        with pytest.raises(_abi.UppeError):
            oracle_mod.OracleStepper(_abi.UPPE_RADIAL, 65, 16, 1, 17)

    def test_call_order(self, oracle_mod):
        st = oracle_mod.OracleStepper(_abi.UPPE_RADIAL, 65, 16, 1, 16)
        with pytest.raises(_abi.UppeError):
            st.set_field(np.zeros((65, 16, 1), complex))     # Hankel tables missing
        with pytest.raises(_abi.UppeError):
            st.advance(0.0, 1.0, 0, 1, 1.0)                # no field


class TestRotationalRaman:
    '''The oracle's rotational susceptibility (the serial recurrence of rotations.cl:80-129) against its closed
    form: rho and cum are prefix sums over m = Nt-1-n, so for each state

        dchi[n] = sum_{m'' <= m} 2 T_j S_j n_g dalpha Im(rho_m'' conj(E^(2m'')) W_j dt),
        rho_m'' = -i dt sum_{m' < m''} F_m' E^(2m'+1),   E = exp(i W_j dt / 2),

    evaluated here with explicit powers and double sums (no recurrence).'''

    @staticmethod
    def _closed_form(rot, Et, ng, dt):
        N = Et.shape[0]
        wdt = rot.Wj * dt
        out = np.zeros(N)
        F = 0.5 * (rot.dalpha / rot.hbar) * Et[::-1] ** 2  # indexed by m
        m = np.arange(N)
        for j in range(2, rot.states):
            Sjj = 2.0 * j * (j - 1) / (15 * (2 * j - 1))
            E = np.exp(0.5j * wdt[j].real - 0.5 * wdt[j].imag)
            a = F * E ** (2 * m + 1)
            rho = -1j * dt * np.concatenate(([0], np.cumsum(a)[:-1]))
            g = 2 * rot.Tj[j] * Sjj * ng * rot.dalpha * np.imag(rho * np.conj(E ** (2 * m)) * wdt[j])
            out += np.cumsum(g)[::-1]
        return out

    def test_against_closed_form(self, oracle_mod):
        from searay_amd import configs
        rot = configs.n2_rotator(states=12)
        nw, n1, n2 = 33, 2, 2
        st = oracle_mod.OracleStepper(_abi.UPPE_CARTESIAN, nw, n1, n2, n1, 0)
        w = np.linspace(0.0, 2.0, nw)
        kz = np.repeat(w[:, None, None], n1 * n2, axis=1).reshape(nw, n1, n2).astype(np.complex128)
        kz[0] = 1.0
        ng = np.array([[1.0, 0.5], [2.0, 0.25]])
        dt = np.pi / w[-1]
        ion = ionization.no_ionization_params()
        st.set_medium(w, kz, np.zeros(nw), ng, np.ones(nw), np.ones(nw), 0.0, dt, ion)
        st.set_rotator(rot, dt)
        rng = np.random.default_rng(11)
        spec = np.exp(-(w - 1.0) ** 2 / 0.05)[:, None, None] * (1 + 0.1 * rng.standard_normal((nw, n1, n2)))
        A = 3e-3 * spec.astype(np.complex128)
        st.set_field(A)
        a, J, chi_t, ne_t = st.finish(0.0)
        st.close()
        E = 1j * kz * a
        Et = np.fft.irfft(E, axis=0)
        for x in range(n1):
            for y in range(n2):
                ref = self._closed_form(rot, Et[:, x, y], ng[x, y], dt)
                assert np.abs(ref).max() > 0
                assert rel_l2(chi_t[:, x, y], ref) < 1e-10

    def test_reference_shaped_rotator(self, oracle_mod):
        '''Any object with the reference Rotator's attributes (rotations.py:27-35) is accepted.'''
        import types
        from searay_amd import configs
        rot = configs.n2_rotator(states=16)
        bare = types.SimpleNamespace(Wj=rot.Wj.copy(), Tj=rot.Tj.copy(), dalpha=rot.dalpha, hbar=rot.hbar)
        vol = configs.radial_raman_air(scale=16, steps=2)
        vol['rotator'] = rot
        vol_bare = configs.radial_raman_air(scale=16, steps=2)
        vol_bare['rotator'] = bare
        a = oracle_mod.oracle_track(vol)
        b = oracle_mod.oracle_track(vol_bare)
        assert rel_l2(b[2][..., -1], a[2][..., -1]) == 0.0

    def test_switched_off(self, oracle_mod):
        from searay_amd import configs
        vol = configs.radial_raman_air(scale=16, steps=2)
        vol_off = configs.radial_raman_air(scale=16, steps=2)
        vol_off['rotator'] = None
        a = oracle_mod.oracle_track(vol)
        b = oracle_mod.oracle_track(vol_off)
        c = oracle_mod.oracle_track(configs.radial_kerr_air(scale=16, steps=2, seed=1003))
        assert rel_l2(b[0][..., -1], c[0][..., -1]) == 0.0
        assert rel_l2(a[0][..., -1], c[0][..., -1]) > 1e-4


class TestRotationalRamanContinuum:
    '''The oracle's rotational susceptibility against the continuous equation its recurrence discretises
    (rotations.cl:80-129), evaluated by quadrature. With m dt = t (m = Nt-1-n: the strip runs backwards in time),
    rho_j(t) = -i int_0^t F(t') exp(i W_j t') dt' and dchi_j/dt = 2 T_j S_j n_g dalpha Im(W_j rho_j conj(exp(i W_j t))),
    W_j = w_j - w_{j-2} + i damp_j. Undamped this is the impulsive Raman response
    chi(t) = -2 n_g dalpha sum_j T_j S_j int_0^t F(t') sin(w_j (t - t')) dt'; the recurrence must converge to it at
    second order in dt. Damped, the reference's factor is exp(-damp (t + t')) (both exponentials carry the decay),
    not the exp(-damp (t - t')) of a decaying coherence: the oracle must converge to the former and stay clear of
    the latter.'''

    T, T0, TAU, B, DALPHA, KT, HBAR, NS, NG = 40.0, 8.0, 1.5, 0.004, 0.7, 0.2, 1.0, 10, 1.3

    def _field(self, t):
        return np.exp(-((t - self.T0) / self.TAU) ** 2)

    def _rotator(self, damp):
        from searay_amd import rotations
        j = np.arange(self.NS)
        return rotations.Rotator(self.B, self.DALPHA, self.KT, 6 - 3 * (j % 2), damp * np.ones(self.NS), self.HBAR)

    def _oracle_chi(self, oracle_mod, nw, rot):
        n = 2 * (nw - 1)
        dt = self.T / n
        w = np.arange(nw) * (2 * np.pi / self.T)
        st = oracle_mod.OracleStepper(_abi.UPPE_CARTESIAN, nw, 1, 1, 1, 0)
        kz = np.ones((nw, 1, 1), np.complex128)  # E = i kz A = i A
        st.set_medium(w, kz, np.zeros(nw), np.array([[self.NG]]), np.ones(nw), np.ones(nw), 0.0, dt,
                      ionization.no_ionization_params())
        st.set_rotator(rot, dt)
        t = np.arange(n) * dt
        st.set_field((np.fft.rfft(self._field(t)[::-1]) / 1j)[:, None, None])
        _, _, chi_t, _ = st.finish(0.0)
        st.close()
        return t, chi_t[::-1, 0, 0]

    def _continuum(self, rot, t, decay_sum):
        from scipy import integrate
        out = 0.0
        for j in range(2, rot.states):
            s = 2.0 * j * (j - 1) / (15 * (2 * j - 1))
            om, g = rot.Wj[j].real, rot.Wj[j].imag

            def f(tp):
                # int_tp^t exp(-g s') (om cos(om (s'-tp)) + g sin(om (s'-tp))) ds', times the decay of rho
                z = (np.exp((-g + 1j * om) * (t - tp)) - 1) / (-g + 1j * om)
                decay = np.exp(-2 * g * tp) if decay_sum else 1.0
                return 0.5 * self.DALPHA / self.HBAR * self._field(tp) ** 2 * decay * (om * z.real + g * z.imag)
            out += -2 * rot.Tj[j] * s * self.NG * self.DALPHA * integrate.quad(f, 0, t, limit=200)[0]
        return out

    def _error(self, oracle_mod, nw, rot, decay_sum=True):
        t, chi = self._oracle_chi(oracle_mod, nw, rot)
        k = max(1, (nw - 1) // 16)
        ref = np.array([self._continuum(rot, x, decay_sum) for x in t[::k]])
        return rel_l2(chi[::k], ref)

    def test_undamped_second_order(self, oracle_mod):
        rot = self._rotator(0.0)
        assert 0.1 < rot.Wj.real.max() * self.T / 256 < 0.2  # W dt of the coarse grid
        e1, e2 = self._error(oracle_mod, 129, rot), self._error(oracle_mod, 257, rot)
        assert e1 < 5e-4
        assert 3.6 < e1 / e2 < 4.4

    def test_damping_form(self, oracle_mod):
        rot = self._rotator(0.01)
        e1, e2 = self._error(oracle_mod, 129, rot), self._error(oracle_mod, 257, rot)
        assert e2 < 1e-3 and 1.8 < e1 / e2 < 2.2
        assert self._error(oracle_mod, 257, rot, decay_sum=False) > 0.1


class TestPlaneWaveRHS:
    '''The UPPE right-hand side assembled by the oracle (E = i kz A, uppe.cl:93-108; the Kerr current of
    update_current, uppe_kernel.py:139-186; S = (i/2) exp(+Im kz z) exp(-i(Re kz + kg) z) J / Re kz, uppe.cl:73-91)
    against the closed-form plane-wave solutions of tests/plane_wave.py, in BK7 (kz differs from w by 50 %):
    the self-phase of a transparent medium, self-phase with absorption (amplitude and phase), and third-harmonic
    growth with phase mismatch. Planting a dropped 1/2, 1/w for 1/Re kz, E = i w A, a flipped exponent sign or
    exp(-Im kz z) in the oracle fails these checks by orders of magnitude beyond their tolerances.'''

    BAND = (6, 11)  # nonlinear band: the fundamental's bin 8 and its neighbours only

    def _spm(self, geometry, absorb, phase, steps):
        from oracle import oracle
        kz = pw.bk7_kz(absorb)
        st, shape, nrm = pw.make(oracle.oracle_factory, geometry, kz, self.BAND)
        c = pw.spm_rate(kz[pw.M], nrm)
        L = 400.0
        u0 = phase / (c * L)                     # nonlinear phase `phase` over L without absorption
        q0 = np.sqrt(u0) * np.exp(0.3j)
        q = pw.run(st, pw.field(shape, q0 / nrm), L, steps)
        st.close()
        expect = pw.spm_solution(c, kz[pw.M], q0, L)
        got = q[pw.M, 0, 0]
        others = np.delete(q.reshape(pw.NW, -1), pw.M, axis=0)
        return got, expect, q0, others, q

    @pytest.mark.parametrize('geometry', ['cartesian', 'radial'])
    def test_self_phase_transparent(self, geometry):
        got, expect, q0, others, q = self._spm(geometry, 0.0, 1.0, 400)
        assert abs(got - expect) < 1e-11 * abs(q0)
        assert abs(abs(got) - abs(q0)) < 1e-12 * abs(q0)
        assert np.angle(got / q0) == pytest.approx(1.0, abs=1e-10)
        assert np.abs(others).max() < 1e-12 * abs(q0)
        if geometry == 'radial':
            # identity basis: every radial cell carries the same plane wave
            assert np.allclose(q[pw.M], q[pw.M, 0, 0], rtol=1e-14, atol=0)

    @pytest.mark.parametrize('geometry', ['cartesian', 'radial'])
    def test_self_phase_absorbing(self, geometry):
        kz1 = pw.bk7_kz(3.8e-3)[pw.M]
        assert 0.8 < 2 * kz1.imag * 400.0 < 1.2
        got, expect, q0, others, q = self._spm(geometry, 3.8e-3, 1.0, 400)
        # |q| falls by the Im kz / Re kz part of kz / Re kz; the phase rate decays as exp(-2 Im kz z)
        assert 1e-4 < 1 - abs(got) / abs(q0) < 1e-3
        assert abs(got - expect) < 1e-11 * abs(q0)
        transparent = np.angle(pw.spm_solution(pw.spm_rate(kz1, 1.0), kz1.real + 0j, 1.0, 400.0))
        assert abs(np.angle(got / q0) - transparent) > 0.1

    def test_third_harmonic(self):
        from oracle import oracle
        kz = pw.bk7_kz()
        st, shape, nrm = pw.make(oracle.oracle_factory, 'cartesian', kz, (0, pw.NW))
        c = pw.spm_rate(kz[pw.M], nrm)
        L = 50.0
        q0 = np.sqrt(1e-4 / (c * L)) * np.exp(0.3j)   # weak: self-phase 1e-4 rad over L
        q = pw.run(st, pw.field(shape, q0 / nrm), L, 400)
        st.close()
        d = 3 * kz[pw.M] - kz[3 * pw.M]
        assert 2.0 < abs(d) * L / np.pi < 3.0           # a few coherence lengths: the growth oscillates
        expect = pw.third_harmonic(kz, q0, 1e-3, nrm, L)
        got = q[3 * pw.M, 0, 0]
        assert abs(got - expect) < 1e-3 * abs(expect)

