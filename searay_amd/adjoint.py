'''Gradients of a UPPE output objective with respect to the input spectral phase, by the stepper's adjoint.

Two objectives, each the squared distance of an output observable to a target:

    spectrum:  L(c) = sum_w (|A(w, r_0, z_f)|^2 - S_tgt(w))^2
    fluence:   L(c) = sum_r (F(r) - F_tgt(r))^2,   F(r) = sum_w w^2 |A(w, r, z_f)|^2  (the time-integrated |E|^2 of
               cell r up to a constant, by Parseval, E = i w A)

with the input A(w, r, 0) = A_base(w, r) exp(i sum_n c_n P_n(x(w))), Legendre polynomials P_n over
x = (w - w_0) / (3 sigma) on |x| <= 1 (zero outside), r_0 the innermost radial cell (the grid has no node at r = 0,
caustic_tools.py:223), and the target the unperturbed observable times (1 + 0.1 xi), xi uniform in [-1, 1) from a
seeded generator.

The forward run goes through uppe_advance_checkpointed (the state every `checkpoint_every` steps), finish gives
A(z_f), and uppe_adjoint carries the cotangent G = dL/dRe A + i dL/dIm A (4 (|A|^2 - S) A on axis, or
4 (F - F_tgt) w^2 A) back to the initial field; then dL/dc_n = -sum_{w,r} P_n(x(w)) Im(conj(G_0) A_0).
``spectral_phase_loss_function`` puts the pair behind a torch.autograd.Function. Radial UPPE media with Kerr, the
rotational Raman susceptibility, photoionisation and the plasma current, on the matrix-core paths.
'''
import numpy as np
from scipy.special import eval_legendre

from . import uppe_kernel


def legendre_basis(w, w0, sigma, n):
    '''P_k(x), k < n, at x = (w - w0)/(3 sigma) where |x| <= 1, zero elsewhere: shape (n, len(w)).'''
    x = (np.asarray(w) - w0) / (3 * sigma)
    inside = np.abs(x) <= 1
    B = np.zeros((n, x.shape[0]))
    for k in range(n):
        B[k, inside] = eval_legendre(k, x[inside])
    return B


class SpectralPhaseProblem:
    '''An output objective (objective='spectrum' or 'fluence') of the input spectral phase on one diagonal
    interval of a radial UPPE workload (vol_dict as for track).'''

    def __init__(self, vol, sigma, precision='fp32', checkpoint_every=10, n_phase=32, w0=1.0, target_seed=1005,
                 target=None, stepper_factory=None, objective='spectrum'):
        if objective not in ('spectrum', 'fluence'):
            raise ValueError("objective must be 'spectrum' or 'fluence'")
        self.objective = objective
        self.rs = uppe_kernel.RunSetup(vol)
        if self.rs.diagnostic_steps != 1:
            raise ValueError('the objective is defined on one diagonal interval (wave grid with two planes)')
        if not self.rs.medium_is_uniform(0):
            raise ValueError('the density varies along z: the adjoint path loads the medium once')
        self.every = int(checkpoint_every)
        self.st = self.rs.make_stepper(stepper_factory or uppe_kernel.hip_stepper_factory(precision))
        self.basis = legendre_basis(self.rs.w, w0, sigma, n_phase)
        self.zi, self.zf, self.L = self.rs.plane(0)
        self._a0 = None
        self._g = None
        self._token = None  # identifies the forward() whose recorded run the device holds
        if target is None:
            s0 = self.spectrum(np.zeros(n_phase))
            rng = np.random.default_rng(target_seed)
            target = s0 * (1 + 0.1 * rng.uniform(-1.0, 1.0, s0.shape))
        self.target = np.asarray(target, dtype=np.float64)

    def field(self, c):
        phase = np.asarray(c, dtype=np.float64) @ self.basis
        return self.rs.A0 * np.exp(1j * phase)[:, None, None]

    def _run(self, c, record):
        a0 = self.field(c)
        self.st.set_field(a0)
        if record:
            self.st.advance_checkpointed(self.zi, self.L, 0, self.rs.subcycles, self.rs.dzmin, self.every)
        else:
            self.st.advance(self.zi, self.L, 0, self.rs.subcycles, self.rs.dzmin)
        a_r, _, _, _ = self.st.finish(self.zf - self.zi, want_diagnostics=False)
        return a0, a_r

    def _observable(self, a_r):
        if self.objective == 'spectrum':
            return np.abs(a_r[:, 0, 0]) ** 2
        return np.sum((self.rs.w ** 2)[:, None] * np.abs(a_r[:, :, 0]) ** 2, axis=0)

    def _cotangent(self, a_r, d):
        '''dL/dRe A + i dL/dIm A of L = sum d^2 at the output field a_r.'''
        g = np.zeros_like(a_r)
        if self.objective == 'spectrum':
            g[:, 0, 0] = 4 * d * a_r[:, 0, 0]
        else:
            g[:, :, 0] = 4 * d[None, :] * (self.rs.w ** 2)[:, None] * a_r[:, :, 0]
        return g

    def spectrum(self, c):
        '''The observable of the run with phase coefficients c: |A(w, r_0, z_f)|^2 ('spectrum') or F(r)
        ('fluence').'''
        _, a_r = self._run(c, False)
        return self._observable(a_r)

    def loss(self, c):
        '''L(c) by a plain forward run.'''
        return float(np.sum((self.spectrum(c) - self.target) ** 2))

    def forward(self, c):
        '''L(c) by a checkpointed forward run, keeping what backward() needs.'''
        a0, a_r = self._run(c, True)
        d = self._observable(a_r) - self.target
        self._a0, self._g = a0, self._cotangent(a_r, d)
        self._token = object()
        return float(np.sum(d ** 2))

    def backward(self):
        '''dL/dc of the last forward() by the adjoint sweep.'''
        if self._g is None:
            raise RuntimeError('backward() needs a forward() first')
        g0 = self.st.adjoint(self.zf - self.zi, self._g)
        self._g = None
        return self.basis @ (-np.imag(np.sum(np.conj(g0) * self._a0, axis=(1, 2))))

    def value_and_grad(self, c):
        loss = self.forward(c)
        return loss, self.backward()

    def close(self):
        self.st.close()


def spectral_phase_loss_function():
    '''torch.autograd.Function class for L(c): forward runs the checkpointed forward, backward the adjoint.'''
    import torch

    class SpectralPhaseLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, c, problem):
            ctx.problem = problem
            ctx.dtype = c.dtype
            ctx.device = c.device
            ctx.c = c.detach().cpu().double().numpy().copy()
            loss = problem.forward(ctx.c)
            ctx.token = problem._token
            return torch.tensor(loss, dtype=c.dtype, device=c.device)

        @staticmethod
        def backward(ctx, grad_out):
            p = ctx.problem
            if p._g is None or p._token is not ctx.token:
                # another forward ran on this problem since (its recorded run replaced this one's): record again
                p.forward(ctx.c)
            g = torch.as_tensor(p.backward(), dtype=ctx.dtype, device=ctx.device)
            return grad_out * g, None

    return SpectralPhaseLoss
