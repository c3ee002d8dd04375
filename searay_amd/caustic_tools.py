'''Grid geometry of the UPPE wave region and its coupling to rays (host side).

Restates SeaRay's ``modules/caustic_tools.py`` for the UPPE boundary: ``FourierTool`` (Cartesian,
caustic_tools.py:82-205) and ``BesselBeamTool`` (cylindrical, caustic_tools.py:207-338). Each gives the grid
(``GetGridInfo``), the incoming field built from ray data on the entrance plane (``GetBoundaryFields``), and the
rays relaunched from the field on the exit plane (``RelaunchRays``, and ``RelaunchRays1``, which also recovers the
transverse wavevector from the field's phase).

Ray data follows the reference's layout: ``xp`` phase space (bundles, rays, 8) = (t, x, y, z, w, kx, ky, kz) with
4 rays per bundle (a primary and three satellites), ``eikonal`` (bundles, 4) = (phase, Ax, Ay, Az), ``vg``
(bundles, rays, 4). The relaunch methods update ``xp``, ``eikonal`` and ``vg`` in place.
'''
import logging

import numpy as np
import scipy.integrate

from . import grid_tools


class CausticTool:
    '''Shared grid bookkeeping (caustic_tools.py:25-42).

    pts = (Nw, N1, N2, Nz), band = (w_low, w_high), center = (x, y, z), size = (Lx, Ly, Lz).'''

    def __init__(self, pts, band, center, size, cl=None):
        self.pts = pts
        w0, dw = grid_tools.cyclic_center_and_width(band[0], band[1])
        self.center = np.array([w0] + list(center), dtype=float)
        self.size = np.array([dw] + list(size), dtype=float)
        self.dx = self.size / np.array(list(pts[:3]) + [1], dtype=float)
        self.default_band = size[0]

    def GetTransverseTool(self):
        return self.T

    def kr2(self):
        return self.T.kr2()

    def frequency_nodes(self):
        return self.center[0] + grid_tools.cyclic_nodes(-self.size[0] / 2, self.size[0] / 2, self.pts[0])

    def ClipRaysToGrid(self, xp):
        '''Bundles whose four rays all lie within the grid's frequency and transverse node range; an axis with one
        node accepts everything (caustic_tools.py:66-77).'''
        wn, xn, yn, _ = self.GetGridInfo()

        def inside(v, nodes):
            return np.logical_or(np.logical_and(v >= nodes[0], v <= nodes[-1]), nodes.shape[0] == 1)
        cond = inside(xp[..., 4], wn) & inside(xp[..., 1], xn) & inside(xp[..., 2], yn)
        keep = cond[:, 0] & cond[:, 1] & cond[:, 2] & cond[:, 3]
        return np.where(keep)[0]

    def _time_reference(self, xps, eiks):
        '''Eikonal phase referred to the arrival of the strongest ray, with that ray put at the centre of the
        time window: psi = phase - w (t_strongest + T/2), T = 2 pi Nw / bandwidth (caustic_tools.py:104-106).'''
        time_window = 2 * np.pi * self.pts[0] / self.size[0]
        largest = np.argmax(np.sum(eiks[:, 1:3] ** 2, axis=1))
        return eiks[:, 0] - xps[:, 4] * (xps[largest, 0] + 0.5 * time_window)

    @staticmethod
    def _advance_rays(xp, vg, impact, L, speed):
        '''Rays leave the exit plane: t advanced by L / speed, z by L (caustic_tools.py:167-170).'''
        vg[impact, ...] = xp[impact, ..., 4:8] / xp[impact, ..., 4:5]
        xp[impact, :, 0] += L / speed
        xp[impact, :, 3] += L

    @staticmethod
    def _phase_gradient(a, nodes, axis, eps):
        '''k along one axis from the field's phase without unwrapping it: Im(conj(a) da) / (eps + |a|)^2, with the
        centred-difference derivative (caustic_tools.py:131-144).'''
        if nodes.shape[0] > 1:
            g = a * np.gradient(np.conj(a), nodes[1] - nodes[0], axis=axis)
            return -np.imag(g) / (eps + np.abs(a)) ** 2, nodes[1] - nodes[0]
        return np.zeros(a.shape), 1.0


class FourierTool(CausticTool):
    '''Cartesian transverse grid: cell-centred x and y over the box (caustic_tools.py:82-96).'''

    def __init__(self, pts, band, center, size, cl=None):
        super().__init__(pts, band, center, size, cl)
        self.T = grid_tools.FourierTransformTool(self.pts, self.dx)

    def GetGridInfo(self):
        w_nodes = self.frequency_nodes()
        x_nodes = grid_tools.cell_centers(-self.size[1] / 2, self.size[1] / 2, self.pts[1])
        y_nodes = grid_tools.cell_centers(-self.size[2] / 2, self.size[2] / 2, self.pts[2])
        w_walls = grid_tools.cell_walls(w_nodes[0], w_nodes[-1], self.pts[0], self.default_band)
        x_walls = grid_tools.cell_walls(x_nodes[0], x_nodes[-1], self.pts[1])
        y_walls = grid_tools.cell_walls(y_nodes[0], y_nodes[-1], self.pts[2])
        ext = np.array([w_walls[0], w_walls[-1], x_walls[0], x_walls[-1], y_walls[0], y_walls[-1]])
        return w_nodes, x_nodes, y_nodes, ext

    def GetBoundaryFields(self, xps, eiks, c):
        '''Field component c (1, 2, 3 = x, y, z) on the (w, x, y) grid from primary rays (xps (bundles, 8), eiks
        (bundles, 4)) already on the entrance plane: amplitude and time-referenced phase each interpolated
        (cubic, per frequency bin), then combined as A exp(i psi) (caustic_tools.py:98-116).'''
        w = xps[:, 4]
        dx = xps[:, 1:4] - self.center[1:4]
        wn, xn, yn, _ = self.GetGridInfo()
        psi_rays = self._time_reference(xps, eiks)
        A, _ = grid_tools.GridFromInterpolation(w, dx[:, 0], dx[:, 1], eiks[:, c], wn, xn, yn)
        psi, _ = grid_tools.GridFromInterpolation(w, dx[:, 0], dx[:, 1], psi_rays, wn, xn, yn)
        return A * np.exp(1j * psi)

    def RelaunchRays1(self, xp, eikonal, vg, a, L, material):
        '''Relaunch rays from the exit-plane field a (Nw, Nx, Ny) with the transverse wavevector of the field's
        phase (caustic_tools.py:118-180). The phase itself comes from del^2 phase = div k, solved spectrally with
        the finite-difference wavenumbers; rays keep their frequency and transverse position and travel at the
        vacuum speed of their new direction.'''
        wn, xn, yn, _ = self.GetGridInfo()
        eps = np.max(np.abs(a)) * 1e-6
        kx, _ = self._phase_gradient(a, xn, 1, eps)
        ky, _ = self._phase_gradient(a, yn, 2, eps)
        iKx = 1j * np.outer(self.T.k_diff(1), np.ones(kx.shape[2]))
        iKy = 1j * np.outer(np.ones(ky.shape[1]), self.T.k_diff(2))
        src = iKx * np.fft.fft(np.fft.fft(kx, axis=1), axis=2) + iKy * np.fft.fft(np.fft.fft(ky, axis=1), axis=2)
        K2m = iKx ** 2 + iKy ** 2
        src[:, 0, 0] = 0.0
        K2m[0, 0] = 1.0
        phase = np.real(np.fft.ifft(np.fft.ifft(src / K2m, axis=2), axis=1))
        impact = self.ClipRaysToGrid(xp)
        w, x, y = xp[impact, :, 4], xp[impact, :, 1], xp[impact, :, 2]
        k1 = grid_tools.DataFromGrid(w, x, y, wn, xn, yn, kx)
        k2 = grid_tools.DataFromGrid(w, x, y, wn, xn, yn, ky)
        evanescent = k1 ** 2 + k2 ** 2 >= w ** 2
        k1[evanescent] = 0.0
        k2[evanescent] = 0.0
        xp[impact, :, 5] = k1
        xp[impact, :, 6] = k2
        xp[impact, :, 7] = np.sqrt(w ** 2 - k1 ** 2 - k2 ** 2)
        v = xp[impact, ..., 5:8] / xp[impact, ..., 4:5]
        self._advance_rays(xp, vg, impact, L, np.sqrt(np.sum(v ** 2, axis=-1)))
        eikonal[impact, 0] = grid_tools.DataFromGrid(w[:, 0], x[:, 0], y[:, 0], wn, xn, yn, phase)
        eikonal[impact, 1] = grid_tools.DataFromGrid(w[:, 0], x[:, 0], y[:, 0], wn, xn, yn, np.abs(a))
        eikonal[impact, 2] = 0.0
        eikonal[impact, 3] = 0.0

    def RelaunchRays(self, xp, eikonal, vg, a, L, material):
        '''Simplified relaunch (caustic_tools.py:182-205): rays on the grid leave along z at the material's group
        speed, with the field's phase and amplitude at their (w, x, y).'''
        logging.warning('Using simplified ray relaunch; only valid for close couplings.')
        wn, xn, yn, _ = self.GetGridInfo()
        impact = self.ClipRaysToGrid(xp)
        w, x, y = xp[impact, :, 4], xp[impact, :, 1], xp[impact, :, 2]
        xp[impact, :, 5] = 0.0
        xp[impact, :, 6] = 0.0
        xp[impact, :, 7] = w
        self._advance_rays(xp, vg, impact, L, material.GroupVelocityMagnitude(w))
        eikonal[impact, 0] = grid_tools.DataFromGrid(w[:, 0], x[:, 0], y[:, 0], wn, xn, yn, np.angle(a))
        eikonal[impact, 1] = grid_tools.DataFromGrid(w[:, 0], x[:, 0], y[:, 0], wn, xn, yn, np.abs(a))
        eikonal[impact, 2] = 0.0
        eikonal[impact, 3] = 0.0


class BesselBeamTool(CausticTool):
    '''Cylindrical grid: radial cell centres on (0, Lx/2) (no node at r=0), azimuth on [-pi, pi)
    (caustic_tools.py:207-229). mmax = N2/2 azimuthal modes; radial spacing dr = Lx/(2 Nr).'''

    def __init__(self, pts, band, center, size, cl=None, modes=None):
        self.pts = pts
        w0, dw = grid_tools.cyclic_center_and_width(band[0], band[1])
        self.center = np.array([w0] + list(center), dtype=float)
        self.size = np.array([dw] + list(size), dtype=float)
        self.mmax = int(self.pts[2] / 2)
        self.default_band = size[0]
        self.T = grid_tools.HankelTransformTool(self.pts[1], 0.5 * self.size[1] / self.pts[1], self.mmax,
                                                cl, Nk=modes)

    def GetGridInfo(self):
        w_nodes = self.frequency_nodes()
        rho_nodes = grid_tools.cell_centers(0.0, self.size[1] / 2, self.pts[1])
        phi_nodes = grid_tools.cyclic_nodes(-np.pi, np.pi, self.pts[2])
        w_walls = grid_tools.cell_walls(w_nodes[0], w_nodes[-1], self.pts[0], self.default_band)
        rho_walls = grid_tools.cell_walls(rho_nodes[0], rho_nodes[-1], self.pts[1])
        phi_walls = grid_tools.cell_walls(phi_nodes[0], phi_nodes[-1], self.pts[2])
        ext = np.array([w_walls[0], w_walls[-1], rho_walls[0], rho_walls[-1], phi_walls[0], phi_walls[-1]])
        return w_nodes, rho_nodes, phi_nodes, ext

    def GetCylCoords(self, xp):
        '''(rho, phi) of phase-space points about the tool's centre (caustic_tools.py:231-235).'''
        dx = xp[..., 1:4] - self.center[1:4]
        return np.sqrt(dx[..., 0] ** 2 + dx[..., 1] ** 2), np.arctan2(dx[..., 1], dx[..., 0])

    def GetBoundaryFields(self, xps, eiks, c):
        '''Field component c on the (w, rho, phi) grid from primary rays on the entrance plane: frequency and
        azimuth binned, rho interpolated linearly; azimuths above 0.9999 pi are folded to negative angles so that
        the cyclic axis favours -pi (caustic_tools.py:237-254).'''
        w = xps[:, 4]
        rho, phi = self.GetCylCoords(xps)
        wn, rn, pn, _ = self.GetGridInfo()
        psi_rays = self._time_reference(xps, eiks)
        phi = np.where(phi > 0.9999 * np.pi, phi - 2 * np.pi, phi)
        A, _ = grid_tools.CylGridFromInterpolation(w, rho, phi, eiks[:, c], wn, rn, pn)
        psi, _ = grid_tools.CylGridFromInterpolation(w, rho, phi, psi_rays, wn, rn, pn)
        return A * np.exp(1j * psi)

    def RelaunchRays1(self, xp, eikonal, vg, a, L, material):
        '''Relaunch with the radial wavevector of the field's phase, the phase integrated outward from a ghost
        cell inside the axis (caustic_tools.py:256-307). Azimuthal phase variation is ignored, as in the
        reference; interpolation failures leave the wavevector zero and the eikonal unchanged.'''
        wn, rn, pn, _ = self.GetGridInfo()
        a, rn = grid_tools.AddGhostCells(a, rn, 1)
        eps = np.max(np.abs(a)) * 1e-6
        kr_grid, dr = self._phase_gradient(a, rn, 1, eps)
        phase = scipy.integrate.cumulative_trapezoid(kr_grid, dx=dr, axis=1, initial=0.0)
        phase -= 0.25 * kr_grid[:, 0:1, :] * dr
        rho, phi = self.GetCylCoords(xp)
        impact = np.where(np.logical_and(rho[:, 0] < rn[-1], xp[:, 0, 4] < wn[-1]))[0]
        w, r, p = xp[impact, :, 4], rho[impact, :], phi[impact, :]
        try:
            kr = grid_tools.DataFromGrid(w, r, p, wn, rn, pn, kr_grid)
        except Exception:
            kr = np.zeros(p.shape)
        kr[kr ** 2 >= w ** 2] = 0.0
        xp[impact, :, 5] = kr * np.cos(p)
        xp[impact, :, 6] = kr * np.sin(p)
        xp[impact, :, 7] = np.sqrt(w ** 2 - xp[impact, :, 5] ** 2 - xp[impact, :, 6] ** 2)
        v = xp[impact, ..., 5:8] / xp[impact, ..., 4:5]
        self._advance_rays(xp, vg, impact, L, np.sqrt(np.sum(v ** 2, axis=-1)))
        try:
            ph = grid_tools.DataFromGrid(w[:, 0], r[:, 0], p[:, 0], wn, rn, pn, phase)
            amp = grid_tools.DataFromGrid(w[:, 0], r[:, 0], p[:, 0], wn, rn, pn, np.abs(a))
            eikonal[impact, 0] = ph
            eikonal[impact, 1] = amp
        except Exception:
            pass
        eikonal[impact, 2] = 0.0
        eikonal[impact, 3] = 0.0

    def RelaunchRays(self, xp, eikonal, vg, a, L, material):
        '''Simplified relaunch (caustic_tools.py:309-338): rays within the radial and frequency range leave along
        z at the material's group speed; the eikonal phase is w t (after the transit) plus the field's phase.'''
        logging.warning('Using simplified ray relaunch; only valid for close couplings.')
        wn, rn, pn, _ = self.GetGridInfo()
        rho, phi = self.GetCylCoords(xp)
        wtest = np.logical_and(xp[:, 0, 4] >= wn[0], xp[:, 0, 4] <= wn[-1])
        impact = np.where(np.logical_and(rho[:, 0] <= rn[-1], wtest))[0]
        w, r, p = xp[impact, :, 4], rho[impact, :], phi[impact, :]
        xp[impact, :, 5] = 0.0
        xp[impact, :, 6] = 0.0
        xp[impact, :, 7] = w
        self._advance_rays(xp, vg, impact, L, material.GroupVelocityMagnitude(w))
        eikonal[impact, 0] = w[:, 0] * xp[impact, 0, 0] + grid_tools.DataFromGrid(w[:, 0], r[:, 0], p[:, 0], wn, rn,
                                                                                  pn, np.angle(a))
        eikonal[impact, 1] = grid_tools.DataFromGrid(w[:, 0], r[:, 0], p[:, 0], wn, rn, pn, np.abs(a))
        eikonal[impact, 2] = 0.0
        eikonal[impact, 3] = 0.0
