'''Grid generators, ray-grid interpolation and transverse mode tools for the UPPE propagator (host side, FP64).

Only the parts of SeaRay's ``modules/grid_tools.py`` that the UPPE path and its ray boundary need are provided:
node/wall generators (grid_tools.py:25-49), the interpolation between ray data and the wave grid that the ray
boundary and the ray relaunch use (grid_tools.py:51-177), the finite-difference Fourier tool's kr^2
(grid_tools.py:218-250) and the discrete Hankel tool's eigenbasis (grid_tools.py:252-298,357-358).
The device-side transforms themselves live in the HIP library; this module only builds tables.
'''
import logging

import numpy as np
import scipy.interpolate
import scipy.linalg


def cyclic_center_and_width(node_low, node_high):
    '''Centre and wall-to-wall width of a periodic node set (grid_tools.py:18-23).'''
    return 0.5 * (node_low + node_high), node_high - node_low


def cyclic_nodes(node_low, node_high, num_nodes):
    '''Nodes of a periodic axis: low end included, high end excluded (grid_tools.py:25-34).'''
    if num_nodes == 1:
        return np.array([0.5 * node_low + 0.5 * node_high])
    step = (node_high - node_low) / num_nodes
    return np.linspace(node_low, node_high - step, num_nodes)


def cell_centers(wall_low, wall_high, cells):
    '''Cell-centred nodes between two walls (grid_tools.py:36-40).'''
    step = (wall_high - wall_low) / cells
    return np.linspace(wall_low + 0.5 * step, wall_high - 0.5 * step, cells)


def cell_walls(node_low, node_high, nodes, default_voxel_width=1000.0):
    '''Walls bracketing a set of nodes (grid_tools.py:42-49).'''
    if nodes == 1:
        step = default_voxel_width
    else:
        step = (node_high - node_low) / (nodes - 1)
    return np.linspace(node_low - 0.5 * step, node_high + 0.5 * step, nodes + 1)


def hypersurface_idx(ary, axis, cell):
    """Index tuple of the one-cell slab ``cell`` along ``axis``, keeping that axis (grid_tools.py:51-55)."""
    idx = [slice(None)] * ary.ndim
    idx[axis] = slice(cell, cell + 1)
    return tuple(idx)


def AddGhostCells(a, nodes, ax):
    """Prepend one ghost cell along ``ax``: a copy of the first cell, at one node spacing below the first node
    (grid_tools.py:81-86)."""
    ghost = a[hypersurface_idx(a, ax, 0)]
    return np.concatenate((ghost, a), axis=ax), np.concatenate(([2 * nodes[0] - nodes[1]], nodes))


def DataFromGrid(w, x, y, wn, xn, yn, the_data):
    """Values of ``the_data`` (shape (len(wn), len(xn), len(yn))) at scattered points (w, x, y): trilinear on the
    regular nodes, linear extrapolation outside them (grid_tools.py:88-110). A single y node is widened to the
    y range of the points, with the data constant in y."""
    if wn.shape[0] == 1:
        raise ValueError('Cannot relaunch rays with monochromatic data.')
    if xn.shape[0] == 1:
        raise ValueError('Missing spatial dimension x.')
    if yn.shape[0] == 1:
        ygrid = np.linspace(np.min(y), np.max(y), 2)
        data = np.concatenate((the_data, the_data), axis=2)
    else:
        ygrid, data = yn, the_data
    f = scipy.interpolate.RegularGridInterpolator((wn, xn, ygrid), data, bounds_error=False, fill_value=None)
    return f(np.stack((w, x, y), axis=-1))


def _node_array(spec, pts):
    return np.linspace(np.min(pts), np.max(pts), spec) if type(spec) is int else spec


def _bins(walls, v):
    """Bin of each value between consecutive walls, walls[i] <= v < walls[i+1]; -1 outside."""
    b = np.searchsorted(walls, v, side='right') - 1
    return np.where((b >= 0) & (b < walls.shape[0] - 1), b, -1)


def _plot_extent(wn, xn, yn):
    ww = cell_walls(wn[0], wn[-1], wn.shape[0])
    wx = cell_walls(xn[0], xn[-1], xn.shape[0])
    wy = cell_walls(yn[0], yn[-1], yn.shape[0])
    return ww, wx, wy, np.array([ww[0], ww[-1], wx[0], wx[-1], wy[0], wy[-1]])


def CylGridFromInterpolation(w, x, y, vals, wn=1, xn=100, yn=100, fill=0.0):
    """Scattered ray data onto a (w, rho, phi) grid (grid_tools.py:112-146). Frequency and azimuth are binned
    between cell walls; within each (w, phi) bin the data is interpolated linearly in rho, taking the value of
    the innermost point below it and zero beyond the outermost. Empty bins are zero. ``wn``, ``xn``, ``yn`` are
    node arrays, or node counts spanning the points. Returns (data, plot extent)."""
    wn, xn, yn = _node_array(wn, w), _node_array(xn, x), _node_array(yn, y)
    ww, wx, wy, ext = _plot_extent(wn, xn, yn)
    out = np.zeros((wn.shape[0], xn.shape[0], yn.shape[0]))
    bw, by = _bins(ww, w), _bins(wy, y)
    ok = np.where((bw >= 0) & (by >= 0))[0]
    key = bw[ok] * yn.shape[0] + by[ok]
    order = np.argsort(key, kind='stable')  # keeps the points of one bin in their original order
    ok, key = ok[order], key[order]
    starts = np.flatnonzero(np.r_[True, key[1:] != key[:-1]]) if key.shape[0] else np.zeros(0, int)
    for s0, s1 in zip(starts, np.r_[starts[1:], key.shape[0]]):
        sel = ok[s0:s1]
        xs, vs = x[sel], vals[sel]
        v0 = vs[np.argmin(xs)]
        f = scipy.interpolate.interp1d(xs, vs, kind='linear', bounds_error=False, fill_value=(v0, 0.0))
        i, j = divmod(int(key[s0]), yn.shape[0])
        out[i, :, j] = f(xn)
    return out, ext


def GridFromInterpolation(w, x, y, vals, wn=1, xn=100, yn=100, fill=0.0):
    """Scattered ray data onto a (w, x, y) grid (grid_tools.py:148-177): frequency binned between cell walls,
    then a cubic (Clough-Tocher) interpolation in (x, y) per bin at the nodes pulled in by 0.9999, ``fill``
    outside the points' hull. A frequency bin without points ends the run, as in the reference. Returns
    (data, plot extent)."""
    wn, xn, yn = _node_array(wn, w), _node_array(xn, x), _node_array(yn, y)
    ww, wx, wy, ext = _plot_extent(wn, xn, yn)
    out = np.zeros((wn.shape[0], xn.shape[0], yn.shape[0]))
    xg = 0.9999 * np.outer(xn, np.ones(yn.shape[0]))
    yg = 0.9999 * np.outer(np.ones(xn.shape[0]), yn)
    bw = _bins(ww, w)
    for i in range(wn.shape[0]):
        sel = np.where(bw == i)[0]
        if sel.shape[0] == 0:
            logging.critical('no rays found in w-bin, w-dimension should be matched for all objects')
            raise SystemExit(1)
        pts = np.stack((x[sel], y[sel]), axis=-1)
        out[i] = scipy.interpolate.griddata(pts, vals[sel], (xg, yg), method='cubic', fill_value=fill)
    return out, ext


class TransverseModeTool:
    '''Common interface: grid shape, spacing and the transverse eigenvalues kr^2 (grid_tools.py:196-216).'''

    def __init__(self, N, dq):
        self.N = N
        self.dq = dq

    def kr2(self):
        return np.zeros(self.N[1:3])


class FourierTransformTool(TransverseModeTool):
    '''Plane-wave basis on a Cartesian grid.

    kr^2 is the eigenvalue of the 3-point finite-difference Laplacian, k_d(i)^2 = 2(1-cos(2 pi i/N))/dx^2,
    with the sign of k_d negative above N/2 (grid_tools.py:236-250). The device transforms are unnormalised
    forward FFTs over x then y and normalised inverse FFTs over y then x (grid_tools.py:220-229).'''

    def k_diff(self, axis):
        num = self.N[axis]
        dx = self.dq[axis]
        idx = np.arange(num)
        sgn = np.where(idx > num // 2, -1.0, 1.0)
        return sgn * np.sqrt(2.0 * (1.0 - np.cos(2.0 * np.pi * idx / num))) / dx

    def kr2(self):
        kx = self.k_diff(1)
        ky = self.k_diff(2)
        return (kx ** 2)[:, None] + (ky ** 2)[None, :]


class HankelTransformTool(TransverseModeTool):
    '''Discrete Hankel basis: eigenvectors of the symmetrised finite-volume radial Laplacian.

    Construction follows grid_tools.py:254-298:
      r_j = (j+1/2) dr, Lambda_j = sqrt(pi((r_j+dr/2)^2-(r_j-dr/2)^2)),
      tridiagonal T (sub A1/(dr V), diag -(A1+A2)/(dr V)-(m/r)^2, super A2/(dr V)) with the
      boundary corrections T2[0]+=T1[0], T2[-1]-=T3[-1], symmetrised by Lambda, and the Nk
      eigenpairs of largest eigenvalue kept in ascending order.
    H has shape (Nk, Nr, Nm) with H[k, j, m] = eigenvector k at radius j; kr2 = -eigenvalue (zero for
    unused modes).

    The forward transform used by the stepper is a_hat_k = Lambda_k^-1 sum_j H[k,j] Lambda_j a_j and the
    inverse a_j = Lambda_j^-1 sum_k H[k,j] Lambda_k a_hat_k (grid_tools.py:308-325).'''

    def __init__(self, Nr, dr, mmax, cl=None, Nk=None):
        if Nk is None:
            Nk = Nr
        self.Nr = Nr
        self.Nk = Nk
        self.dr = dr
        self.mmax = mmax
        self.N = (None, Nr, max(1, 2 * mmax), None)
        r = cell_centers(0.0, Nr * dr, Nr)
        inner_area = 2.0 * np.pi * (r - 0.5 * dr)
        outer_area = 2.0 * np.pi * (r + 0.5 * dr)
        vol = np.pi * ((r + 0.5 * dr) ** 2 - (r - 0.5 * dr) ** 2)
        self.r = r
        self.Lambda = np.sqrt(vol)
        nm = 1 if mmax == 0 else 2 * mmax
        self.vals = np.zeros((Nr, nm))
        Hi = np.zeros((Nr, Nk, nm))
        for m in range(mmax + 1):
            sub = inner_area / (dr * vol)
            diag = -(inner_area + outer_area) / (dr * vol) - (m / r) ** 2
            sup = outer_area / (dr * vol)
            diag[0] += sub[0]
            diag[Nr - 1] -= sup[Nr - 1]
            # similarity transform by the root volumes makes the operator symmetric
            sub[1:] *= self.Lambda[1:] / self.Lambda[:-1]
            sup[:-1] *= self.Lambda[:-1] / self.Lambda[1:]
            band = np.zeros((2, Nr))
            band[0, :] = sub          # upper-band storage: row 0 holds the off-diagonal (first entry unused)
            band[1, :] = diag
            vals, vecs = scipy.linalg.eig_banded(band, select='i', select_range=(Nr - Nk, Nr - 1))
            self.vals[:Nk, m] = vals
            Hi[:, :, m] = vecs
        for m in range(1, mmax):
            self.vals[:Nk, -m] = self.vals[:Nk, m]
            Hi[:, :, -m] = Hi[:, :, m]
        self.H = np.ascontiguousarray(Hi.swapaxes(0, 1))

    def kr2(self):
        return -self.vals

    def tridiagonal(self, m=0):
        '''Return (diag, offdiag) of the symmetrised operator for azimuthal mode m (for checks).'''
        Nr, dr = self.Nr, self.dr
        r = self.r
        inner_area = 2.0 * np.pi * (r - 0.5 * dr)
        outer_area = 2.0 * np.pi * (r + 0.5 * dr)
        vol = self.Lambda ** 2
        sub = inner_area / (dr * vol)
        diag = -(inner_area + outer_area) / (dr * vol) - (m / r) ** 2
        sup = outer_area / (dr * vol)
        diag[0] += sub[0]
        diag[Nr - 1] -= sup[Nr - 1]
        off = sup[:-1] * self.Lambda[:-1] / self.Lambda[1:]
        return diag, off

    def kspace_host(self, a):
        '''FP64 host forward transform of a (Nw, Nr, 1) array (m=0 only), for tests and set-up.'''
        H = self.H[:, :, 0]
        lam = self.Lambda
        out = np.zeros(a.shape, dtype=np.complex128)
        out[:, :self.Nk, 0] = np.einsum('kj,fj->fk', H, a[:, :, 0] * lam[None, :]) / lam[None, :self.Nk]
        return out

    def rspace_host(self, a):
        '''FP64 host inverse transform of a (Nw, Nr, 1) array (m=0 only), for tests and set-up.'''
        H = self.H[:, :, 0]
        lam = self.Lambda
        out = np.einsum('kj,fk->fj', H, a[:, :self.Nk, 0] * lam[None, :self.Nk]) / lam[None, :]
        return out[:, :, None]
