'''Grid geometry of the UPPE wave region (host side).

Restates the grid-information half of SeaRay's ``modules/caustic_tools.py``: ``FourierTool`` (Cartesian,
caustic_tools.py:82-96) and ``BesselBeamTool`` (cylindrical, caustic_tools.py:207-229). The ray coupling
methods (``GetBoundaryFields``, ``RelaunchRays``) are not part of the UPPE stepper and are not provided.
'''
import numpy as np
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
