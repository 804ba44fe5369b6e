"""Mesh coordinates on the GPU (mirrors pyminiweather/mesh.py:5-112).

Same getters and caching as the reference's ``MeshData``; arrays are float64
torch tensors built like numpy.linspace (i * step + start), then meshgrid'ed
with x fastest.
"""
import torch

from .data import linspace


class MeshData:
    def __init__(self, params, device=None):
        self.params = params
        self.xlen = params["xlen"]
        self.zlen = params["zlen"]
        self.dx = params["dx"]
        self.dz = params["dz"]
        self.nx = params["nx"]
        self.nz = params["nz"]
        self.hs = params["hs"]
        self.device = device if device is not None else \
            torch.device("cuda", torch.cuda.current_device())
        self.mesh_int_ext = None
        self.mesh_cell_centers = None
        self.mesh_vertical_cell_edges = None
        self.mesh_vertical_cell_centers_int_ext = None

    # ---- (start, stop, num) of each coordinate, as numpy.linspace takes them
    def _spec_int_ext(self):
        hs, dx, dz = self.hs, self.dx, self.dz
        return ((-hs * dx, (self.nx + hs) * dx, self.nx + 2 * hs),
                (-hs * dz, (self.nz + hs) * dz, self.nz + 2 * hs))

    def _spec_cell_centers(self):
        return ((self.dx / 2.0, self.xlen + self.dx / 2.0, self.nx),
                (self.dz / 2.0, self.zlen + self.dz / 2.0, self.nz))

    def _spec_vertical_cell_edges(self):
        return (0.0, (self.nz + 1) * self.dz, self.nz + 1)

    def _spec_vertical_cell_centers_int_ext(self):
        hs, dz = self.hs, self.dz
        return ((-hs + 0.5) * dz, (self.nz + hs + 0.5) * dz, self.nz + 2 * hs)

    def linspace_geometry(self):
        """The 16 doubles pmw_init takes: xlen, zlen, dx, dz, then the
        (start, step) of each coordinate family."""
        def ss(spec):
            start, stop, num = spec
            return [float(start), (float(stop) - float(start)) / num]
        xe, ze = self._spec_int_ext()
        xc, zc = self._spec_cell_centers()
        return ([float(self.xlen), float(self.zlen), float(self.dx), float(self.dz)]
                + ss(xe) + ss(ze) + ss(self._spec_vertical_cell_centers_int_ext())
                + ss(self._spec_vertical_cell_edges()) + ss(xc) + ss(zc))

    def _grid(self, xs, zs):
        x = linspace(*xs, self.device)
        z = linspace(*zs, self.device)
        zz, xx = torch.meshgrid(z, x, indexing="ij")
        return [xx, zz]

    def get_mesh_int_ext(self):
        """Cell left edges incl. ghost cells (mesh.py:25-51)."""
        if self.mesh_int_ext is None:
            self.mesh_int_ext = self._grid(*self._spec_int_ext())
        return self.mesh_int_ext

    def get_mesh_cell_centers(self):
        """Interior cell centres (mesh.py:53-75)."""
        if self.mesh_cell_centers is None:
            self.mesh_cell_centers = self._grid(*self._spec_cell_centers())
        return self.mesh_cell_centers

    def get_mesh_vertical_cell_edges(self):
        """z of the interior cell edges (mesh.py:77-92)."""
        if self.mesh_vertical_cell_edges is None:
            self.mesh_vertical_cell_edges = linspace(
                *self._spec_vertical_cell_edges(), self.device)
        return self.mesh_vertical_cell_edges

    def get_mesh_vertical_cell_centers_int_ext(self):
        """z of the cell centres incl. ghost cells (mesh.py:94-112)."""
        if self.mesh_vertical_cell_centers_int_ext is None:
            self.mesh_vertical_cell_centers_int_ext = linspace(
                *self._spec_vertical_cell_centers_int_ext(), self.device)
        return self.mesh_vertical_cell_centers_int_ext
