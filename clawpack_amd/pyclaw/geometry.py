"""Grid containers (upstream pyclaw/src/pyclaw/geometry.py): Dimension, Grid, Patch, Domain.

Uniform Cartesian grids only: delta = (upper - lower) / num_cells, cell i has centre
lower + (i + 1/2) * delta.
"""
import numpy as np


class Dimension:
    def __init__(self, lower, upper, num_cells, name="x", on_lower_boundary=None,
                 on_upper_boundary=None, units=None):
        if upper <= lower:
            raise ValueError("Dimension: upper must be greater than lower")
        if int(num_cells) < 1:
            raise ValueError("Dimension: num_cells must be positive")
        self.name = name
        self.lower = float(lower)
        self.upper = float(upper)
        self.num_cells = int(num_cells)
        self.on_lower_boundary = True if on_lower_boundary is None else on_lower_boundary
        self.on_upper_boundary = True if on_upper_boundary is None else on_upper_boundary
        self.units = units

    @property
    def delta(self):
        return (self.upper - self.lower) / self.num_cells

    @property
    def centers(self):
        return self.lower + (np.arange(self.num_cells) + 0.5) * self.delta

    @property
    def nodes(self):
        return self.lower + np.arange(self.num_cells + 1) * self.delta

    def __repr__(self):
        return "Dimension(%s: [%g, %g], %d cells)" % (self.name, self.lower, self.upper, self.num_cells)


class Grid:
    def __init__(self, dimensions):
        if isinstance(dimensions, Dimension):
            dimensions = [dimensions]
        self.dimensions = list(dimensions)
        for d in self.dimensions:
            setattr(self, d.name, d)

    @property
    def num_dim(self):
        return len(self.dimensions)

    @property
    def num_cells(self):
        return [d.num_cells for d in self.dimensions]

    @property
    def lower(self):
        return [d.lower for d in self.dimensions]

    @property
    def upper(self):
        return [d.upper for d in self.dimensions]

    @property
    def delta(self):
        return [d.delta for d in self.dimensions]

    @property
    def on_lower_boundary(self):
        return [d.on_lower_boundary for d in self.dimensions]

    @property
    def on_upper_boundary(self):
        return [d.on_upper_boundary for d in self.dimensions]

    @property
    def c_centers(self):
        """Cell centres as an 'ij' meshgrid (computational == physical for this grid)."""
        return np.meshgrid(*[d.centers for d in self.dimensions], indexing="ij")

    @property
    def p_centers(self):
        return self.c_centers

    @property
    def c_nodes(self):
        return np.meshgrid(*[d.nodes for d in self.dimensions], indexing="ij")

    @property
    def p_nodes(self):
        return self.c_nodes


class Patch:
    def __init__(self, dimensions):
        self.grid = Grid(dimensions)
        self.dimensions = self.grid.dimensions
        self.level = 1
        self.patch_index = 1

    def __getattr__(self, name):
        if name in ("grid", "dimensions"):
            raise AttributeError(name)
        return getattr(self.grid, name)


class Domain:
    """Domain(dimension_list) or Domain(lower, upper, num_cells) as in pyclaw."""

    def __init__(self, *args):
        if len(args) == 1:
            geom = args[0]
            if isinstance(geom, Patch):
                self.patches = [geom]
            else:
                self.patches = [Patch(geom)]
        elif len(args) == 3:
            lower, upper, num_cells = args
            names = ["x", "y", "z"][: len(lower)]
            dims = [Dimension(lo, up, n, name=nm) for lo, up, n, nm in zip(lower, upper, num_cells, names)]
            self.patches = [Patch(dims)]
        else:
            raise ValueError("Domain(dimensions) or Domain(lower, upper, num_cells)")
        self.patch = self.patches[0]
        self.grid = self.patch.grid

    @property
    def num_dim(self):
        return self.grid.num_dim
