"""CpabSequential (libcpab/sequential.py:15-143): a chain of Cpab transforms.

Each stage integrates the previous stage's output grid.  Unlike the
reference, whose fast path returns no gradient w.r.t. the points
(libcpab/pytorch/transformer.py:202) and so trains only the last theta, every
stage here back-propagates dL/dpoints, so all thetas receive gradients.  The
stages do not keep trajectories or Jacobians: the backward of each stage
recomputes its trajectories from the stage's input grid.
"""
from .cpab import Cpab
from .pytorch.hip_ops import TransformGrid


class CpabSequential(object):
    def __init__(self, *cpab):
        self.n_cpab = len(cpab)
        self.cpab = cpab
        for i in range(self.n_cpab):
            assert isinstance(self.cpab[i], Cpab), "Class {0} is not a member of the cpab core class".format(i)
        self.ndim = self.cpab[0].params.ndim
        for i in range(1, self.n_cpab):
            assert self.ndim == self.cpab[i].params.ndim, (
                "Mismatching dimensionality of transformers. Transformer 1 have dimensionality {0} but "
                "transformer {1} have dimensionality {2}".format(self.ndim, i + 1, self.cpab[i].params.ndim))
        self.backend = self.cpab[0].backend
        self.backend_name = self.cpab[0].backend_name
        for i in range(1, self.n_cpab):
            assert self.backend_name == self.cpab[i].backend_name, "Mismatch in backend"

    def get_theta_dim(self):
        return [c.get_theta_dim() for c in self.cpab]

    def get_params(self):
        return [c.get_params() for c in self.cpab]

    def get_basis(self):
        return [c.get_basis() for c in self.cpab]

    def uniform_meshgrid(self, n_points):
        return self.cpab[0].uniform_meshgrid(n_points)

    def sample_transformation(self, n_sample, means=None, covs=None):
        means = self.n_cpab * [None] if means is None else means
        covs = self.n_cpab * [None] if covs is None else covs
        assert len(means) == self.n_cpab and len(covs) == self.n_cpab, \
            "The number of supplied means/covariances should be equal to the number of transformations"
        return [c.sample_transformation(n_sample, m, s) for c, m, s in zip(self.cpab, means, covs)]

    def identity(self, n_sample, epsilon=0):
        return [c.identity(n_sample, epsilon) for c in self.cpab]

    def _stage(self, i, grid, theta, grid_shape=None):
        c = self.cpab[i]
        c._check_type(grid)
        c._check_device(grid)
        c._check_type(theta)
        c._check_device(theta)
        if getattr(c.params, "method", "closed_form") != "closed_form" or c.params.use_slow:
            return c.transform_grid(grid, theta)
        return TransformGrid.apply(grid, theta, c.params, 1.0, False, grid_shape)

    def transform_grid(self, grid, thetas, output_all=False, _grid_shape=None):
        self._assert_theta_shape(thetas)
        if not output_all:
            for i in range(self.n_cpab):
                grid = self._stage(i, grid, thetas[i], _grid_shape)
            return grid
        grids = [self._stage(0, grid, thetas[0], _grid_shape)]
        for i in range(1, self.n_cpab):
            grids.append(self._stage(i, grids[-1], thetas[i], _grid_shape))
        return grids

    def transform_data(self, data, thetas, outsize, output_all=False):
        self._assert_theta_shape(thetas)
        grid = self.uniform_meshgrid(outsize)
        # every stage's points stay in meshgrid order: 2D stages walk them in tiles
        shape = list(outsize) if self.ndim == 2 else None
        grid_t = self.transform_grid(grid, thetas, output_all=output_all, _grid_shape=shape)
        if not output_all:
            return self.cpab[-1].interpolate(data, grid_t, outsize)
        return [self.cpab[i].interpolate(data, grid_t[i], outsize) for i in range(self.n_cpab)]

    def _assert_theta_shape(self, thetas):
        assert len(thetas) == self.n_cpab, "Number of parametrizations needed are {0}".format(self.n_cpab)
        batch = thetas[0].shape[0]
        for i in range(1, len(thetas)):
            assert batch == thetas[i].shape[0], "Batch size should be the same for all theta's"

    def __repr__(self):
        return "\n".join("======= Transformer {0} =======\n{1}".format(i + 1, c) for i, c in enumerate(self.cpab))
