"""The Cpab facade with the reference's API (libcpab/cpab.py:15-539).

Backend 'pytorch' with device 'gpu' runs every hot operation (integration,
its gradient, interpolation, the fused transform_data) as HIP kernels on the
MI355X through libcpab_hip.so.  Other backend/device combinations are outside
this build and raise NotImplementedError at construction.
"""
import os

import numpy as np

from .core.tessellation import make_tessellation


class params:
    """Plain attribute container (libcpab/core/utility.py:20-24)."""

    def __repr__(self):
        return str({k: v for k, v in self.__dict__.items() if not k.startswith("_")})


def _cache_dir():
    d = os.environ.get("LIBCPAB_AMD_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "libcpab_amd"))
    try:
        os.makedirs(d, exist_ok=True)
        return d
    except OSError:
        return None


def basis_file_name(nc, zero_boundary, volume_perservation):
    # same naming as libcpab/core/tesselation.py:60-65 (stored as .npz, no pickle)
    return ("cpab_basis_dim" + str(len(nc)) + "_tess" + "_".join(str(e) for e in nc) + "_" +
            "vo" + str(int(not zero_boundary)) + "_" + "zb" + str(int(zero_boundary)) + "_" +
            "vp" + str(int(volume_perservation)))


def load_or_build_basis(nc, zero_boundary, volume_perservation, override=False):
    """L, B and the facet table of a tessellation, cached as .npz."""
    d = _cache_dir()
    path = os.path.join(d, basis_file_name(nc, zero_boundary, volume_perservation) + ".npz") if d else None
    if path and os.path.exists(path) and not override:
        with np.load(path, allow_pickle=False) as f:
            cached = {k: f[k] for k in f.files}
        if "amap" in cached:
            return cached
    tess = make_tessellation(nc, zero_boundary=zero_boundary, volume_perservation=volume_perservation)
    planes, nbr = tess.facet_table()
    out = {"L": tess.L, "B": tess.B, "planes": planes, "nbr": nbr, "amap": tess.amap,
           "centers": tess.get_cell_centers()}
    if path:
        tmp = path + ".%d.tmp.npz" % os.getpid()
        np.savez(tmp, **out)
        os.replace(tmp, path)
    return out


class Cpab(object):
    """Continuous piecewise-affine transformations (libcpab/cpab.py:15-52).

    Arguments:
        tess_size: list, number of cells per dimension (1, 2 or 3 entries)
        backend: 'numpy' (the reference's default), 'tensorflow' or 'pytorch';
            this build provides 'pytorch' only
        device: 'cpu' (the reference's default) or 'gpu'; this build provides
            'gpu' only, so Cpab(tess_size) with the defaults raises
            NotImplementedError as the other unprovided combinations do
        zero_boundary: bool, zero normal velocity on the domain boundary
        volume_perservation: bool, trace-free cell fields
        override: bool, rebuild the cached basis
    """

    def __init__(self, tess_size, backend="numpy", device="cpu", zero_boundary=True,
                 volume_perservation=False, override=False, basis=None):
        self._check_input(tess_size, backend, device, zero_boundary, volume_perservation, override)
        p = params()
        p.nc = list(tess_size)
        p.ndim = len(tess_size)
        p.Ashape = [p.ndim, p.ndim + 1]
        p.valid_outside = not zero_boundary
        p.zero_boundary = zero_boundary
        p.volume_perservation = volume_perservation
        p.domain_max = [1 for _ in p.nc]
        p.domain_min = [0 for _ in p.nc]
        p.inc = [(p.domain_max[i] - p.domain_min[i]) / p.nc[i] for i in range(p.ndim)]
        p.nstepsolver = 50
        p.numeric_grad = False
        p.use_slow = False
        # this build's solver knobs
        p.method = "closed_form"
        p.store_jacobian = True
        p.launch = "batched"
        p.raise_unfinished = False
        if p.ndim == 1:
            p.nC = p.nc[0]
            p.params_pr_cell = 2
        elif p.ndim == 2:
            p.nC = 4 * int(np.prod(p.nc))
            p.params_pr_cell = 6
        else:
            p.nC = 5 * int(np.prod(p.nc))
            p.params_pr_cell = 12
        geo = load_or_build_basis(p.nc, zero_boundary, volume_perservation, override)
        if basis is not None:
            basis = np.asarray(basis, dtype=np.float64)
            assert basis.shape[0] == geo["B"].shape[0], "basis must have D = %d rows" % geo["B"].shape[0]
            geo = dict(geo, B=basis)
        p.constrain_mat = geo["L"]
        p.basis = geo["B"]
        p.facet_planes = geo["planes"]
        p.facet_nbr = geo["nbr"]
        # 3D without zero boundary: the simplex each region outside the domain uses
        p.facet_amap = np.asarray(geo["amap"], dtype=np.int32)
        p.D, p.d = p.basis.shape
        self._centers = geo["centers"]
        self.params = p
        self.backend_name = backend
        from .pytorch import functions as backend_module
        self.backend = backend_module
        self.device = device.lower()
        self.backend.assert_version()
        from . import _native
        _native.lib()  # fail now, not at the first call, if the HIP library is missing

    # -- parameters ---------------------------------------------------------
    def get_theta_dim(self):
        return self.params.d

    def get_theta_size(self):
        """Alias advertised by the reference README (README.md:101)."""
        return self.params.d

    def get_params(self):
        return self.params

    def get_basis(self):
        return self.params.basis

    def unfinished_points(self, reset=True):
        """Points that used up the closed-form integrator's chunk budget (512
        chunks and facet crossings per point; their outputs and gradients are
        NaN) in this transformer's forward calls since the last reset.  Reads
        device counters (a synchronisation).  set_solver_params(
        raise_unfinished=True) checks after every forward call instead and
        raises CpabError."""
        from .pytorch.hip_ops import unfinished_points
        return unfinished_points(self.params, reset)

    def set_solver_params(self, nstepsolver=50, numeric_grad=False, use_slow=False, method=None,
                          store_jacobian=None, launch=None, raise_unfinished=None):
        """libcpab/cpab.py:140-157.  use_slow=True selects the reference's
        slow path (libcpab/pytorch/transformer.py:90-128): nstepsolver steps
        of expm(A~/nstepsolver) with the torch cell lookup, differentiated
        exactly w.r.t. theta and the points (HIP kernels for the steps and
        their adjoint, torch autograd through matrix_exp).  Otherwise
        method='numeric' runs the reference's compiled fixed-step path
        (nstepsolver steps; numeric_grad=True for its finite-difference
        gradient) and the default method='closed_form' integrates exactly,
        ignoring nstepsolver and numeric_grad.  Beyond the reference:
        store_jacobian (closed form: keep the forward's Jacobians for the
        backward), launch ('batched' or 'plain', one launch sequence per
        theta) and raise_unfinished (check every forward for points that
        used up the chunk budget and raise CpabError)."""
        assert nstepsolver > 0, "nstepsolver must be a positive number"
        assert type(nstepsolver) == int, "nstepsolver must be integer"
        assert type(numeric_grad) == bool, "numeric_grad must be bool"
        assert type(use_slow) == bool, "use_slow must be bool"
        self.params.nstepsolver = nstepsolver
        self.params.numeric_grad = numeric_grad
        self.params.use_slow = use_slow
        if method is not None:
            assert method in ("closed_form", "numeric"), "method must be 'closed_form' or 'numeric'"
            self.params.method = method
        if store_jacobian is not None:
            self.params.store_jacobian = bool(store_jacobian)
        if launch is not None:
            assert launch in ("batched", "plain"), "launch must be 'batched' or 'plain'"
            self.params.launch = launch
        if raise_unfinished is not None:
            self.params.raise_unfinished = bool(raise_unfinished)

    # -- sampling -----------------------------------------------------------
    def uniform_meshgrid(self, n_points):
        return self.backend.uniform_meshgrid(self.params.ndim, self.params.domain_min, self.params.domain_max,
                                             n_points, self.device)

    def sample_transformation(self, n_sample=1, mean=None, cov=None):
        if mean is not None:
            self._check_type(mean)
            self._check_device(mean)
        if cov is not None:
            self._check_type(cov)
            self._check_device(cov)
        samples = self.backend.sample_transformation(self.params.d, n_sample, mean, cov, self.device)
        return self.backend.to(samples, device=self.device)

    def sample_transformation_with_prior(self, n_sample=1, mean=None, length_scale=0.1, output_variance=1):
        """Smooth prior over cell parameters (libcpab/cpab.py:192-241): samples
        of N(mean, prior_covariance(length_scale, output_variance))."""
        import torch
        cov_theta = self.prior_covariance(length_scale, output_variance).to(torch.float32)
        dev = self.backend._device(self.device)
        return self.sample_transformation(n_sample, mean=mean, cov=cov_theta.to(dev))

    def prior_covariance(self, length_scale=0.1, output_variance=1):
        """The prior's covariance in theta space [d, d] (float64, host): squared
        exponential kernel over squared cell-centre distances on the diagonal
        of each cell block of the [D, D] parameter covariance, 100 x the
        largest distance elsewhere, pushed to theta space as B^T C B
        (libcpab/cpab.py:211-234, built here without the nC^2 block loop)."""
        import torch
        p = self.params
        centers = torch.tensor(self._centers, dtype=torch.float64)
        dist = self.backend.pdist(centers)
        ppc = p.params_pr_cell
        big = 100 * dist.max()
        cov_init = torch.full((p.D, p.D), float(big), dtype=torch.float64)
        # block (i, j) has dist[i, j] on its diagonal
        eye = torch.eye(ppc, dtype=torch.bool)
        blocks = cov_init.view(p.nC, ppc, p.nC, ppc).permute(0, 2, 1, 3)
        blocks[:, :, eye] = dist[:, :, None].expand(p.nC, p.nC, ppc)
        cov_init = blocks.permute(0, 2, 1, 3).reshape(p.D, p.D)
        cov_avees = output_variance ** 2 * torch.exp(-(cov_init / (2 * length_scale ** 2)))
        B = torch.tensor(p.basis, dtype=torch.float64)
        return B.t() @ cov_avees @ B

    def identity(self, n_sample=1, epsilon=0):
        return self.backend.identity(self.params.d, n_sample, epsilon, self.device)

    # -- the hot path --------------------------------------------------------
    def transform_grid(self, grid, theta):
        """Integrate grid [ndim, nP] or [n_theta, ndim, nP] under theta [n_theta, d]
        (libcpab/cpab.py:257-277).  Returns [n_theta, ndim, nP]."""
        self._check_type(grid)
        self._check_device(grid)
        self._check_type(theta)
        self._check_device(theta)
        if len(grid.shape) == 3:
            assert grid.shape[0] == theta.shape[0], (
                "When passing a 3D grid, expects the first dimension to be of same length as the first "
                "dimension of theta")
        return self.backend.transformer(grid, theta, self.params)

    def interpolate(self, data, grid, outsize):
        """Linear interpolation of data [n_batch, C, *size] at grid [n_batch, ndim, nP]
        (libcpab/cpab.py:280-301).  Returns [n_batch, C, *outsize]."""
        self._check_type(data)
        self._check_device(data)
        self._check_type(grid)
        self._check_device(grid)
        return self.backend.interpolate(self.params.ndim, data, grid, outsize)

    def transform_data(self, data, theta, outsize):
        """uniform_meshgrid -> transform_grid -> interpolate (libcpab/cpab.py:304-331),
        fused into one kernel per direction for the closed-form method."""
        self._check_type(data)
        self._check_device(data)
        self._check_type(theta)
        self._check_device(theta)
        if getattr(self.params, "method", "closed_form") == "closed_form" and not self.params.use_slow:
            return self.backend.transform_data(data, theta, self.params, outsize)
        grid = self.uniform_meshgrid(outsize)
        grid_t = self.transform_grid(grid, theta)
        return self.interpolate(data, grid_t, outsize)

    def calc_vectorfield(self, grid, theta):
        self._check_type(grid)
        self._check_device(grid)
        self._check_type(theta)
        self._check_device(theta)
        return self.backend.calc_vectorfield(grid, theta, self.params)

    def visualize_vectorfield(self, *args, **kwargs):
        raise NotImplementedError("plotting (libcpab/cpab.py:349-476) is outside this build")

    visualize_deformgrid = visualize_vectorfield
    visualize_tesselation = visualize_vectorfield

    # -- checks ---------------------------------------------------------------
    def _check_input(self, tess_size, backend, device, zero_boundary, volume_perservation, override):
        # libcpab/cpab.py:479-502
        assert 0 < len(tess_size) <= 3, "Transformer only supports 1D, 2D or 3D"
        assert type(tess_size) in (list, tuple), "Argument tess_size must be a list or tuple"
        assert all(type(e) == int for e in tess_size), "All elements of tess_size must be integers"
        assert all(e > 0 for e in tess_size), "All elements of tess_size must be positive"
        assert backend in ["numpy", "tensorflow", "pytorch"], \
            "Unknown backend, choose between 'numpy', 'tensorflow' or 'pytorch'"
        assert device in ["cpu", "gpu"], "Unknown device, choose between 'cpu' or 'gpu'"
        if backend == "numpy":
            assert device == "cpu", "Cannot use gpu with numpy backend"
        assert type(zero_boundary) == bool, "Argument zero_boundary must be True or False"
        assert type(volume_perservation) == bool, "Argument volume_perservation must be True or False"
        assert type(override) == bool, "Argument override must be True or False"
        if backend != "pytorch" or device != "gpu":
            raise NotImplementedError("this build provides backend='pytorch', device='gpu' only (got backend=%r, "
                                      "device=%r)" % (backend, device))

    def _check_type(self, x):
        assert isinstance(x, self.backend.backend_type()), \
            "Input has type {0} but expected type {1}".format(type(x), self.backend.backend_type())

    def _check_device(self, x):
        assert self.backend.check_device(x, self.device), \
            "Input is placed on device {0} but the class expects it to be on device {1}".format(
                str(x.device), self.device)

    def __repr__(self):
        p = self.params
        return ("\n        CPAB transformer class. \n            Parameters:\n"
                "                Tesselation size:           {0}\n"
                "                Total number of cells:      {1}\n"
                "                Theta size:                 {2}\n"
                "                Domain lower bound:         {3}\n"
                "                Domain upper bound:         {4}\n"
                "                Zero Boundary:              {5}\n"
                "                Volume perservation:        {6}\n"
                "            Backend:                        {7}\n").format(
            p.nc, p.nC, p.d, p.domain_min, p.domain_max, p.zero_boundary, p.volume_perservation, self.backend_name)
