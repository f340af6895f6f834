"""Differentiable dose from a dose-influence matrix held on the device.

A Dij built by ``Context.dij_build`` is a CSR matrix with one row per spot
(the orientation of the reference's ``io::save_to_npz``,
moqui/base/mqi_io.hpp:249-320).  With D the voxels x spots influence
matrix, the dose of spot weights w is D w and the gradient of an objective
f(D w) with respect to w is D^T (df/dd).  Both products run in the
library's kernels (FP64 accumulation, float32 inputs, fixed summation
order) on torch's current stream; ``dose`` wraps them in a
``torch.autograd.Function``, and ``quadratic_objective`` forms the
quadratic objective and its dose gradient inside the D w pass.

    mat = DijMatrix(ctx, scenario=0)
    w = torch.rand(mat.n_spots, device="cuda", requires_grad=True)
    f = quadratic_objective(mat, w, prescription)   # fused: one pass, no dose vector
    f.backward()                      # w.grad = 2 D^T (D w - p) / n

Robust optimisation: ``worst_case_objective`` over the Dij of several
scenarios (max over scenarios, subgradient of the worst one).  A scenario
run with a history budget factor b holds b times the dose per unit weight
(the tables sum over histories, as the reference's unit-weight Dij does);
``DijMatrix(ctx, s, scale=1 / b)`` puts every scenario on one scale.
"""
import torch


class DijMatrix:
    """Handle on the device CSR of one scenario of a context's Dij; the
    products use scale * D."""

    def __init__(self, ctx, scenario=0, scale=1.0):
        self.ctx = ctx
        self.scenario = int(scenario)
        self.scale = float(scale)
        self.nnz, self.n_spots, self.n_voxels = ctx.dij_info(self.scenario)

    def _vec(self, x, n):
        if x.dim() != 1 or x.numel() != n:
            raise ValueError("expected a vector of %d entries, got shape %s" % (n, tuple(x.shape)))
        if not x.is_cuda:
            raise ValueError("the sparse products take device tensors")
        return x.detach().to(torch.float32).contiguous()

    def _product(self, transpose, x, n_out, out_dtype):
        y = torch.empty(n_out, dtype=torch.float64, device=x.device)
        stream = torch.cuda.current_stream(x.device).cuda_stream
        self.ctx.dij_matvec(self.scenario, int(transpose), x.data_ptr(), y.data_ptr(), y_double=True,
                            alpha=self.scale, stream=stream)
        return y if out_dtype == torch.float64 else y.to(out_dtype)

    def matvec(self, w, out_dtype=torch.float64):
        """Dose scale * D w (n_voxels)."""
        return self._product(False, self._vec(w, self.n_spots), self.n_voxels, out_dtype)

    def quadratic(self, w, prescription, voxel_weights=None, want_grad=True):
        """f = mean_v c_v (scale (D w)_v - p_v)^2 as a device scalar (float64)
        and, with want_grad, df/dd = 2 c (d - p) / n_voxels (float32, the
        input of rmatvec), in one pass over the matrix (mqi_dij_quadratic)."""
        w = self._vec(w, self.n_spots)
        p = prescription.detach().to(device=w.device, dtype=torch.float64).contiguous()
        if p.dim() != 1 or p.numel() != self.n_voxels:
            raise ValueError("expected a prescription of %d voxels, got shape %s" % (self.n_voxels, tuple(p.shape)))
        c = None
        if voxel_weights is not None:
            c = voxel_weights.detach().to(device=w.device, dtype=torch.float32).contiguous()
            if c.dim() != 1 or c.numel() != self.n_voxels:
                raise ValueError("expected voxel weights of %d voxels" % self.n_voxels)
        f = torch.empty((), dtype=torch.float64, device=w.device)
        g = torch.empty(self.n_voxels, dtype=torch.float32, device=w.device) if want_grad else None
        stream = torch.cuda.current_stream(w.device).cuda_stream
        self.ctx.dij_quadratic(self.scenario, w.data_ptr(), p.data_ptr(), c.data_ptr() if c is not None else 0,
                               self.scale, g.data_ptr() if g is not None else 0, f.data_ptr(), stream=stream)
        return f, g

    def rmatvec(self, g, out_dtype=torch.float64):
        """scale * D^T g (n_spots)."""
        return self._product(True, self._vec(g, self.n_voxels), self.n_spots, out_dtype)


class _Dose(torch.autograd.Function):
    @staticmethod
    def forward(fctx, w, mat):
        fctx.mat = mat
        fctx.w_dtype = w.dtype
        return mat.matvec(w, torch.float64)

    @staticmethod
    def backward(fctx, grad_dose):
        grad_w = fctx.mat.rmatvec(grad_dose, torch.float64)
        return grad_w.to(fctx.w_dtype), None


def dose(mat, w):
    """Differentiable dose D w (float64, n_voxels) of spot weights w."""
    return _Dose.apply(w, mat)


class _Quadratic(torch.autograd.Function):
    @staticmethod
    def forward(fctx, w, mat, prescription, voxel_weights):
        f, g = mat.quadratic(w, prescription, voxel_weights, want_grad=fctx.needs_input_grad[0])
        fctx.mat = mat
        fctx.w_dtype = w.dtype
        if g is not None:
            fctx.save_for_backward(g)
        return f

    @staticmethod
    def backward(fctx, grad_f):
        (g,) = fctx.saved_tensors
        grad_w = fctx.mat.rmatvec(g, torch.float64) * grad_f
        return grad_w.to(fctx.w_dtype), None, None, None


def quadratic_objective(mat, w, prescription, voxel_weights=None):
    """f(w) = mean_v c_v (D w - p)_v^2 (c = 1 when voxel_weights is None), in
    one pass over the matrix (the dose is not formed); the gradient is
    D^T (2 c (D w - p) / n)."""
    return _Quadratic.apply(w, mat, prescription, voxel_weights)


def worst_case_objective(mats, w, prescription, voxel_weights=None):
    """max over scenarios of the quadratic objective (differentiable: the
    gradient is the worst scenario's), and the per-scenario values (a
    detached vector).  The values are evaluated without a graph and only the
    worst scenario's objective is formed again with one, so a backward pass
    runs one D^T g and not one per scenario (a graph over every scenario
    would carry zero gradients through the others).  The products are
    deterministic: the returned maximum equals its entry in the values."""
    with torch.no_grad():
        vals = torch.stack([m.quadratic(w, prescription, voxel_weights, want_grad=False)[0] for m in mats])
    k = int(torch.argmax(vals).item())
    return quadratic_objective(mats[k], w, prescription, voxel_weights), vals
