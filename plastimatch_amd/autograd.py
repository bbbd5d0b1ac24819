"""The MSE score of a B-spline stage as a differentiable torch function of
its coefficient vector (the extra/python binding of the HIP flavor).

The backward pass is the library's explicit adjoint of the warp-and-score
forward: the exact trilinear derivative dc/dv (flavor p semantics,
src/register/bspline_mse.cxx:1149-1384, src/base/interpolate.cxx:162-244),
condensed onto the knots and reduced in FP64, scaled by 2/N
(bspline_score_normalize, src/register/bspline_mse.cxx:43-73).  Two ways to
get the warp into the backward pass are offered:

* ``"recompute"`` (default): the forward pass evaluates only the score; the
  backward pass interpolates the displacement again from the coefficients.
  Nothing beyond the coefficients is kept between the passes.
* ``"stored"``: the forward pass also writes the dense displacement field
  (12 bytes per fixed voxel, bspline_interpolate_vf semantics) and the
  backward pass reads it instead of interpolating.

Both return the same gradient up to float rounding; the choice trades
memory (the stored field) against the backward pass's interpolation.
"""
import torch

from .bspline import BsplineHip

MODES = ("recompute", "stored")


def _device_scalar(dtype, dev):
    return torch.zeros(1, dtype=dtype, device=dev)


class BsplineMSEFunction(torch.autograd.Function):
    """score = BsplineMSEFunction.apply(coeff, stage, mode)

    coeff: 1-D float32 or float64 tensor on the stage's GPU, 3 * num_knots
    entries, interleaved coeff[3 * knot + d].  stage: a BsplineHip.  Returns
    a 0-dim float64 tensor (the MSE normalized by the overlap count N);
    coeff.grad comes back in coeff's dtype.
    """

    @staticmethod
    def forward(ctx, coeff, stage, mode="recompute"):
        if not isinstance(stage, BsplineHip):
            raise TypeError("stage must be a BsplineHip")
        if mode not in MODES:
            raise ValueError(f"mode must be one of {MODES}")
        if coeff.dim() != 1 or coeff.numel() != stage.bxf.num_coeff:
            raise ValueError("coeff must be a vector of num_coeff entries")
        if not coeff.is_cuda:
            raise ValueError("coeff must live on the GPU")
        dev = coeff.device
        c32 = coeff.detach().to(torch.float32).contiguous()
        stream = torch.cuda.current_stream(dev).cuda_stream
        score = _device_scalar(torch.float64, dev)
        nvox = _device_scalar(torch.int64, dev)
        vf = None
        if mode == "stored":
            nf = stage.fixed_voxels
            vf = torch.empty(3 * nf, dtype=torch.float32, device=dev)
            stage.interpolate_vf(c32, vf, stream=stream)
        stage.score_device(c32, score, None, nvox, "mse", "none", stream=stream)
        ctx.stage = stage
        ctx.mode = mode
        ctx.coeff_dtype = coeff.dtype
        ctx.save_for_backward(c32, vf if vf is not None else c32.new_empty(0))
        ctx.nvox = nvox
        return score.reshape(())

    @staticmethod
    def backward(ctx, grad_out):
        c32, vf = ctx.saved_tensors
        stage = ctx.stage
        dev = c32.device
        stream = torch.cuda.current_stream(dev).cuda_stream
        grad = torch.empty(c32.numel(), dtype=torch.float64, device=dev)
        score = _device_scalar(torch.float64, dev)
        nvox = _device_scalar(torch.int64, dev)
        if ctx.mode == "stored":
            stage.score_vf_device(vf, score, grad, nvox, stream=stream)
        else:
            stage.score_device(c32, score, grad, nvox, "mse", "exact", stream=stream)
        g = grad * grad_out.to(torch.float64)
        return g.to(ctx.coeff_dtype), None, None


def mse(coeff, stage, mode="recompute"):
    """Differentiable MSE score of ``stage`` at ``coeff`` (see BsplineMSEFunction)."""
    return BsplineMSEFunction.apply(coeff, stage, mode)
