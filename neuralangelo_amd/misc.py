"""Drop-in SDF regularisers (projects/neuralangelo/utils/misc.py:74-89) on a HIP kernel."""
from . import functional as Fn


def _n_per_ray(x, outside):
    if outside is None:
        return 1
    return x[..., 0].numel() // max(outside.numel(), 1)


def eikonal_loss(gradients, outside=None):
    """mean((|g| - 1)^2 * !outside) over all samples, NaN/Inf -> 0."""
    g = gradients.reshape(-1, 3)
    eik, _ = Fn.SDFLosses.apply(g, None, outside.reshape(-1) if outside is not None else None,
                                _n_per_ray(gradients, outside))
    return eik


def curvature_loss(hessian, outside=None):
    """mean(|sum_d hess_d| * !outside) over all samples, NaN/Inf -> 0."""
    h = hessian.reshape(-1, 3)
    _, curv = Fn.SDFLosses.apply(h, h, outside.reshape(-1) if outside is not None else None,
                                 _n_per_ray(hessian, outside))
    return curv


def sdf_losses(gradients, hessians, outside):
    """Both regularisers in one pass: (eikonal, curvature)."""
    return Fn.SDFLosses.apply(gradients.reshape(-1, 3),
                              hessians.reshape(-1, 3) if hessians is not None else None,
                              outside.reshape(-1) if outside is not None else None,
                              _n_per_ray(gradients, outside))


def to_full_image(image, image_size=None, from_vec=True):
    """misc.py:110-117: [B,HW,...,K] (from_vec) or [B,H,W,...,K] -> [B,K,H,W,...]."""
    if from_vec:
        if image_size is None:
            raise ValueError("image_size is needed to unflatten the pixel axis")
        image = image.unflatten(1, tuple(image_size))
    return image.movedim(-1, 1)
