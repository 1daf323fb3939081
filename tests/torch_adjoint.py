"""torch.autograd wrapper of the adjoint pass (test harness only).

Parameters: a 6-vector [Mandelbulb power, Mandelbox scale, Mandelbox min radius, camera
translation x, y, z]. Forward renders every view at those parameters and returns the depth
loss against a target; backward returns the loss gradient the HIP adjoint computed with the
implicit-function theorem (no march steps stored, no autograd tape through the march).
"""
import ctypes as C

import numpy as np
import torch

import numericdream_amd as nd


def apply_params(base_views, p):
    """Copies of the base Uniforms with the parameters applied."""
    out = []
    for b in base_views:
        u = nd.Uniforms()
        C.memmove(C.addressof(u), C.addressof(b), C.sizeof(u))
        bulb = box = None
        for i in range(max(int(u.instance_count), 1)):
            code = int(u.instances[i].code)
            if code == 0 and bulb is None:
                bulb = i
            if code == 1 and box is None:
                box = i
        if bulb is not None:
            u.instances[bulb].params[0] = float(p[0])
        if box is not None:
            u.instances[box].params[0] = float(p[1])
            u.instances[box].params[1] = float(p[2])
        for k in range(3):
            u.camera_pos[k] = b.camera_pos[k] + float(p[3 + k])
        out.append(u)
    return out


class DepthLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, renderer, base_views, target):
        views = apply_params(base_views, params.detach().cpu().numpy().astype(np.float64))
        loss, grad, used, _ = renderer.adjoint(views, target=target)
        ctx.save_for_backward(torch.tensor(grad, dtype=params.dtype, device=params.device))
        ctx.used = used
        return params.new_tensor(loss)

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g * grad, None, None, None


def render_depth(renderer, views, device):
    total = sum(int(v.resolution[0]) * int(v.resolution[1]) for v in views)
    depth = torch.zeros(total, dtype=torch.float32, device=device)
    renderer.adjoint(views, depth_out=depth)
    return depth
