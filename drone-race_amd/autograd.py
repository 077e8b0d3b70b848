"""torch.autograd.Function over the engine's adjoint rollout kernel.

RolloutCost.apply(U, engine, instance, ckpt_every) returns J (one cost per
control sequence U[n][T][4], a float32 CUDA tensor); backward multiplies the
incoming gradient into dJ/dU, which the forward call already produced with
the hand-written reverse-mode kernel (checkpoints every ckpt_every steps).
Both run on torch's current stream.
"""
import torch


class RolloutCost(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, engine, instance, ckpt_every=10):
        T = engine.horizon(instance)
        if not U.is_cuda or U.dtype != torch.float32 or U.dim() != 3 or U.shape[0] < 1 or tuple(U.shape[1:]) != (T, 4):
            raise ValueError(f"U must be a float32 CUDA tensor of shape [n, {T}, 4] for instance {instance}")
        U = U.contiguous()
        n = U.shape[0]
        J = torch.empty(n, dtype=torch.float32, device=U.device)
        G = torch.empty_like(U)
        stream = torch.cuda.current_stream(U.device).cuda_stream
        engine.cost_grad_device(instance, U.data_ptr(), n, U.numel(), J.data_ptr(), G.data_ptr(), ckpt_every, stream)
        ctx.save_for_backward(G)
        return J

    @staticmethod
    def backward(ctx, gJ):
        (G,) = ctx.saved_tensors
        return gJ.reshape(-1, 1, 1) * G, None, None, None


def rollout_cost(U, engine, instance=0, ckpt_every=10):
    return RolloutCost.apply(U, engine, instance, ckpt_every)
