"""PyTorch binding of the differentiable tick: final agent positions as a
function of per-road speed multipliers and the soft-avoidance strength, with
the hand-written adjoint of libmsim (msim_batch_diff_*) as its backward.

    run = DiffRun(batch, ticks=1000, checkpoint_every=50)
    pos = run(theta, alpha)          # theta: [n_roads] cuda, alpha: 0-d cuda
    loss = ((pos - target) ** 2).mean(); loss.backward()

theta holds one multiplier per packed road of the batch (scenarios sharing a
map tie theirs with theta_map.repeat(n_scen)).  Every call simulates from the
batch's state at setup.  The gradient freezes the run's decisions (arrival
ticks, route choices) and neighbour cells.  arrival_gradient="carry"
differentiates each arrival as if it carried its leftover time onto the next
road (a surrogate: with "frozen", the snap to the target cuts the gradient).
"""
import torch


class _DiffFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, alpha, batch):
        th = theta.detach().to(torch.float32).contiguous()
        out = torch.empty(batch.n, 2, device=th.device, dtype=torch.float32)
        torch.cuda.synchronize(th.device)
        batch.diff_forward(th.data_ptr(), float(alpha.detach().item()), out.data_ptr())
        ctx.batch = batch
        ctx.generation = batch.diff_generation()
        return out

    @staticmethod
    def backward(ctx, grad_pos):
        b = ctx.batch
        if b.diff_generation() != ctx.generation:
            # the handle keeps one run's checkpoints: the adjoint would be
            # that of a later forward call, not of this one
            raise RuntimeError("DiffRun: another forward run was made on this handle after the one being "
                               "differentiated; call backward before the next forward")
        g = grad_pos.detach().to(torch.float32).contiguous()
        gt = torch.empty(max(1, b.n_roads), device=g.device, dtype=torch.float64)
        ga = torch.empty(1, device=g.device, dtype=torch.float64)
        torch.cuda.synchronize(g.device)
        b.diff_backward(g.data_ptr(), gt.data_ptr(), ga.data_ptr())
        return gt[: b.n_roads].to(torch.float32), ga.reshape(()).to(torch.float32), None


class DiffRun:
    """A batched handle set up for differentiable runs of `ticks` ticks."""

    def __init__(self, batch, ticks, checkpoint_every=50, keep_schedule=False, arrival_gradient="frozen"):
        if arrival_gradient not in ("frozen", "carry"):
            raise ValueError("arrival_gradient must be 'frozen' or 'carry'")
        batch.diff_setup(ticks, checkpoint_every, keep_schedule)
        batch.diff_set_arrival_gradient(1 if arrival_gradient == "carry" else 0)
        self.batch = batch
        self.ticks = ticks

    def __call__(self, theta, alpha):
        return _DiffFn.apply(theta, alpha, self.batch)

    def replay(self, theta, alpha):
        """Final positions under the decisions of the last forward call
        (needs keep_schedule); no gradient."""
        th = theta.detach().to(torch.float32).contiguous()
        out = torch.empty(self.batch.n, 2, device=th.device, dtype=torch.float32)
        torch.cuda.synchronize(th.device)
        self.batch.diff_forward(th.data_ptr(), float(alpha), out.data_ptr(), replay=True)
        return out
