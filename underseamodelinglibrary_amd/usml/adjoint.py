"""Eigenray travel times and transmission losses as differentiable functions
of a sound speed profile's node speeds.

The reference has no derivatives; this is the differentiation form built
beside its wave_queue path.  `profile_runs` runs one scenario family over a
batch of 1-D sound speed profiles (one fan per profile, stepped together as
a fan_batch) with checkpoints of every fan's state every K steps, and turns
gradients with respect to the eigenrays back into gradients with respect to
the node speeds through the library's adjoint (wq3d_adjoint_gradient:
checkpoint replay, tape-recorded eigenray construction and reflections,
hand-written step adjoints; include/wq3d.h).  `eigenray_function` puts that
behind torch.autograd: forward(speeds [P, n]) -> (travel_time [E],
intensity [E, F]) over the E eigenrays of all P runs, in drain order per run
and run after run; backward returns d loss / d speeds.

The eigenray set itself (which (step, target, DE, AZ) pairs are eigenrays,
which branches compute_offsets and the walks take) is fixed by the forward
run; the gradient is that of the forward run's arithmetic on that set.
"""
import numpy as np
import torch

from .waveq3d import fan_batch, wave_queue


class profile_runs:
    """Forward runs of `make_scenario(speeds)` for each row of a [P, n]
    array of node speeds, kept for the backward pass.

    checkpoint_every: steps between state checkpoints (the replay segment
    length of the backward pass).  batched: step the P fans as one fan_batch
    (fused launches) instead of one after another.
    """

    def __init__(self, make_scenario, checkpoint_every=100, batched=True, lib=None):
        self.make_scenario = make_scenario
        self.checkpoint_every = int(checkpoint_every)
        self.batched = batched
        self.lib = lib
        self.queues = []
        self.records = []
        self.nf = 0
        self.n_nodes = 0

    def close(self):
        for q in self.queues:
            q.close()
        self.queues = []
        self.records = []

    def run(self, speeds):
        speeds = np.ascontiguousarray(speeds, dtype=np.float64)
        if speeds.ndim != 2:
            raise ValueError("speeds must be [profiles, nodes]")
        self.close()
        scs = [self.make_scenario(row) for row in speeds]
        self.n_nodes = speeds.shape[1]
        self.nf = scs[0].freq.size()
        steps = scs[0].steps
        if any(s.steps != steps for s in scs):
            raise ValueError("every profile's run must have the same number of steps")
        self.queues = [wave_queue(*s.args(), lib=self.lib) for s in scs]
        for q in self.queues:
            q.adjoint_checkpoint()
        batch = fan_batch(self.queues) if self.batched and len(self.queues) > 1 else None
        try:
            done = 0
            while done < steps:
                n = min(self.checkpoint_every, steps - done)
                if batch is not None:
                    batch.step(n)
                else:
                    for q in self.queues:
                        q.step(n)
                done += n
                if done < steps:
                    for q in self.queues:
                        q.adjoint_checkpoint()
        finally:
            if batch is not None:
                batch.close()
        self.records = []
        for q in self.queues:
            recs, inten, _ = q.drain_eigenrays()
            self.records.append((recs, inten.reshape(len(recs), self.nf)))
        return self.records

    def counts(self):
        return [len(r) for r, _ in self.records]

    def gradient(self, grad_time, grad_intensity, workers=4, release=False):
        """d loss / d node speeds [P, n] from the gradients with respect to
        every run's eigenrays (concatenated in run order).  `workers` fans
        run their backward passes at once (each on its own stream; about 12
        GB of device workspace each at config 5's size, pooled between
        passes).  release: free every fan's checkpoints once all passes are
        done (no second backward); freeing device memory synchronises the
        device, so it waits until no pass is running."""
        from concurrent.futures import ThreadPoolExecutor
        grad_time = np.asarray(grad_time, dtype=np.float64).reshape(-1)
        grad_intensity = np.asarray(grad_intensity, dtype=np.float64).reshape(-1, self.nf)
        out = np.zeros((len(self.queues), self.n_nodes))
        starts = np.concatenate([[0], np.cumsum(self.counts())]).astype(int)

        def one(p):
            q, (recs, inten) = self.queues[p], self.records[p]
            a, b = starts[p], starts[p + 1]
            out[p] = q.adjoint_gradient(recs, inten, grad_time[a:b], grad_intensity[a:b], self.n_nodes)

        with ThreadPoolExecutor(max_workers=max(1, int(workers))) as pool:
            list(pool.map(one, range(len(self.queues))))
        if release:
            for q in self.queues:
                q.adjoint_release()
        return out


class eigenray_function(torch.autograd.Function):
    """(travel_time [E], intensity [E, F]) = f(speeds [P, n]) over the
    eigenrays of `runs` (a profile_runs); backward through the adjoint."""

    @staticmethod
    def forward(ctx, speeds, runs):
        recs = runs.run(speeds.detach().cpu().numpy())
        ctx.runs = runs
        ctx.device, ctx.dtype = speeds.device, speeds.dtype
        tt = np.concatenate([r["travel_time"] for r, _ in recs]) if recs else np.zeros(0)
        ti = np.concatenate([i for _, i in recs]) if recs else np.zeros((0, runs.nf))
        return (torch.from_numpy(tt).to(speeds.device, speeds.dtype),
                torch.from_numpy(ti).to(speeds.device, speeds.dtype))

    @staticmethod
    def backward(ctx, grad_time, grad_intensity):
        gt = grad_time.detach().cpu().double().numpy() if grad_time is not None else 0.0
        gi = grad_intensity.detach().cpu().double().numpy() if grad_intensity is not None else 0.0
        n = sum(ctx.runs.counts())
        if np.isscalar(gt):
            gt = np.zeros(n)
        if np.isscalar(gi):
            gi = np.zeros((n, ctx.runs.nf))
        g = ctx.runs.gradient(gt, gi, release=True)
        return torch.from_numpy(g).to(ctx.device, ctx.dtype), None
