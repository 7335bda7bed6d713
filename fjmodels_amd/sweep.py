"""Model sweeps: many single-component models with seeded random DF
parameters and uneven grids, advanced by the batched path (fj_batch_*).

The reference runs one model per invocation of fJmodels.cpp (oneComp,
fJmodels.cpp:86-144); a sweep is that run repeated over a parameter draw.
Here the models are packed, in draw order, into as few batches as the
per-batch node limit of fj_batch_create allows (64 M velocity nodes, i.e.
about 6,300 grid points of 22^3 nodes), and each batch advances all its
models with one launch per iteration phase, replayed from a HIP graph.
Under FJ_PREC_FAST all models go into one batch, which runs them one after
another through the fast node phase's kernels.
"""
import time

import numpy as np

from .capi import Batch, df_params, FJ_PREC_FAST
from .scf import log_grid

NODES_PER_POINT = 22 ** 3
BATCH_NODE_LIMIT = 64 << 20
# fast-phase batches keep one model's node records at a time and a 4-byte
# sort key per node: no node limit of their own
FAST_BATCH_NODE_LIMIT = 1 << 40
KINDS = ("Hernquist", "Isochrone", "NFW")


def draw_models(n, seed=0, nr=(24, 96), ntheta=(12, 32), lmax=(8, 10, 12), delta=(0.3, 1.2), chi=(0.0, 2.0),
                rmin=0.01, rmax=100.0, rot_form=1):
    """n model specifications for Batch / Sweep, drawn in order from
    numpy.random.default_rng(seed): model type uniform over Hernquist,
    Isochrone, NFW (A, B of df/DF.cpp:40-57, guess potential of the same
    type, M = 1, r0 = 1); h(J) deltas dphi, dz ~ U(delta); chi ~ U(chi) with
    the README rotation form; nr ~ U{nr}, ntheta ~ U{ntheta} (both ends
    included), l_max uniform over `lmax` (NPOLY = l_max / 2 + 1); nr log
    radii in [rmin, rmax]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        kind = KINDS[int(rng.integers(3))]
        dphi, dz = (float(x) for x in rng.uniform(delta[0], delta[1], 2))
        c = float(rng.uniform(chi[0], chi[1]))
        r = int(rng.integers(nr[0], nr[1] + 1))
        t = int(rng.integers(ntheta[0], ntheta[1] + 1))
        lm = int(lmax[int(rng.integers(len(lmax)))])
        out.append({"ar": log_grid(r, rmin, rmax), "ntheta": t, "npoly": lm // 2 + 1, "model": kind,
                    "df": df_params(kind, dphi_h=dphi, dz_h=dz, chi=c, rot_form=rot_form)})
    return out


def model_nodes(m):
    return len(m["ar"]) * m["ntheta"] * NODES_PER_POINT


def partition(models, node_limit=BATCH_NODE_LIMIT):
    """Consecutive groups of model indices, each holding at most node_limit
    nodes; a model larger than the limit cannot be batched (ValueError)."""
    groups, cur, used = [], [], 0
    for k, m in enumerate(models):
        n = model_nodes(m)
        if n > node_limit:
            raise ValueError("model %d has %d nodes, more than the batch limit %d" % (k, n, node_limit))
        if cur and used + n > node_limit:
            groups.append(cur)
            cur, used = [], 0
        cur.append(k)
        used += n
    if cur:
        groups.append(cur)
    return groups


class Sweep:
    """Runs `models` through the batched path, batch after batch, on ctx's
    device with ctx's precision.  run(niter) returns each model's tables
    after niter SCF iterations (Batch.get) and the timings: iteration time
    summed over the batches (iterations 2..niter: the first includes the
    graph capture, or under FJ_PREC_FAST the first pass's lane order before
    any turning points are known), setup time (guess potentials, Poisson,
    Delta tables) apart."""

    def __init__(self, ctx, models, node_limit=None, autotune=False):
        """node_limit: velocity nodes per batch; by default BATCH_NODE_LIMIT,
        or FAST_BATCH_NODE_LIMIT under FJ_PREC_FAST, whose batches hold no
        node records (the whole 256-model draw is one batch)."""
        self.ctx, self.models, self.autotune = ctx, models, autotune
        if node_limit is None:
            fast = getattr(ctx, "precision", None) == FJ_PREC_FAST
            node_limit = FAST_BATCH_NODE_LIMIT if fast else BATCH_NODE_LIMIT
        self.groups = partition(models, node_limit)

    def run(self, niter=2, progress=None):
        """progress: optional callable, called after each batch with
        (batches done, batches, seconds so far)."""
        if niter < 1:
            raise ValueError("niter >= 1")
        t_start = time.perf_counter()
        results = [None] * len(self.models)
        t_setup = t_first = t_iter = 0.0
        ms_gpu = 0.0
        for gi, g in enumerate(self.groups):
            t0 = time.perf_counter()
            b = Batch(self.ctx, [self.models[k] for k in g])
            try:
                b.setup()
                if self.autotune:
                    b.autotune()
                t1 = time.perf_counter()
                b.iterate(1)
                t2 = time.perf_counter()
                if niter > 1:
                    b.iterate(niter - 1)
                t3 = time.perf_counter()
                ms_gpu += b.stats()["ms_last_iter"]
                for j, k in enumerate(g):
                    results[k] = b.get(j)
            finally:
                b.close()
            t_setup += t1 - t0
            t_first += t2 - t1
            t_iter += t3 - t2
            if progress:
                progress(gi + 1, len(self.groups), time.perf_counter() - t_start)
        pts = sum(len(m["ar"]) * m["ntheta"] for m in self.models)
        st = {"models": len(self.models), "batches": len(self.groups), "grid_points": pts,
              "nodes": pts * NODES_PER_POINT, "s_setup": t_setup, "s_first_iteration": t_first,
              "ms_last_iteration_gpu": ms_gpu}
        if niter > 1:
            st["ms_per_iteration"] = 1e3 * t_iter / (niter - 1)
            st["value"] = pts / (t_iter / (niter - 1))
        return results, st
