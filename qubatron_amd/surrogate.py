"""Host side of the learned shadow surrogate: online training from exact
renders and the mismatch measurement against the exact shadow bits.

The network, its kernels and the render-path decision live in libqbt.so
(surrogate.hip); these helpers only sequence C-ABI calls.
"""
import numpy as np

from . import _ffi


def train_online(ctx, views, lights, basecube, nsamples=1 << 24, stride=32, steps_per_frame=16, batch=4096,
                 lr=2e-3, seed=0, init=True, final_steps=0, final_batch=None, final_lr=None, decay=0.0):
    """Render each view exactly with sample collection, and after each frame
    take steps_per_frame Adam steps on minibatches of everything collected so
    far (once the store holds a batch); then final_steps more on the whole
    store (batch final_batch, learning rate final_lr; both default to the
    online ones), the rate falling along a cosine from final_lr to
    decay * final_lr over 16 equal chunks when decay > 0. Returns a dict of
    training figures."""
    if init:
        ctx.surrogate_init(seed)
    ctx.surrogate_reserve(nsamples, stride)
    losses, steps = [], 0
    for i, v in enumerate(views):
        ctx.render(v, lights, _ffi.OUT_FLAGS | _ffi.RENDER_COLLECT)
        n, _ = ctx.surrogate_samples()
        if n >= batch and steps_per_frame > 0:
            losses.append(ctx.surrogate_train(basecube, steps_per_frame, batch, lr, seed * 1000003 + i))
            steps += steps_per_frame
    if final_steps > 0:
        fb = final_batch or batch
        flr = final_lr if final_lr is not None else lr
        chunks = 16 if decay > 0 else 1
        for k in range(chunks):
            n_k = final_steps * (k + 1) // chunks - final_steps * k // chunks
            if n_k <= 0:
                continue
            f = 0.5 * (1.0 + np.cos(np.pi * (k + 0.5) / chunks)) if decay > 0 else 1.0
            rate = flr * (decay + (1.0 - decay) * f)
            losses.append(ctx.surrogate_train(basecube, n_k, fb, rate, seed * 1000003 + len(views) + 7 * k))
        steps += final_steps
    n, lit = ctx.surrogate_samples()
    return {"samples": n, "lit_fraction": lit / max(n, 1), "steps": steps, "frames": len(views),
            "loss_first": losses[0] if losses else None, "loss_last": losses[-1] if losses else None}


def shadow_mismatch(exact, surrogate, nlights):
    """Shadow-bit mismatch between an exact and a surrogate frame (dicts from
    Context.read_frame with 'flags' and 'shadow'): over the (shaded pixel,
    light) queries of the exact frame. Returns (mismatches, queries)."""
    fe = exact["flags"]
    shaded = ((fe & _ffi.PX_SHADED) != 0) & ((fe & _ffi.PX_DISCARD) == 0)
    diff = (exact["shadow"] ^ surrogate["shadow"])[shaded]
    bad = sum(int(np.count_nonzero((diff >> l) & 1)) for l in range(nlights))
    return bad, int(np.count_nonzero(shaded)) * nlights


def choose_band(p, y, target=0.01, steps=500):
    """The uncertainty band [lo, hi] that decides the most queries while the
    decided ones that disagree with their label stay within target of all
    queries: a query is decided 0 when p < lo and 1 when p > hi (the render
    path's rule). lo runs over a grid of steps values in (0, 0.5], hi over
    [0.5, 1). Returns (lo, hi, decided fraction, mismatch fraction)."""
    p = np.asarray(p, dtype=np.float64).ravel()
    y = np.asarray(y).ravel().astype(bool)
    n = p.size
    if n == 0:
        raise ValueError("choose_band: no samples")
    lo = np.linspace(0.0, 0.5, steps + 1)[1:]
    hi = np.linspace(0.5, 1.0, steps + 1)[:-1]
    ps, pp, pn = np.sort(p), np.sort(p[y]), np.sort(p[~y])
    n0 = np.searchsorted(ps, lo, side="left")                 # p < lo: decided 0
    e0 = np.searchsorted(pp, lo, side="left")                 # ... labelled 1
    n1 = n - np.searchsorted(ps, hi, side="right")            # p > hi: decided 1
    e1 = pn.size - np.searchsorted(pn, hi, side="right")      # ... labelled 0
    err = e0[:, None] + e1[None, :]
    dec = np.where(err <= target * n, n0[:, None] + n1[None, :], -1)
    i, j = np.unravel_index(int(np.argmax(dec)), dec.shape)
    if dec[i, j] < 0:  # no band of the grid meets the target: decide (almost) nothing
        a, b = 1e-6, 1.0 - 1e-6
        d0, d1 = p < a, p > b
        return a, b, float(np.count_nonzero(d0) + np.count_nonzero(d1)) / n, \
            float(np.count_nonzero(d0 & y) + np.count_nonzero(d1 & ~y)) / n
    return float(lo[i]), float(hi[j]), float(n0[i] + n1[j]) / n, float(e0[i] + e1[j]) / n


def calibrate_band(ctx, views, lights, basecube, target=0.01, nsamples=1 << 23):
    """Collect every stride-th shadow query of exact renders of views (frames
    the network was not trained on), predict them, and set the band that
    choose_band picks for them. Empties the training store. Returns a dict of
    the band and the held-out figures."""
    ctx.render(views[0], lights, _ffi.OUT_FLAGS)
    per_frame = max(1, ctx.stats()["shadow_rays"])
    stride = max(1, int(round(len(views) * per_frame / nsamples)))
    ctx.surrogate_reserve(nsamples, stride)
    for v in views:
        ctx.render(v, lights, _ffi.OUT_FLAGS | _ffi.RENDER_COLLECT)
    s = ctx.surrogate_read_samples()
    p = ctx.surrogate_predict(s[:, 0:3], s[:, 3:6], s[:, 6:9], basecube)
    lo, hi, dec, mis = choose_band(p, s[:, 9] > 0.5, target)
    ctx.surrogate_set_band(lo, hi)
    return {"band": [lo, hi], "frames": len(views), "samples": int(s.shape[0]), "stride": stride,
            "target_mismatch": target, "held_out_decided": round(dec, 4), "held_out_mismatch": round(mis, 5)}

