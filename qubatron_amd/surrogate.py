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
