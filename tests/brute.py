"""Brute-force references for the traversal, independent of both the oracle's
and the GPU's traversal code: enumerate every occupied leaf of a node array
and intersect rays with the leaf boxes in float64."""
import numpy as np


def leaves(nodes, basecube, levels):
    """(node index, lo xyz, hi xyz) of every leaf reachable from the root."""
    bx, by, bz, w = (float(v) for v in basecube)
    u = w / 2 ** levels
    out_n, out_c = [], []
    stack = [(0, 0, 0, 0, 0)]
    while stack:
        n, l, ix, iy, iz = stack.pop()
        if l == levels:
            out_n.append(n)
            out_c.append((ix, iy, iz))
            continue
        for o in range(8):
            c = int(nodes[n, o])
            if c > 0:
                stack.append((c, l + 1, 2 * ix + (o & 1), 2 * iy + ((o >> 1) & 1), 2 * iz + (o >> 2)))
    c = np.array(out_c, np.float64).reshape(-1, 3)
    lo = np.stack([bx + c[:, 0] * u, by - (c[:, 1] + 1) * u, bz - (c[:, 2] + 1) * u], 1)
    hi = np.stack([bx + (c[:, 0] + 1) * u, by - c[:, 1] * u, bz - c[:, 2] * u], 1)
    return np.array(out_n, np.int64), lo, hi


def first_hits(org, dirs, lo, hi, eps=1e-6):
    """For each ray: index of the leaf with the smallest entry parameter
    (clamped at 0), or -1; and a flag marking degenerate rays (the best two
    entries within eps of each other, a grazing best hit, or an entry within
    eps of 0) that the comparison skips."""
    org = np.asarray(org, np.float64)
    dirs = np.asarray(dirs, np.float64)
    best = np.full(org.shape[0], -1, np.int64)
    degen = np.zeros(org.shape[0], bool)
    tent = np.full(org.shape[0], np.inf)
    for r in range(org.shape[0]):
        p, d = org[r], dirs[r]
        with np.errstate(divide="ignore", invalid="ignore"):
            t1 = (lo - p) / d
            t2 = (hi - p) / d
        tmin = np.minimum(t1, t2)
        tmax = np.maximum(t1, t2)
        z = d == 0
        if z.any():
            inside = (p[z] > lo[:, z]) & (p[z] < hi[:, z])
            tmin[:, z] = np.where(inside, -np.inf, np.inf)
            tmax[:, z] = np.where(inside, np.inf, -np.inf)
        tn = tmin.max(1)
        tf = tmax.min(1)
        ok = (tn < tf) & (tf > 0)
        if not ok.any():
            continue
        te = np.where(ok, np.maximum(tn, 0.0), np.inf)
        order = np.argsort(te)
        b = order[0]
        best[r] = b
        tent[r] = te[b]
        scale = max(1.0, abs(te[b]))
        if len(order) > 1 and te[order[1]] - te[b] < eps * scale:
            degen[r] = True
        if tf[b] - max(tn[b], 0.0) < eps * scale:
            degen[r] = True
        if 0.0 < tn[b] < eps * scale or abs(tn[b]) < eps * scale:
            degen[r] = True
    return best, tent, degen
