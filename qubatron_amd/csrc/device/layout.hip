/*
 * layout.hip - compact, breadth-first device layout of an uploaded octree.
 *
 * The reference numbers nodes in creation order (octree.c:121-124), which
 * for scanned or generated scenes scatters the children of a node over the
 * whole array: every 32-byte child block the traversal reads sits in its own
 * cache line. The relayout renumbers nodes level by level, children of a
 * node contiguous and in octet order, and stores per node only
 *
 *     rec  = { index of the first child, 8-bit child mask }   (8 bytes)
 *     model = slot 8 of the reference node                     (4 bytes)
 *     ref  = the reference node index                          (4 bytes)
 *
 * so child o of node j is rec.x + popcount(mask & ((1 << o) - 1)). Every level
 * comes out in hierarchical octet order, so nodes that are close in space are
 * close in memory and rays of neighbouring pixels share cache lines.
 * Traversal results are unchanged (the same children exist, the same leaves
 * carry the same slot 8); reported leaf indices are mapped back through ref.
 *
 * Only a proper tree can be renumbered this way: if a node is reachable
 * twice (shared child, cycle) the relayout reports failure and the direct
 * layout is used.
 */
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/rocprim.hpp>

#include "qbt.h"
#include "qbt_internal.hpp"

#define HIPCHK(x) QBT_HIPCHK(x)

namespace {

__global__ void k_count(const int4* __restrict__ child, const uint32_t* __restrict__ old_of_new, uint32_t s,
                        uint32_t k, uint32_t* __restrict__ cnt, uint32_t* __restrict__ maskv)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    uint32_t o = old_of_new[s + i];
    int4     a = child[2 * (size_t) o], b = child[2 * (size_t) o + 1];
    uint32_t m = (a.x > 0 ? 1u : 0u) | (a.y > 0 ? 2u : 0u) | (a.z > 0 ? 4u : 0u) | (a.w > 0 ? 8u : 0u) |
                 (b.x > 0 ? 16u : 0u) | (b.y > 0 ? 32u : 0u) | (b.z > 0 ? 64u : 0u) | (b.w > 0 ? 128u : 0u);
    maskv[i] = m;
    cnt[i]   = (uint32_t) __popc(m);
}

__global__ void k_emit(const int4* __restrict__ child, const int* __restrict__ model,
                       uint32_t* __restrict__ old_of_new, uint32_t s, uint32_t e, uint32_t k,
                       const uint32_t* __restrict__ off, const uint32_t* __restrict__ maskv, uint32_t cap,
                       uint2* __restrict__ rec, int* __restrict__ model_c, int* __restrict__ ref,
                       uint32_t* __restrict__ seen, unsigned int* __restrict__ bad)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    uint32_t j = s + i, o = old_of_new[j], m = maskv[i];
    uint32_t first = m ? e + off[i] : 0u;
    rec[j]         = make_uint2(first, m);
    model_c[j]     = model[o];
    ref[j]         = (int) o;
    if (!m) return;
    const int* c = reinterpret_cast<const int*>(child) + 8 * (size_t) o;
    uint32_t   r = 0;
    for (int q = 0; q < 8; q++) {
        if (!((m >> q) & 1u)) continue;
        uint32_t co = (uint32_t) c[q];
        uint32_t bit = 1u << (co & 31u);
        if (atomicOr(&seen[co >> 5], bit) & bit) atomicAdd(bad, 1u); /* reached twice: not a tree */
        uint32_t nj = first + r++;
        if (nj < cap) old_of_new[nj] = co;
        else atomicAdd(bad, 1u);
    }
}

__global__ void k_fill32(uint32_t* p, size_t n, uint32_t v)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned nblk(size_t n) { return (unsigned) ((n + 255) / 256); }

} // namespace

/* (Re)build the compact layout of tree t. Returns QBT_OK with cmp.ok set to
 * whether the array is a proper tree; errors only on allocation/HIP
 * failures. */
int qbt_ctx::relayout(int t)
{
    TreeBuf&     tb = tree[t];
    CompactTree& cm = tb.cmp;
    cm.ok           = false;
    cm.n            = 0;
    const size_t n  = tb.n;
    if (n == 0 || n > (size_t) INT32_MAX) return QBT_OK;
    hipStream_t s = stream;
    int         rc;
    /* persistent outputs */
    if ((rc = dev_alloc(cm.rec, n * 8)) || (rc = dev_alloc(cm.model, n * 4)) || (rc = dev_alloc(cm.ref, n * 4)))
        return rc;
    /* scratch */
    DevBuf oon, cnt, off, msk, seen, bad, tmp;
    auto   cleanup = [&]() {
        for (DevBuf* b : {&oon, &cnt, &off, &msk, &seen, &bad, &tmp}) dev_free(*b);
    };
    const size_t words = (n + 31) / 32;
    if ((rc = dev_alloc(oon, n * 4)) || (rc = dev_alloc(cnt, n * 4)) || (rc = dev_alloc(off, n * 4)) ||
        (rc = dev_alloc(msk, n * 4)) || (rc = dev_alloc(seen, words * 4)) || (rc = dev_alloc(bad, 256))) {
        cleanup();
        return rc;
    }
    size_t tb_bytes = 0;
    if (rocprim::exclusive_scan(nullptr, tb_bytes, (uint32_t*) cnt.p, (uint32_t*) off.p, 0u, n,
                                rocprim::plus<uint32_t>(), s) != hipSuccess ||
        (rc = dev_alloc(tmp, tb_bytes ? tb_bytes : 16))) {
        cleanup();
        return rc ? rc : QBT_EDEVICE;
    }
    k_fill32<<<nblk(words), 256, 0, s>>>((uint32_t*) seen.p, words, 0u);
    k_fill32<<<1, 64, 0, s>>>((uint32_t*) bad.p, 2, 0u);
    k_fill32<<<1, 64, 0, s>>>((uint32_t*) oon.p, 1, 0u); /* new 0 = reference root 0 */
    uint32_t lo = 0, hi = 1;
    bool     ok = true;
    for (int depth = 0; hi > lo; depth++) {
        if (depth > 64) { /* deeper than any octree: a cycle */
            ok = false;
            break;
        }
        const uint32_t k = hi - lo;
        k_count<<<nblk(k), 256, 0, s>>>((const int4*) tb.child.p, (const uint32_t*) oon.p, lo, k,
                                         (uint32_t*) cnt.p, (uint32_t*) msk.p);
        size_t b2 = tb_bytes;
        if (rocprim::exclusive_scan(tmp.p, b2, (uint32_t*) cnt.p, (uint32_t*) off.p, 0u, k, rocprim::plus<uint32_t>(),
                                    s) != hipSuccess) {
            cleanup();
            return QBT_EDEVICE;
        }
        uint32_t last[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(&last[0], (uint32_t*) off.p + (k - 1), 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(&last[1], (uint32_t*) cnt.p + (k - 1), 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const uint64_t total = (uint64_t) last[0] + last[1];
        if ((uint64_t) hi + total > n) { /* more nodes than the array holds: shared children */
            ok = false;
            break;
        }
        k_emit<<<nblk(k), 256, 0, s>>>((const int4*) tb.child.p, (const int*) tb.model.p, (uint32_t*) oon.p, lo, hi, k,
                                        (const uint32_t*) off.p, (const uint32_t*) msk.p, (uint32_t) n,
                                        (uint2*) cm.rec.p, (int*) cm.model.p, (int*) cm.ref.p, (uint32_t*) seen.p,
                                        (unsigned int*) bad.p);
        HIPCHK(hipGetLastError());
        lo = hi;
        hi = (uint32_t) (hi + total);
    }
    unsigned int nbad = 0;
    HIPCHK(hipMemcpyAsync(&nbad, bad.p, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    cleanup();
    /* the single-tree traversal keeps first-child indices in 30 bits */
    cm.ok = ok && nbad == 0 && hi < (1u << 30);
    cm.n  = cm.ok ? hi : 0;
    return QBT_OK;
}
