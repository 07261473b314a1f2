/*
 * layout.hip - compact, breadth-first device layout of an uploaded octree.
 *
 * The reference numbers nodes in creation order (octree.c:121-124), which
 * for scanned or generated scenes scatters the children of a node over the
 * whole array: every 32-byte child block the traversal reads sits in its own
 * cache line. The relayout renumbers nodes level by level, children of a
 * node contiguous and in octet order, and stores per node only
 *
 *     rec  = { index of the first child, live | alloc << 8 }  (8 bytes)
 *     model = slot 8 of the reference node                     (4 bytes)
 *     ref  = the reference node index                          (4 bytes)
 *
 * where alloc marks the slots of the node's child block and live the
 * children that exist (equal after a relayout; an in-place range update can
 * clear a live bit), so child o of node j is
 * rec.x + popcount(alloc & ((1 << o) - 1)). Every level
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

#include <algorithm>
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
                       int* __restrict__ r2c, uint32_t* __restrict__ seen, unsigned int* __restrict__ bad)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    uint32_t j = s + i, o = old_of_new[j], m = maskv[i];
    uint32_t first = m ? e + off[i] : 0u;
    rec[j]         = make_uint2(first, m | m << 8); /* live = alloc */
    model_c[j]     = model[o];
    ref[j]         = (int) o;
    r2c[o]         = (int) j;
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

/* ---------------- in-place patch of range uploads ------------------------ */

/* r2c entries of nodes linked by this patch carry kNew until their own
 * children are laid out */
constexpr int kNew = 1 << 30;

/* patch state words: failure, records in use, grow-list and new-list lengths */
enum { PS_FAIL = 0, PS_LEN = 1, PS_GROW = 2, PS_NEW = 3, PS_WORDS = 64 };

/* link reference node v at record p: claim r2c[v] (v must not be in the
 * tree), fill the record's ref and slot 8, and queue v for k_patch_new */
__device__ __forceinline__ void link_new(int v, uint32_t p, size_t n, const int* __restrict__ model,
                                         int* __restrict__ model_c, int* __restrict__ ref, uint2* __restrict__ rec,
                                         int* __restrict__ r2c, uint2* __restrict__ newl, uint32_t newcap,
                                         unsigned* __restrict__ ps)
{
    if (v <= 0 || (size_t) v >= n || atomicCAS(&r2c[v], -1, (int) p | kNew) != -1) { /* shared or cyclic */
        atomicOr(&ps[PS_FAIL], 1u);
        return;
    }
    ref[p]     = v;
    model_c[p] = model[v];
    rec[p]     = make_uint2(0u, 0u);
    const unsigned k = atomicAdd(&ps[PS_NEW], 1u);
    if (k < newcap) newl[k] = make_uint2((uint32_t) v, p);
    else atomicOr(&ps[PS_FAIL], 1u);
}

/* Phase 1, one node of the uploaded ranges that the layout holds (below
 * ref_n, with a record): its slot 8, and its slots against its record. A
 * zeroed slot clears the live bit (the block keeps the slot, so the ranks of
 * the others stay); a slot holding its old child again sets it. A slot that
 * now holds another node, or a slot the block lacks, needs a new block: the
 * node goes to the grow list with a snapshot of its record, and its record
 * is left to phase 2. */
__global__ void k_patch_live(const uint2* __restrict__ ranges, const uint32_t* __restrict__ roff, int nr,
                             uint32_t total, size_t ref_n, const int4* __restrict__ child,
                             const int* __restrict__ model, uint2* __restrict__ rec, int* __restrict__ model_c,
                             const int* __restrict__ ref, const int* __restrict__ r2c, uint4* __restrict__ grow,
                             unsigned* __restrict__ ps)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const size_t i = range_item(ranges, roff, nr, g);
    if (i >= ref_n) return; /* appended nodes are laid out below their parents */
    const int ci = r2c[i];
    if (ci < 0) return; /* not in the tree */
    model_c[ci]          = model[i];
    const uint2    r     = rec[ci];
    const uint32_t alloc = (r.y >> 8) & 0xFFu;
    const int*     c     = reinterpret_cast<const int*>(child) + 8 * i;
    uint32_t       live  = 0;
    bool           relink = false;
    for (int o = 0; o < 8; o++) {
        const uint32_t bit = 1u << o;
        const int      v   = c[o];
        if (v == 0) continue;
        if (!(alloc & bit) || ref[r.x + (uint32_t) __popc(alloc & (bit - 1u))] != v) relink = true;
        else live |= bit;
    }
    if (!relink) {
        rec[ci].y = live | alloc << 8;
        return;
    }
    const unsigned k = atomicAdd(&ps[PS_GROW], 1u);
    grow[k]          = make_uint4((uint32_t) i, (uint32_t) ci, r.x, r.y); /* at most one per item: fits */
}

/* Phase 2a, one grow-list node: a new child block at the end of the records
 * with a slot for every nonzero slot and every slot the old block had (a
 * zeroed one stays, not live); the old children's records are copied there
 * (the old block stays unused), the others linked. The node's new record is
 * left in grow[k].zw for phase 2b, since its own record may be moving in
 * this phase (its parent may be on the list). */
__global__ void k_patch_grow(uint4* __restrict__ grow, size_t n, const int4* __restrict__ child,
                             const int* __restrict__ model, uint2* __restrict__ rec, int* __restrict__ model_c,
                             int* __restrict__ ref, int* __restrict__ r2c, uint32_t cap, uint2* __restrict__ newl,
                             uint32_t newcap, unsigned* __restrict__ ps)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ps[PS_GROW]) return;
    const uint4    e     = grow[k];
    const uint32_t alloc = (e.w >> 8) & 0xFFu;
    const int*     c     = reinterpret_cast<const int*>(child) + 8 * (size_t) e.x;
    uint32_t       na    = alloc;
    for (int o = 0; o < 8; o++)
        if (c[o] > 0) na |= 1u << o;
    const uint32_t base = atomicAdd(&ps[PS_LEN], (unsigned) __popc(na));
    if ((size_t) base + (size_t) __popc(na) > cap) {
        atomicOr(&ps[PS_FAIL], 1u);
        return;
    }
    uint32_t live = 0;
    for (int o = 0; o < 8; o++) {
        const uint32_t bit = 1u << o;
        if (!(na & bit)) continue;
        const uint32_t np = base + (uint32_t) __popc(na & (bit - 1u));
        const int      v  = c[o];
        int            cur = -1;
        if (alloc & bit) { /* carry the old child's record over */
            const uint32_t op = e.z + (uint32_t) __popc(alloc & (bit - 1u));
            cur               = ref[op];
            rec[np]           = rec[op];
            model_c[np]       = model_c[op];
            ref[np]           = cur;
            if (cur > 0 && r2c[cur] == (int) op) r2c[cur] = (int) np;
        }
        if (v == 0) continue; /* zeroed: not live */
        live |= bit;
        if (v == cur) continue;
        if (cur > 0 && r2c[cur] == (int) np) r2c[cur] = -1; /* replaced: the old child leaves the tree */
        link_new(v, np, n, model, model_c, ref, rec, r2c, newl, newcap, ps);
    }
    grow[k].z = base;
    grow[k].w = live | na << 8;
}

/* Phase 2b: the grow-list nodes' records, where they are now (a node its
 * parent replaced in phase 2a has left the tree: nothing to write) */
__global__ void k_patch_commit(const uint4* __restrict__ grow, uint2* __restrict__ rec, const int* __restrict__ r2c,
                               const unsigned* __restrict__ ps)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ps[PS_GROW]) return;
    const uint4 e = grow[k];
    const int   p = r2c[e.x];
    if (p >= 0) rec[p & ~kNew] = make_uint2(e.z, e.w);
}

/* Phase 3: the subtrees below the nodes linked in phase 2, depth first, one
 * list entry per thread: each node's children get a block of records at the
 * end (live = alloc = its nonzero slots) and are claimed in r2c. */
__global__ void k_patch_new(const uint2* __restrict__ newl, uint32_t newcap, size_t n, const int4* __restrict__ child,
                            const int* __restrict__ model, uint2* __restrict__ rec, int* __restrict__ model_c,
                            int* __restrict__ ref, int* __restrict__ r2c, uint32_t cap, unsigned* __restrict__ ps)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ps[PS_NEW] || e >= newcap) return;
    constexpr int kStack = 192; /* pending nodes; deeper or wider subtrees relayout instead */
    uint2         stk[kStack];
    int           sp = 0;
    stk[sp++]        = newl[e];
    while (sp > 0) {
        const uint2 u = stk[--sp];
        const int*  c = reinterpret_cast<const int*>(child) + 8 * (size_t) u.x;
        uint32_t    m = 0;
        for (int o = 0; o < 8; o++)
            if (c[o] > 0) m |= 1u << o;
        uint32_t base = 0;
        if (m) {
            base = atomicAdd(&ps[PS_LEN], (unsigned) __popc(m));
            if ((size_t) base + (size_t) __popc(m) > cap) {
                atomicOr(&ps[PS_FAIL], 1u);
                return;
            }
        }
        rec[u.y]   = make_uint2(base, m | m << 8);
        r2c[u.x]   = (int) u.y; /* laid out */
        uint32_t k = 0;
        for (int o = 0; o < 8; o++) {
            if (!((m >> o) & 1u)) continue;
            const int      v = c[o];
            const uint32_t p = base + k++;
            if ((size_t) v >= n || atomicCAS(&r2c[v], -1, (int) p | kNew) != -1 || sp >= kStack) {
                atomicOr(&ps[PS_FAIL], 1u);
                return;
            }
            ref[p]     = v;
            model_c[p] = model[v];
            stk[sp++]  = make_uint2((uint32_t) v, p);
        }
    }
}

inline unsigned nblk(size_t n) { return (unsigned) ((n + 255) / 256); }

} // namespace

int qbt_ctx::compact_reserve(int t, size_t nrec, size_t nref)
{
    CompactTree& cm = tree[t].cmp;
    int          rc;
    if (nrec > cm.cap) { /* room for the records in-place patches append */
        const size_t cap = nrec + nrec / 8 + 65536;
        if ((rc = dev_alloc(cm.rec, cap * 8)) || (rc = dev_alloc(cm.model, cap * 4)) || (rc = dev_alloc(cm.ref, cap * 4)))
            return rc;
        cm.cap = cap;
    }
    if (nref > cm.r2c_cap) {
        const size_t cap = nref > tree[t].cap_nodes ? nref : tree[t].cap_nodes;
        if ((rc = dev_alloc(cm.r2c, cap * 4))) return rc;
        cm.r2c_cap = cap;
    }
    return QBT_OK;
}

/* (Re)build the compact layout of tree t. Returns QBT_OK with cmp.ok set to
 * whether the array is a proper tree; errors only on allocation/HIP
 * failures. Level by level: the node count of the next level is read back
 * before its launch (one small synchronisation per level); the work buffers
 * are kept for the next relayout. */
int qbt_ctx::relayout(int t)
{
    TreeBuf&     tb = tree[t];
    CompactTree& cm = tb.cmp;
    cm.ok           = false;
    cm.n            = 0;
    cm.ref_n        = 0;
    tb.patch.clear();
    tb.relayouts++;
    const size_t n  = tb.n;
    if (n == 0 || n > (size_t) INT32_MAX) return QBT_OK;
    hipStream_t s = stream;
    int         rc;
    if ((rc = compact_reserve(t, n, n))) return rc;
    /* work buffers, carved from one kept allocation */
    const size_t words = (n + 31) / 32;
    size_t       tb_bytes = 0;
    if (rocprim::exclusive_scan(nullptr, tb_bytes, (uint32_t*) nullptr, (uint32_t*) nullptr, 0u, n,
                                rocprim::plus<uint32_t>(), s) != hipSuccess)
        return QBT_EDEVICE;
    auto   al      = [](size_t b) { return (b + 255) & ~(size_t) 255; };
    size_t o_oon = 0, o_cnt = o_oon + al(n * 4), o_off = o_cnt + al(n * 4), o_msk = o_off + al(n * 4);
    size_t o_seen = o_msk + al(n * 4), o_bad = o_seen + al(words * 4), o_tmp = o_bad + 256;
    if ((rc = dev_alloc(cm.scratch, o_tmp + al(tb_bytes ? tb_bytes : 16)))) return rc;
    char*     sb   = (char*) cm.scratch.p;
    uint32_t* oon  = (uint32_t*) (sb + o_oon);
    uint32_t* cnt  = (uint32_t*) (sb + o_cnt);
    uint32_t* off  = (uint32_t*) (sb + o_off);
    uint32_t* msk  = (uint32_t*) (sb + o_msk);
    uint32_t* seen = (uint32_t*) (sb + o_seen);
    uint32_t* bad  = (uint32_t*) (sb + o_bad);
    void*     tmp  = sb + o_tmp;
    k_fill32<<<nblk(words), 256, 0, s>>>(seen, words, 0u);
    k_fill32<<<nblk(n), 256, 0, s>>>((uint32_t*) cm.r2c.p, n, ~0u); /* -1: not in the tree */
    k_fill32<<<1, 64, 0, s>>>(bad, 2, 0u);
    k_fill32<<<1, 64, 0, s>>>(oon, 1, 0u); /* new 0 = reference root 0 */
    uint32_t lo = 0, hi = 1;
    bool     ok = true;
    for (int depth = 0; hi > lo; depth++) {
        if (depth > 64) { /* deeper than any octree: a cycle */
            ok = false;
            break;
        }
        const uint32_t k = hi - lo;
        k_count<<<nblk(k), 256, 0, s>>>((const int4*) tb.child.p, oon, lo, k, cnt, msk);
        size_t b2 = tb_bytes;
        if (rocprim::exclusive_scan(tmp, b2, cnt, off, 0u, k, rocprim::plus<uint32_t>(), s) != hipSuccess)
            return QBT_EDEVICE;
        uint32_t last[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(&last[0], off + (k - 1), 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(&last[1], cnt + (k - 1), 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const uint64_t total = (uint64_t) last[0] + last[1];
        if ((uint64_t) hi + total > n) { /* more nodes than the array holds: shared children */
            ok = false;
            break;
        }
        k_emit<<<nblk(k), 256, 0, s>>>((const int4*) tb.child.p, (const int*) tb.model.p, oon, lo, hi, k, off, msk,
                                        (uint32_t) n, (uint2*) cm.rec.p, (int*) cm.model.p, (int*) cm.ref.p,
                                        (int*) cm.r2c.p, seen, (unsigned int*) bad);
        HIPCHK(hipGetLastError());
        lo = hi;
        hi = (uint32_t) (hi + total);
    }
    unsigned int nbad = 0;
    HIPCHK(hipMemcpyAsync(&nbad, bad, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    /* the single-tree traversal keeps first-child indices in 30 bits */
    cm.ok    = ok && nbad == 0 && hi < (1u << 30);
    cm.n     = cm.ok ? hi : 0;
    cm.ref_n = cm.ok ? n : 0;
    return QBT_OK;
}

/* The range uploads since the last layout, in place: k_patch_live,
 * k_patch_grow, k_patch_commit, k_patch_new. Records are appended up to the
 * capacity; *done = false when the uploads link a node that is already in
 * the tree (shared child, cycle), need more records than the capacity holds,
 * or the layout is not current. One synchronisation reads the outcome. */
int qbt_ctx::patch_layout(int t, bool* done)
{
    TreeBuf&     tb = tree[t];
    CompactTree& cm = tb.cmp;
    *done           = false;
    if (!cm.ok || cm.ref_n == 0 || tb.n < cm.ref_n || tb.n > (size_t) INT32_MAX) return QBT_OK;
    hipStream_t  s = stream;
    int          rc;
    const size_t n = tb.n;
    /* the layout's arrays must already hold every record and reference node
     * (growing them would drop their contents; the array only outgrows them
     * by a regrowth, which relays out) */
    if (cm.n > cm.cap || n > cm.r2c_cap) return QBT_OK;
    if (cm.cap >= (1u << 30)) return QBT_OK; /* first-child indices are 30-bit */
    /* the ranges (merged on upload) below ref_n and their prefix offsets */
    std::vector<uint2>    rg;
    std::vector<uint32_t> roff;
    uint64_t              total = 0;
    std::sort(tb.patch.begin(), tb.patch.end()); /* disjoint ranges: every node once */
    std::vector<std::pair<size_t, size_t>> merged;
    for (auto& r : tb.patch) {
        if (!merged.empty() && r.first <= merged.back().second)
            merged.back().second = r.second > merged.back().second ? r.second : merged.back().second;
        else merged.push_back(r);
    }
    for (auto& r : merged) {
        const size_t lo = r.first, hi = r.second < cm.ref_n ? r.second : cm.ref_n;
        if (hi <= lo) continue;
        rg.push_back(make_uint2((uint32_t) lo, (uint32_t) hi));
        roff.push_back((uint32_t) total);
        total += hi - lo;
    }
    if (total > (1u << 28)) return QBT_OK; /* as much work as a relayout */
    /* a grow entry per dirty node, up to 8 new children each */
    const uint64_t newcap64 = total * 8 < (1u << 26) ? total * 8 + 8 : (1u << 26);
    const uint32_t newcap   = (uint32_t) newcap64;
    auto           al       = [](size_t b) { return (b + 255) & ~(size_t) 255; };
    const size_t   o_ps = 0, o_rg = o_ps + al(PS_WORDS * 4), o_ro = o_rg + al(rg.size() * 8 + 8);
    const size_t   o_gr = o_ro + al(roff.size() * 4 + 4), o_nl = o_gr + al(total * 16 + 16);
    if ((rc = dev_alloc(cm.scratch, o_nl + (size_t) newcap * 8))) return rc;
    char*     sb   = (char*) cm.scratch.p;
    unsigned* ps   = (unsigned*) (sb + o_ps);
    uint2*    drg  = (uint2*) (sb + o_rg);
    uint32_t* dro  = (uint32_t*) (sb + o_ro);
    uint4*    grow = (uint4*) (sb + o_gr);
    uint2*    newl = (uint2*) (sb + o_nl);
    unsigned  hps[PS_WORDS] = {0};
    hps[PS_LEN]             = (unsigned) cm.n;
    HIPCHK(hipMemcpyAsync(ps, hps, sizeof(hps), hipMemcpyHostToDevice, s));
    if (!rg.empty()) {
        HIPCHK(hipMemcpyAsync(drg, rg.data(), rg.size() * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dro, roff.data(), roff.size() * 4, hipMemcpyHostToDevice, s));
    }
    if (n > cm.ref_n) /* nodes appended since the layout: not in the tree yet */
        k_fill32<<<nblk(n - cm.ref_n), 256, 0, s>>>((uint32_t*) cm.r2c.p + cm.ref_n, n - cm.ref_n, ~0u);
    const int4* ch = (const int4*) tb.child.p;
    const int*  md = (const int*) tb.model.p;
    uint2*      rc_ = (uint2*) cm.rec.p;
    int*        mc = (int*) cm.model.p;
    int*        rf = (int*) cm.ref.p;
    int*        r2 = (int*) cm.r2c.p;
    if (total) {
        const unsigned nb = nblk(total);
        k_patch_live<<<nb, 256, 0, s>>>(drg, dro, (int) rg.size(), (uint32_t) total, cm.ref_n, ch, md, rc_, mc, rf, r2,
                                        grow, ps);
        k_patch_grow<<<nb, 256, 0, s>>>(grow, n, ch, md, rc_, mc, rf, r2, (uint32_t) cm.cap, newl, newcap, ps);
        k_patch_commit<<<nb, 256, 0, s>>>(grow, rc_, r2, ps);
        k_patch_new<<<nblk(newcap), 256, 0, s>>>(newl, newcap, n, ch, md, rc_, mc, rf, r2, (uint32_t) cm.cap, ps);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(hps, ps, 4 * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    tb.patch.clear();
    if (hps[PS_FAIL]) return QBT_OK;
    cm.n     = hps[PS_LEN];
    cm.ref_n = n;
    tb.patches++;
    *done = true;
    return QBT_OK;
}
