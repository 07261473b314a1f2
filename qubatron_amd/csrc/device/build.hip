/*
 * build.hip - octree construction on the GPU (qbt_build_octree).
 *
 * Produces, node for node, the array that inserting points 0..n-1 in index
 * order with octree_insert_point (src/qubatron/octree.c:95-147) produces:
 *
 *  - every point's octet path is computed with the reference's fp32 formula
 *    ((int)(p/size)%2 with size halved per level, octree.c:102-109) and
 *    packed into a key, 3 bits per level, first level highest;
 *  - a stable radix sort groups equal prefixes while keeping index order;
 *  - the node of a depth-d prefix is created by the smallest point index in
 *    its group (the first inserter); per depth, a reduce-by-key with min over
 *    the sorted keys gives the groups and their first inserters, walking up
 *    from the leaves (a parent group is a run of child groups);
 *  - a point that is the first inserter at depth d is so at every deeper
 *    depth, so point i creates nodes at depths dmin_i..L consecutively and the
 *    node index of (i, d) is 1 + sum_{j<i} count_j + (d - dmin_i): an
 *    exclusive scan of the per-point counts in index order;
 *  - slot 8 of a node is its first inserter (the model index), and its
 *    parent's slot for the last octet of the prefix points to it.
 */
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/rocprim.hpp>

#include "qbt.h"
#include "qbt_internal.hpp"

#pragma clang fp contract(off)

/* Scratch blocks of the builder: a build takes about 80 scratch buffers
 * (keys, values and groups per level), and allocating and freeing them each
 * time cost more than the build itself for a few million points (the
 * reference rebuilds its dynamic tree every frame). Each build carves its
 * buffers from the blocks in order, adding a block where none fits. A
 * dynamic-tree build keeps its blocks, up to kKeepBytes, for the next
 * per-frame rebuild; a static build releases them when it ends. Kept blocks
 * are idle between builds, and qbt_ctx::dev_alloc releases them (trim)
 * before it would report QBT_ENOMEM. */
struct BuildState {
    std::vector<DevBuf> blocks;
    bool                active = false; /* a build is carving from the blocks */
};

void qbt_ctx::builder_free()
{
    if (build)
        for (auto& b : build->blocks) dev_free(b);
    delete build;
    build = nullptr;
}

bool qbt_ctx::builder_trim()
{
    if (!build || build->active || build->blocks.empty()) return false;
    (void) hipStreamSynchronize(stream);
    for (auto& b : build->blocks) dev_free(b);
    build->blocks.clear();
    return true;
}

namespace {

constexpr int kMaxBuildLevels = 21; /* 3 bits per level in a 64-bit key */

__global__ void k_paths(const float* __restrict__ xyz, size_t n, float edge, int L, uint64_t* __restrict__ keys,
                        uint32_t* __restrict__ vals, unsigned int* __restrict__ bad)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float    x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    float    size = edge;
    uint64_t key  = 0;
    bool     ok   = true;
    for (int l = 0; l < L; l++) {
        size = size / 2.0f;
        float qx = x / size, qy = y / size, qz = z / size;
        /* outside the int range (or NaN) the reference's conversion is undefined */
        ok = ok && (qx > -2147483520.0f && qx < 2147483520.0f) && (qy > -2147483520.0f && qy < 2147483520.0f) &&
             (qz > -2147483520.0f && qz < 2147483520.0f);
        int o  = ok ? (int) qx % 2 : 0;
        int yi = ok ? (int) qy % 2 : 1;
        int zi = ok ? (int) qz % 2 : 1;
        if (yi == 0) o += 2;
        if (zi == 0) o += 4;
        if (o < 0) ok = false;
        key = (key << 3) | (uint64_t) (o & 7);
    }
    if (!ok) atomicAdd(bad, 1u);
    keys[i] = key;
    vals[i] = (uint32_t) i;
}

/* Nodes per depth from the sorted keys in one pass: point i starts a new
 * node at every depth from d0 down to L, where d0 is the first octet in
 * which its key differs from the previous one (depth 1 is the top 3 bits;
 * an equal key starts none). bins[d0] counts them; the number of nodes at
 * depth d is the sum of bins[1..d]. */
__global__ void k_depth_counts(const uint64_t* __restrict__ keys, size_t n, int L, unsigned* __restrict__ bins)
{
    __shared__ unsigned h[32];
    if (threadIdx.x < 32) h[threadIdx.x] = 0u;
    __syncthreads();
    /* grid-stride over the keys (a bounded grid: the global adds at the end
     * are few); one LDS add per (wave, depth): most points share one bin */
    const size_t stride = (size_t) gridDim.x * blockDim.x;
    for (size_t base = (size_t) blockIdx.x * blockDim.x; base < n; base += stride) {
        const size_t i  = base + threadIdx.x;
        int          d0 = 0;
        if (i < n) {
            d0 = 1;
            if (i > 0) {
                const uint64_t x = keys[i] ^ keys[i - 1];
                d0 = x == 0ull ? 0 : L - (63 - __clzll((long long) x)) / 3;
            }
        }
        for (int d = 1; d <= L; d++) {
            const uint64_t b = __ballot(d0 == d);
            if ((threadIdx.x & 63) == 0 && b) atomicAdd(&h[d], (unsigned) __popcll(b));
        }
    }
    __syncthreads();
    if (threadIdx.x >= 1 && threadIdx.x <= (unsigned) L && h[threadIdx.x]) atomicAdd(&bins[threadIdx.x], h[threadIdx.x]);
}
__global__ void k_shift_keys(const uint64_t* __restrict__ in, size_t n, uint64_t* __restrict__ out)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] >> 3;
}

/* head[j] = 1 where the parent key of child group j starts a new parent group */
__global__ void k_heads(const uint64_t* __restrict__ pk, size_t n, uint32_t* __restrict__ head)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) head[i] = (i == 0 || pk[i] != pk[i - 1]) ? 1u : 0u;
}

__global__ void k_fill_u32v(uint32_t* p, size_t n, uint32_t v)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_dmin(const uint32_t* __restrict__ first, size_t n, uint32_t depth, uint32_t* __restrict__ dmin)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicMin(&dmin[first[i]], depth);
}

__global__ void k_counts(const uint32_t* __restrict__ dmin, size_t n, uint32_t L, uint32_t* __restrict__ cnt)
{
    size_t i = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = dmin[i] <= L ? L + 1 - dmin[i] : 0u;
}

/* node index of every group at one depth, slot 8, and the parent's child slot */
__global__ void k_link(const uint64_t* __restrict__ ukey, const uint32_t* __restrict__ first, size_t n, uint32_t depth,
                       const uint32_t* __restrict__ base, const uint32_t* __restrict__ dmin,
                       const uint32_t* __restrict__ parent_pos, const uint32_t* __restrict__ parent_node,
                       uint32_t* __restrict__ node_out, int* __restrict__ child, int* __restrict__ model)
{
    size_t j = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t f   = first[j];
    uint32_t idx = 1u + base[f] + (depth - dmin[f]);
    node_out[j]  = idx;
    model[idx]   = (int) f;
    uint32_t par = depth == 1 ? 0u : parent_node[parent_pos[j] - 1u]; /* parent_pos is 1-based */
    child[8 * (size_t) par + (ukey[j] & 7u)] = (int) idx;
}

/* The compact layout (layout.hip's format) of the built tree, straight from
 * the groups: the groups of a depth in key order are that depth's nodes in
 * breadth-first hierarchical order, so group j of depth d is record
 * off_d + j, and the children of a node are the consecutive groups of the
 * next depth with its position as parent. Per group: its record (children
 * fill it from the next depth's launch), slot 8, reference index, and its
 * bit and, for the first child, the block start in the parent's record. */
__global__ void k_cmp_root(uint2* __restrict__ rec, int* __restrict__ model_c, int* __restrict__ ref,
                           int* __restrict__ r2c)
{
    rec[0]     = make_uint2(0u, 0u);
    model_c[0] = 0;
    ref[0]     = 0;
    r2c[0]     = 0;
}

__global__ void k_cmp_depth(const uint64_t* __restrict__ ukey, const uint32_t* __restrict__ first,
                            const uint32_t* __restrict__ node, const uint32_t* __restrict__ parent_pos, size_t n,
                            uint32_t off_parent, uint32_t off_here, uint2* __restrict__ rec, int* __restrict__ model_c,
                            int* __restrict__ ref, int* __restrict__ r2c)
{
    size_t j = (size_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t ci = off_here + (uint32_t) j;
    rec[ci]           = make_uint2(0u, 0u);
    model_c[ci]       = (int) first[j];
    ref[ci]           = (int) node[j];
    r2c[node[j]]      = (int) ci;
    const uint32_t pp = parent_pos ? parent_pos[j] - 1u : 0u; /* parent_pos is 1-based */
    const uint32_t p  = off_parent + pp;
    atomicOr(&rec[p].y, 0x101u << (uint32_t) (ukey[j] & 7u)); /* live and alloc */
    if (j == 0 || (parent_pos && parent_pos[j] != parent_pos[j - 1])) rec[p].x = ci;
}

inline unsigned nblk(size_t n) { return (unsigned) ((n + 255) / 256); }

} // namespace

#define HIPCHK(x) QBT_HIPCHK(x)

/* one build's scratch, carved from the context's builder blocks */
constexpr size_t kBlockBytes = (size_t) 64 << 20;
constexpr size_t kKeepBytes  = (size_t) 2 << 30;
struct Scratch {
    qbt_ctx*    c;
    BuildState* bs;
    size_t      bi = 0, off = 0;
    bool        keep; /* keep the blocks for the next build (dynamic tree) */
    Scratch(qbt_ctx* cc, bool keep_blocks) : c(cc), bs(cc->build), keep(keep_blocks) { bs->active = true; }
    ~Scratch()
    {
        /* the caller has synchronised the stream (or no kernel uses a block
         * yet); a dynamic build keeps the first blocks up to kKeepBytes */
        size_t kept = 0, i = 0;
        if (keep)
            for (; i < bs->blocks.size() && kept + bs->blocks[i].cap <= kKeepBytes; i++) kept += bs->blocks[i].cap;
        for (size_t j = i; j < bs->blocks.size(); j++) c->dev_free(bs->blocks[j]);
        bs->blocks.resize(i);
        bs->active = false;
    }
    void* get(size_t bytes, int* rc)
    {
        bytes = ((bytes ? bytes : 16) + 255) & ~(size_t) 255;
        for (; bi < bs->blocks.size(); bi++, off = 0) {
            if (off + bytes <= bs->blocks[bi].cap) {
                void* p = (char*) bs->blocks[bi].p + off;
                off += bytes;
                *rc = QBT_OK;
                return p;
            }
        }
        DevBuf b;
        *rc = c->dev_alloc(b, bytes > kBlockBytes ? bytes : kBlockBytes);
        if (*rc) return nullptr;
        bs->blocks.push_back(b);
        bi  = bs->blocks.size() - 1;
        off = bytes;
        return b.p;
    }
};

extern "C" int qbt_build_octree(qbt_ctx* c, int tree, const float basecube[4], int levels, size_t npoints,
                                size_t* nodes_out)
{
    if (!c || !basecube) return QBT_EINVAL;
    if (tree != QBT_BUFFER_STATIC_OCTREE && tree != QBT_BUFFER_DYNAMIC_OCTREE) return QBT_EINVAL;
    if (levels < 0 || levels > kMaxBuildLevels) return QBT_ENOTSUP;
    const int t = tree == QBT_BUFFER_STATIC_OCTREE ? 0 : 1;
    if (npoints > c->pts[t].n || npoints >= (size_t) UINT32_MAX) return QBT_EINVAL;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t    s = c->stream;
    const uint32_t L = (uint32_t) levels;
    int            rc = QBT_OK;
    if ((rc = c->flush_pending(t, true))) return rc; /* the build replaces the array */
    if (!c->build) c->build = new BuildState();
    Scratch        sc(c, t == 1);
    size_t         n = npoints;

    /* 1. paths and stable sort */
    uint64_t*     k0   = (uint64_t*) sc.get(n * 8, &rc);
    uint64_t*     k1   = rc ? nullptr : (uint64_t*) sc.get(n * 8, &rc);
    uint32_t*     v0   = rc ? nullptr : (uint32_t*) sc.get(n * 4, &rc);
    uint32_t*     v1   = rc ? nullptr : (uint32_t*) sc.get(n * 4, &rc);
    unsigned int* bad  = rc ? nullptr : (unsigned int*) sc.get(64, &rc);
    uint32_t*     ucnt = rc ? nullptr : (uint32_t*) sc.get(64, &rc);
    if (rc) return rc;
    k_fill_u32v<<<1, 64, 0, s>>>(bad, 1, 0u);
    if (n) k_paths<<<nblk(n), 256, 0, s>>>((const float*) c->pts[t].xyz.p, n, basecube[3], levels, k0, v0, bad);
    HIPCHK(hipGetLastError());
    unsigned int nbad = 0;
    HIPCHK(hipMemcpyAsync(&nbad, bad, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (nbad) return QBT_ERANGE;

    size_t tmp_bytes = 0, tb;
    if (n && L) {
        HIPCHK(rocprim::radix_sort_pairs(nullptr, tb, k0, k1, v0, v1, n, 0, 3 * levels, s));
        tmp_bytes = tb;
    }
    /* the largest reduce_by_key / scan temp over the sizes used below */
    HIPCHK(rocprim::reduce_by_key(nullptr, tb, k0, v0, n, k1, v1, ucnt, rocprim::minimum<uint32_t>(),
                                  rocprim::equal_to<uint64_t>(), s));
    tmp_bytes = tb > tmp_bytes ? tb : tmp_bytes;
    HIPCHK(rocprim::exclusive_scan(nullptr, tb, v0, v1, 0u, n, rocprim::plus<uint32_t>(), s));
    tmp_bytes = tb > tmp_bytes ? tb : tmp_bytes;
    HIPCHK(rocprim::inclusive_scan(nullptr, tb, v0, v1, n, rocprim::plus<uint32_t>(), s));
    tmp_bytes = tb > tmp_bytes ? tb : tmp_bytes;
    void* tmp = sc.get(tmp_bytes, &rc);
    if (rc) return rc;

    const uint64_t* skeys = k0;
    const uint32_t* svals = v0;
    if (n && L) {
        tb = tmp_bytes;
        HIPCHK(rocprim::radix_sort_pairs(tmp, tb, k0, k1, v0, v1, n, 0, 3 * levels, s));
        skeys = k1;
        svals = v1;
    }

    /* 2. groups per depth, leaves first: ukey[d], first[d] (d = 1..L) */
    std::vector<uint64_t*> ukey(L + 2, nullptr);
    std::vector<uint32_t*> first(L + 2, nullptr);
    std::vector<uint32_t*> ppos(L + 2, nullptr); /* parent position of each group at depth d */
    std::vector<size_t>    cnt(L + 2, 0);
    if (n && L) {
        /* the group count of every depth at once: one readback for the build */
        unsigned* bins = (unsigned*) sc.get(64 * 4, &rc);
        if (rc) return rc;
        k_fill_u32v<<<1, 64, 0, s>>>(bins, 64, 0u);
        k_depth_counts<<<std::min<size_t>(nblk(n), 1024), 256, 0, s>>>(skeys, n, (int) L, bins);
        HIPCHK(hipGetLastError());
        unsigned hb[64];
        HIPCHK(hipMemcpyAsync(hb, bins, sizeof(hb), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        size_t acc = 0;
        for (uint32_t d = 1; d <= L; d++) cnt[d] = acc += hb[d];
        ukey[L]  = (uint64_t*) sc.get(n * 8, &rc);
        first[L] = rc ? nullptr : (uint32_t*) sc.get(n * 4, &rc);
        if (rc) return rc;
        tb = tmp_bytes;
        HIPCHK(rocprim::reduce_by_key(tmp, tb, skeys, svals, n, ukey[L], first[L], ucnt, rocprim::minimum<uint32_t>(),
                                      rocprim::equal_to<uint64_t>(), s));
        for (uint32_t d = L; d > 1; d--) {
            size_t    m  = cnt[d];
            uint64_t* pk = (uint64_t*) sc.get(m * 8, &rc);
            uint32_t* hd = rc ? nullptr : (uint32_t*) sc.get(m * 4, &rc);
            ppos[d]      = rc ? nullptr : (uint32_t*) sc.get(m * 4, &rc);
            if (rc) return rc;
            k_shift_keys<<<nblk(m), 256, 0, s>>>(ukey[d], m, pk);
            k_heads<<<nblk(m), 256, 0, s>>>(pk, m, hd);
            HIPCHK(hipGetLastError());
            tb = tmp_bytes;
            HIPCHK(rocprim::inclusive_scan(tmp, tb, hd, ppos[d], m, rocprim::plus<uint32_t>(), s));
            ukey[d - 1]  = (uint64_t*) sc.get(m * 8, &rc);
            first[d - 1] = rc ? nullptr : (uint32_t*) sc.get(m * 4, &rc);
            if (rc) return rc;
            tb = tmp_bytes;
            HIPCHK(rocprim::reduce_by_key(tmp, tb, pk, first[d], m, ukey[d - 1], first[d - 1], ucnt,
                                          rocprim::minimum<uint32_t>(), rocprim::equal_to<uint64_t>(), s));
        }
    }

    /* 3. per-point node counts, exclusive scan in index order */
    uint32_t* dmin = (uint32_t*) sc.get(n * 4, &rc);
    uint32_t* pcnt = rc ? nullptr : (uint32_t*) sc.get(n * 4, &rc);
    uint32_t* base = rc ? nullptr : (uint32_t*) sc.get((n + 1) * 4, &rc);
    if (rc) return rc;
    size_t total = 1;
    if (n && L) {
        k_fill_u32v<<<nblk(n), 256, 0, s>>>(dmin, n, L + 1);
        for (uint32_t d = 1; d <= L; d++) k_dmin<<<nblk(cnt[d]), 256, 0, s>>>(first[d], cnt[d], d, dmin);
        k_counts<<<nblk(n), 256, 0, s>>>(dmin, n, L, pcnt);
        HIPCHK(hipGetLastError());
        tb = tmp_bytes;
        HIPCHK(rocprim::exclusive_scan(tmp, tb, pcnt, base, 0u, n, rocprim::plus<uint32_t>(), s));
        for (uint32_t d = 1; d <= L; d++) total += cnt[d];
    }
    if (total >= (size_t) INT32_MAX) return QBT_ENOTSUP;

    /* 4. the node arrays: zero, then link depth by depth (parents first) */
    TreeBuf& tr = c->tree[t];
    if (total > tr.cap_nodes) {
        size_t cap = total + total / 4 + 1024;
        c->dev_free(tr.child);
        c->dev_free(tr.model);
        tr.cap_nodes = 0;
        tr.n         = 0;
        if ((rc = c->dev_alloc(tr.child, cap * 32)) || (rc = c->dev_alloc(tr.model, cap * 4))) return rc;
        tr.cap_nodes = cap;
    }
    k_fill_u32v<<<nblk(total * 8), 256, 0, s>>>((uint32_t*) tr.child.p, total * 8, 0u);
    k_fill_u32v<<<nblk(total), 256, 0, s>>>((uint32_t*) tr.model.p, total, 0u);
    HIPCHK(hipGetLastError());
    /* 5. the compact layout, unless its 30-bit first-child indices overflow */
    CompactTree& cm  = tr.cmp;
    const bool   cmp = total < (1u << 30);
    cm.ok = false, cm.n = 0, cm.ref_n = 0;
    tr.n  = total; /* compact_reserve sizes the reference map by it */
    if (cmp) {
        if ((rc = c->compact_reserve(t, total, total))) return rc;
        k_cmp_root<<<1, 1, 0, s>>>((uint2*) cm.rec.p, (int*) cm.model.p, (int*) cm.ref.p, (int*) cm.r2c.p);
        HIPCHK(hipGetLastError());
    }
    if (n && L) {
        std::vector<uint32_t*> node(L + 2, nullptr);
        for (uint32_t d = 1; d <= L; d++) {
            node[d] = (uint32_t*) sc.get(cnt[d] * 4, &rc);
            if (rc) return rc;
        }
        for (uint32_t d = 1; d <= L; d++) {
            const uint32_t* pp = d == 1 ? nullptr : ppos[d];
            const uint32_t* pn = d == 1 ? nullptr : node[d - 1];
            k_link<<<nblk(cnt[d]), 256, 0, s>>>(ukey[d], first[d], cnt[d], d, base, dmin, pp, pn, node[d],
                                               (int*) tr.child.p, (int*) tr.model.p);
            HIPCHK(hipGetLastError());
        }
        if (cmp) {
            uint32_t off_parent = 0, off_here = 1; /* the root is record 0 */
            for (uint32_t d = 1; d <= L; d++) {
                k_cmp_depth<<<nblk(cnt[d]), 256, 0, s>>>(ukey[d], first[d], node[d], d == 1 ? nullptr : ppos[d],
                                                         cnt[d], off_parent, off_here, (uint2*) cm.rec.p,
                                                         (int*) cm.model.p, (int*) cm.ref.p, (int*) cm.r2c.p);
                HIPCHK(hipGetLastError());
                off_parent = off_here;
                off_here += (uint32_t) cnt[d];
            }
        }
    }
    HIPCHK(hipStreamSynchronize(s));
    tr.validated  = false;
    tr.full_check = true;
    tr.patch.clear();
    tr.cmp_stale = !cmp;
    if (cmp) cm.ok = true, cm.n = total, cm.ref_n = total;
    if (nodes_out) *nodes_out = total;
    return QBT_OK;
}
