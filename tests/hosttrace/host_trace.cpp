/*
 * host_trace.cpp - the GPU traversal code (qubatron_amd/csrc/device/trace.hpp)
 * compiled for the host, for CPU tests against the oracle. It builds the
 * direct and the compact (breadth-first) device layouts of a node array the
 * way the library does, and traces rays one by one with a per-ray stack.
 * Test infrastructure only.
 */
#include <cstring>
#include <vector>

#include "trace.hpp"

using namespace qbt;

namespace {

struct Layouts {
    std::vector<int4>  child; /* direct: 2 per node */
    std::vector<int>   model;
    std::vector<uint2> rec; /* compact */
    std::vector<int>   cmodel, ref;
};

/* direct split + breadth-first relayout (children contiguous, octet order) */
bool build(const int32_t* nodes, size_t len, Layouts& L)
{
    L.child.resize(2 * len);
    L.model.resize(len);
    for (size_t i = 0; i < len; i++) {
        const int32_t* n = nodes + 12 * i;
        L.child[2 * i]     = make_int4(n[0], n[1], n[2], n[3]);
        L.child[2 * i + 1] = make_int4(n[4], n[5], n[6], n[7]);
        L.model[i]         = n[8];
    }
    std::vector<int> order = {0};
    std::vector<char> seen(len, 0);
    seen[0] = 1;
    L.rec.assign(len, make_uint2(0, 0));
    for (size_t k = 0; k < order.size(); k++) {
        const int32_t* n     = nodes + 12 * (size_t) order[k];
        unsigned       mask  = 0;
        const unsigned first = (unsigned) order.size();
        for (int o = 0; o < 8; o++) {
            if (n[o] > 0) {
                if ((size_t) n[o] >= len || seen[n[o]]) return false;
                seen[n[o]] = 1;
                mask |= 1u << o;
                order.push_back(n[o]);
            }
        }
        L.rec[k] = make_uint2(mask ? first : 0u, mask | mask << 8);
    }
    L.rec.resize(order.size());
    L.cmodel.resize(order.size());
    L.ref.resize(order.size());
    for (size_t k = 0; k < order.size(); k++) {
        L.cmodel[k] = nodes[12 * (size_t) order[k] + 8];
        L.ref[k]    = order[k];
    }
    return true;
}

TreeDev tree_of(const Layouts& L, bool cmp)
{
    TreeDev t;
    t.child = L.child.data();
    t.model = cmp ? L.cmodel.data() : L.model.data();
    t.n     = (int) (cmp ? L.rec.size() : L.model.size());
    t.rec   = L.rec.data();
    t.ref   = L.ref.data();
    return t;
}

template <bool CMP, bool DUAL, bool GRID>
void run(const SceneDev& s, const int* ref, const float* org, const float* dir, size_t n, int32_t* status,
         int32_t* leaf, float* isp, int64_t* exps)
{
    std::vector<uint32_t> mem(3 * (size_t) (s.maxlevel > 0 ? s.maxlevel : 1));
    Stack                 st;
    const int             L = s.maxlevel > 0 ? s.maxlevel : 1;
    st.stride               = 1;
    st.pw                   = mem.data();
    st.ns                   = reinterpret_cast<int*>(mem.data()) + L;
    st.nd                   = DUAL ? reinterpret_cast<int*>(mem.data()) + 2 * L : st.ns;
    for (size_t i = 0; i < n; i++) {
        Ray r = {org[3 * i], org[3 * i + 1], org[3 * i + 2], dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
        Hit h;
        int stt   = trace<CMP, DUAL, false, GRID>(s, r, st, h);
        status[i] = stt;
        leaf[i]   = stt == ST_HIT ? (CMP ? ref[h.ns] : h.ns) : -1;
        exps[i]   = h.exps;
        const float4 p = stt == ST_HIT ? h.isp : make_float4(0.f, 0.f, 0.f, 0.f);
        isp[4 * i] = p.x, isp[4 * i + 1] = p.y, isp[4 * i + 2] = p.z, isp[4 * i + 3] = p.w;
    }
}

} // namespace

/* Trace n rays with the shader semantics. layout 0 = direct, 1 = compact
 * (both trees must be proper trees); grid selects the closed-form cube
 * coordinates (only valid where the library's grid check passes). dyn may be
 * NULL (root-only dynamic tree). Returns 0, or -1 if a layout cannot be
 * built. */
extern "C" int ht_trace(const int32_t* nodes, size_t len, const int32_t* dyn, size_t dlen, const float basecube[4],
                        int maxlevel, int layout, int grid, const float* org, const float* dir, size_t n,
                        int32_t* status, int32_t* leaf, float* isp, int64_t* exps)
{
    static const int32_t root[12] = {0};
    Layouts              S, D;
    if (!build(nodes, len, S)) return -1;
    const bool dual = dyn && dlen > 1;
    if (!build(dual ? dyn : root, dual ? dlen : 1, D)) return -1;
    const bool cmp = layout == 1;
    SceneDev   s;
    memset(&s, 0, sizeof(s));
    s.stat     = tree_of(S, cmp);
    s.dyn      = tree_of(D, cmp);
    s.bx       = basecube[0];
    s.by       = basecube[1];
    s.bz       = basecube[2];
    s.bw       = basecube[3];
    s.maxlevel = maxlevel;
    s.iter_cap = 1 << 24;
    const int* ref = S.ref.data();
#define QBT_RUN(C, D_, G) run<C, D_, G>(s, ref, org, dir, n, status, leaf, isp, exps)
    if (cmp) {
        if (dual) grid ? QBT_RUN(true, true, true) : QBT_RUN(true, true, false);
        else grid ? QBT_RUN(true, false, true) : QBT_RUN(true, false, false);
    } else {
        if (dual) grid ? QBT_RUN(false, true, true) : QBT_RUN(false, true, false);
        else grid ? QBT_RUN(false, false, true) : QBT_RUN(false, false, false);
    }
#undef QBT_RUN
    return 0;
}

/* The compact layout after in-place range updates that zeroed child slots:
 * the records and child blocks of `nodes` (the tree at the last relayout),
 * with the live bits and slot 8 of `live_nodes` (the same array with some
 * slots zeroed). Traces n rays on it (single tree) and on the direct layout
 * of live_nodes; both must equal the oracle on live_nodes. Returns -1 if a
 * layout cannot be built or live_nodes sets a slot the block lacks. */
extern "C" int ht_trace_patched(const int32_t* nodes, const int32_t* live_nodes, size_t len, const float basecube[4],
                                int maxlevel, int layout, int grid, const float* org, const float* dir, size_t n,
                                int32_t* status, int32_t* leaf, float* isp, int64_t* exps)
{
    static const int32_t root[12] = {0};
    Layouts              S, D, Lv;
    if (!build(nodes, len, S) || !build(root, 1, D)) return -1;
    for (size_t k = 0; k < S.rec.size(); k++) {
        const int32_t* nl    = live_nodes + 12 * (size_t) S.ref[k];
        const unsigned alloc = S.rec[k].y & 0xFFu;
        unsigned       live  = 0;
        for (int o = 0; o < 8; o++)
            if (nl[o] > 0) live |= 1u << o;
        if (live & ~alloc) return -1;
        S.rec[k].y  = live | alloc << 8;
        S.cmodel[k] = nl[8];
    }
    for (size_t i = 0; i < len; i++) { /* the direct layout is the edited array itself */
        const int32_t* nl  = live_nodes + 12 * i;
        S.child[2 * i]     = make_int4(nl[0], nl[1], nl[2], nl[3]);
        S.child[2 * i + 1] = make_int4(nl[4], nl[5], nl[6], nl[7]);
        S.model[i]         = nl[8];
    }
    const bool cmp = layout == 1;
    SceneDev   s;
    memset(&s, 0, sizeof(s));
    s.stat     = tree_of(S, cmp);
    s.dyn      = tree_of(D, cmp);
    s.bx       = basecube[0];
    s.by       = basecube[1];
    s.bz       = basecube[2];
    s.bw       = basecube[3];
    s.maxlevel = maxlevel;
    s.iter_cap = 1 << 24;
    const int* ref = S.ref.data();
    if (cmp) grid ? run<true, false, true>(s, ref, org, dir, n, status, leaf, isp, exps)
             : run<true, false, false>(s, ref, org, dir, n, status, leaf, isp, exps);
    else grid ? run<false, false, true>(s, ref, org, dir, n, status, leaf, isp, exps)
              : run<false, false, false>(s, ref, org, dir, n, status, leaf, isp, exps);
    (void) Lv;
    return 0;
}
