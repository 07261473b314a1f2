/*
 * quotient.cpp - the plane-test quotient of qubatron_amd/csrc/device/trace.hpp
 * (plane_k: t = (c - p) / d from a per-ray reciprocal and two FMAs) compiled
 * for the host, compared with IEEE division (tests/test_quotient.py). Test
 * infrastructure only.
 */
#include <cstring>

#include "trace.hpp"

using namespace qbt;

namespace {

/* plane_k's t for a = c - 0 along axis x, and whether its fast path ran */
inline float plane_t(float a, float d, int* fast)
{
    Ray r{};
    r.dx = d, r.dy = 1.0f, r.dz = 1.0f;
    r.ix = recip_or_nan(d), r.iy = 1.0f, r.iz = 1.0f;
    if (fast) {
        const float q = a * r.ix;
        *fast = (std::fabs(q) >= 0x1p-80f && std::fabs(q) <= 0x1p+120f) ? 1 : 0;
    }
    return plane_k<kSemShader>(0, a, r).w;
}

inline uint32_t bits(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    return u;
}

inline float from_bits(uint32_t u)
{
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

} // namespace

extern "C" {

/* t[i] = plane_k's quotient of a[i] / d[i]; fast[i] = 1 where the
 * reciprocal path ran (d[i] != 0) */
void qc_quotients(const float* a, const float* d, size_t n, float* t, int* fast)
{
    for (size_t i = 0; i < n; i++) t[i] = plane_t(a[i], d[i], fast ? fast + i : nullptr);
}

/* Every significand of a in [1, 2) against the significands of d in [1, 2)
 * numbered mb_first + k * mb_step (k < mb_count; significand j is 1 + j 2^-23):
 * the number of pairs whose quotient differs from IEEE a / d, and the first
 * max_ex of them (a, d bit patterns) in ex. */
uint64_t qc_sweep(uint32_t mb_first, uint32_t mb_step, uint32_t mb_count, uint32_t* ex, int max_ex)
{
    uint64_t bad = 0;
    for (uint32_t k = 0; k < mb_count; k++) {
        const uint32_t mb = mb_first + k * mb_step;
        if (mb >= (1u << 23)) break;
        const float d = from_bits(0x3F800000u | mb);
        for (uint32_t ma = 0; ma < (1u << 23); ma++) {
            const float a = from_bits(0x3F800000u | ma);
            const float t = plane_t(a, d, nullptr);
            if (bits(t) != bits(a / d)) {
                if ((int64_t) bad < max_ex) ex[2 * bad] = bits(a), ex[2 * bad + 1] = bits(d);
                bad++;
            }
        }
    }
    return bad;
}

} // extern "C"
