/*
 * quotient_gpu.hip - device build of the plane-test quotient of
 * qubatron_amd/csrc/device/trace.hpp (plane_k), compared with the device's
 * IEEE division over whole rows of the significand table: every a in [1, 2)
 * against divisor significands d in [1, 2). Within the range where plane_k
 * takes its reciprocal path every step scales with the operands' exponents,
 * so the 2^23 x 2^23 table decides correctness for all such operands
 * (tests/test_gpu_quotient.py). Test infrastructure only; built with the
 * library's own flags (qubatron_amd/csrc/flags.mk).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "trace.hpp"

using namespace qbt;

namespace {

constexpr int kPerThread = 256; /* a significands per thread */

/* rows [mb0, mb0 + nrows): thread g handles row mb0 + g / 2^15 and the a
 * significands 256 (g mod 2^15) ... + 255 */
__global__ void __launch_bounds__(256) k_quot_rows(uint32_t mb0, uint32_t nrows, unsigned long long* bad,
                                                    uint32_t* ex, int max_ex)
{
    const uint64_t g   = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t row = (uint32_t) (g >> 15);
    if (row >= nrows) return;
    const uint32_t mb = mb0 + row;
    if (mb >= (1u << 23)) return;
    const float d = __uint_as_float(0x3F800000u | mb);
    Ray r{};
    r.dx = d, r.dy = 1.0f, r.dz = 1.0f;
    r.ix = recip_or_nan(d), r.iy = 1.0f, r.iz = 1.0f;
    const uint32_t ma0 = (uint32_t) (g & 0x7FFFu) * kPerThread;
    unsigned       nbad = 0;
    for (int k = 0; k < kPerThread; k++) {
        const float a = __uint_as_float(0x3F800000u | (ma0 + (uint32_t) k));
        const float t = plane_k<kSemShader>(0, a, r).w;
        const float q = a / d; /* correctly rounded (-fhip-fp32-correctly-rounded-divide-sqrt) */
        if (__float_as_uint(t) != __float_as_uint(q)) {
            const unsigned long long i = atomicAdd(bad, 1ull);
            if (i < (unsigned long long) max_ex) ex[2 * i] = __float_as_uint(a), ex[2 * i + 1] = __float_as_uint(d);
            nbad++;
        }
    }
    (void) nbad;
}

} // namespace

extern "C" {

/* Rows [mb0, mb0 + nrows) of the table, launched 4096 rows at a time on the
 * null stream; *bad receives the mismatching pairs, ex (2 x max_ex words)
 * the first ones' bit patterns. Returns 0, or -1 on a HIP error. */
int qg_rows(uint32_t mb0, uint32_t nrows, unsigned long long* bad, uint32_t* ex, int max_ex)
{
    unsigned long long* dbad = nullptr;
    uint32_t*           dex  = nullptr;
    if (hipMalloc(&dbad, sizeof(*dbad)) != hipSuccess) return -1;
    if (hipMalloc(&dex, 8 * (size_t) (max_ex > 0 ? max_ex : 1)) != hipSuccess) {
        (void) hipFree(dbad);
        return -1;
    }
    int rc = 0;
    if (hipMemset(dbad, 0, sizeof(*dbad)) != hipSuccess) rc = -1;
    for (uint32_t r = 0; rc == 0 && r < nrows; r += 4096) {
        const uint32_t n       = nrows - r < 4096 ? nrows - r : 4096;
        const uint64_t threads = (uint64_t) n << 15;
        k_quot_rows<<<(unsigned) (threads / 256), 256>>>(mb0 + r, n, dbad, dex, max_ex);
        if (hipGetLastError() != hipSuccess) rc = -1;
    }
    if (rc == 0 && hipDeviceSynchronize() != hipSuccess) rc = -1;
    if (rc == 0 && hipMemcpy(bad, dbad, sizeof(*dbad), hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    if (rc == 0 && max_ex > 0 && hipMemcpy(ex, dex, 8 * (size_t) max_ex, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    (void) hipFree(dbad);
    (void) hipFree(dex);
    return rc;
}

} // extern "C"
