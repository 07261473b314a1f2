/*
 * Flat point-model files and the qmc quantisation step that produces them.
 *
 * .pnt/.nrm/.col: raw little-endian float32 triples, one per point, the same
 * count in each file (src/qubatron/model.c:53-83). .rng: int32 triples
 * (first point index, x cell, y cell) written whenever the (x, y) cell column
 * changes in the sorted order (src/qubatron/qmc.c:291-305).
 */
#include "qbt_host.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int64_t qbt_flat_count(const char* pnt_path)
{
    if (!pnt_path) return QBT_EINVAL;
    FILE* f = fopen(pnt_path, "rb");
    if (!f) return QBT_EIO;
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return QBT_EIO;
    }
    long bytes = ftell(f);
    fclose(f);
    if (bytes < 0) return QBT_EIO;
    return (int64_t) (bytes / (3 * (long) sizeof(float)));
}

static int read_exact(const char* path, void* dst, size_t bytes)
{
    FILE* f = fopen(path, "rb");
    if (!f) return QBT_EIO;
    size_t got = fread(dst, 1, bytes, f);
    fclose(f);
    return got == bytes ? QBT_OK : QBT_EIO;
}

int64_t qbt_flat_read(const char* pnt_path, const char* col_path, const char* nrm_path, float* pnt, float* col,
                      float* nrm, size_t cap)
{
    if (!pnt_path || !col_path || !nrm_path || !pnt || !col || !nrm) return QBT_EINVAL;
    int64_t n = qbt_flat_count(pnt_path);
    if (n < 0) return n;
    if ((size_t) n > cap) return QBT_EINVAL;
    size_t bytes = (size_t) n * 3 * sizeof(float);
    int    rc;
    if ((rc = read_exact(pnt_path, pnt, bytes))) return rc;
    if ((rc = read_exact(col_path, col, bytes))) return rc;
    if ((rc = read_exact(nrm_path, nrm, bytes))) return rc;
    return n;
}

static int write_all(const char* path, const void* src, size_t bytes)
{
    FILE* f = fopen(path, "wb");
    if (!f) return QBT_EIO;
    size_t put = fwrite(src, 1, bytes, f);
    int    ok  = fclose(f) == 0 && put == bytes;
    return ok ? QBT_OK : QBT_EIO;
}

int qbt_flat_write(const char* pnt_path, const char* col_path, const char* nrm_path, const float* pnt,
                   const float* col, const float* nrm, size_t n)
{
    if (!pnt_path || !col_path || !nrm_path || (n && (!pnt || !col || !nrm))) return QBT_EINVAL;
    size_t bytes = n * 3 * sizeof(float);
    int    rc;
    if ((rc = write_all(pnt_path, pnt, bytes))) return rc;
    if ((rc = write_all(col_path, col, bytes))) return rc;
    return write_all(nrm_path, nrm, bytes);
}

typedef struct cellkey {
    float  ix, iy, iz;
    size_t src;
} cellkey;

static int cmp_cell(const void* a, const void* b)
{
    const cellkey* p = (const cellkey*) a;
    const cellkey* q = (const cellkey*) b;
    if (p->ix != q->ix) return p->ix < q->ix ? -1 : 1;
    if (p->iy != q->iy) return p->iy < q->iy ? -1 : 1;
    if (p->iz != q->iz) return p->iz < q->iz ? -1 : 1;
    return p->src < q->src ? -1 : (p->src > q->src); /* ties keep input order */
}

int64_t qbt_qmc_quantize(float* pnt, float* col, float* nrm, size_t n, float size, int levels, int32_t* rng_out,
                         size_t* rng_count)
{
    if ((n && (!pnt || !col || !nrm)) || levels < 0 || levels > 29 || !(size > 0.0f)) return QBT_EINVAL;
    int division = 2;
    for (int i = 0; i < levels; i++) division *= 2;
    float    precision = (float) size / (float) division; /* qmc.c:217-218 */
    cellkey* keys      = (cellkey*) malloc(sizeof(cellkey) * (n ? n : 1));
    if (!keys) return QBT_ENOMEM;
    size_t kept = 0;
    for (size_t i = 0; i < n; i++) {
        float xi = floorf(pnt[3 * i] / precision);
        float yi = floorf(pnt[3 * i + 1] / precision);
        float zi = floorf(pnt[3 * i + 2] / precision);
        if (xi >= 0.0f && xi < (float) division && yi >= 0.0f && yi < (float) division && zi >= 0.0f &&
            zi < (float) division) /* qmc.c:62-88: out-of-cube points dropped */
            keys[kept++] = (cellkey){xi, yi, zi, i};
    }
    qsort(keys, kept, sizeof(cellkey), cmp_cell);

    float* tp = (float*) malloc(sizeof(float) * 9 * (kept ? kept : 1));
    if (!tp) {
        free(keys);
        return QBT_ENOMEM;
    }
    size_t  npi = 0, nr = 0;
    cellkey prev = {0.0f, 0.0f, 0.0f, 0};
    for (size_t i = 0; i < kept; i++) {
        cellkey a = keys[i];
        if (!(i > 0 && prev.ix == a.ix && prev.iy == a.iy && prev.iz == a.iz)) {
            /* column change vs the previous record, including the zero-initialised
             * "previous" of the first point (qmc.c:293-305) */
            if (prev.ix != a.ix || prev.iy != a.iy) {
                if (rng_out) {
                    rng_out[3 * nr]     = (int32_t) npi;
                    rng_out[3 * nr + 1] = (int32_t) a.ix;
                    rng_out[3 * nr + 2] = (int32_t) a.iy;
                }
                nr++;
            }
            memcpy(tp + 9 * npi, pnt + 3 * a.src, 3 * sizeof(float));
            memcpy(tp + 9 * npi + 3, col + 3 * a.src, 3 * sizeof(float));
            memcpy(tp + 9 * npi + 6, nrm + 3 * a.src, 3 * sizeof(float));
            npi++;
        }
        prev = a;
    }
    for (size_t i = 0; i < npi; i++) {
        memcpy(pnt + 3 * i, tp + 9 * i, 3 * sizeof(float));
        memcpy(col + 3 * i, tp + 9 * i + 3, 3 * sizeof(float));
        memcpy(nrm + 3 * i, tp + 9 * i + 6, 3 * sizeof(float));
    }
    free(tp);
    free(keys);
    if (rng_count) *rng_count = nr;
    return (int64_t) npi;
}
