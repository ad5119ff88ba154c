// tdw_checkpoint.cpp -- td checkpoint files in the layout of
// hpc-engine/tdwslda_checkpoint.c (one rank per file, HowMany = 1).
//
// File checkpoint_<r> (r = rank; 0 here), native byte order, no padding
// between fields (tdwslda_checkpoint.c:230-280):
//   0   double t0
//   8   double mu[2]
//   24  double ec
//   32  double kF
//   40  int    nwf
//   44  double eF
//   52  double Effg
//   60  double beta
//   68  fbeta          (nwf + p) doubles
//       kz  (dim <= 2) (nwf + p) doubles   first half of the kt array
//       ky  (dim == 1) (nwf + p) doubles   second half
//       wf, fkm1, fkm2, fkm3 [, fkm4 if scheme >= AB4AM5] [, fkm5 if >= AB5AM5]
//                      2 (nwf + p) N complex doubles each
//       potentials     12 N doubles
// where each per-state section is sized for nwf + p states and only the
// first nwf are written; p = 1 for a single rank (tdwslda_checkpoint.c:282-283:
// rank 0 is not below nwf % 1).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "tdw_internal.hpp"

namespace {

struct File {
    FILE *f = nullptr;
    ~File() { if (f) fclose(f); }
};

constexpr long HEADER = 8 + 16 + 8 + 8 + 4 + 8 + 8 + 8;   // 68

// device <-> file through a bounded pinned staging buffer
int dev_to_file(FILE *f, long off, const void *d, size_t bytes, void *stage, size_t cap, hipStream_t st)
{
    if (fseek(f, off, SEEK_SET)) return TDW_EINVAL;
    const char *src = (const char *)d;
    for (size_t done = 0; done < bytes;) {
        const size_t m = bytes - done < cap ? bytes - done : cap;
        if (hipMemcpyAsync(stage, src + done, m, hipMemcpyDeviceToHost, st) != hipSuccess) return TDW_EINVAL;
        if (hipStreamSynchronize(st) != hipSuccess) return TDW_EINVAL;
        if (fwrite(stage, 1, m, f) != m) return TDW_EINVAL;
        done += m;
    }
    return 0;
}

int file_to_dev(FILE *f, long off, void *d, size_t bytes, void *stage, size_t cap, hipStream_t st)
{
    if (fseek(f, off, SEEK_SET)) return TDW_EINVAL;
    char *dst = (char *)d;
    for (size_t done = 0; done < bytes;) {
        const size_t m = bytes - done < cap ? bytes - done : cap;
        if (fread(stage, 1, m, f) != m) return TDW_EINVAL;
        if (hipMemcpyAsync(dst + done, stage, m, hipMemcpyHostToDevice, st) != hipSuccess) return TDW_EINVAL;
        if (hipStreamSynchronize(st) != hipSuccess) return TDW_EINVAL;
        done += m;
    }
    return 0;
}

struct Layout {
    long fbeta, kz, ky, arrays, pots;
    size_t arr_bytes, arr_span;
    int narr;
};

Layout layout(int dim, int scheme, int nwf, size_t N)
{
    Layout L;
    const long pad = nwf + 1;
    L.fbeta = HEADER;
    long off = HEADER + 8 * pad;
    L.kz = L.ky = -1;
    if (dim <= 2) { L.kz = off; off += 8 * pad; }
    if (dim == 1) { L.ky = off; off += 8 * pad; }
    L.arrays = off;
    L.arr_bytes = (size_t)2 * nwf * N * 16;
    L.arr_span = (size_t)2 * pad * N * 16;
    L.narr = 1 + 3 + (scheme >= TDW_AB4AM5 ? 1 : 0) + (scheme >= TDW_AB5AM5 ? 1 : 0);
    L.pots = L.arrays + (long)(L.narr * L.arr_span);
    return L;
}

struct Stage {
    void *p = nullptr;
    size_t cap = 0;
    ~Stage() { if (p) hipHostFree(p); }
    int alloc(size_t c)
    {
        cap = c;
        return hipHostMalloc(&p, c, 0) == hipSuccess ? 0 : TDW_ENOMEM;
    }
};

}  // namespace

extern "C" int tdw_checkpoint_nwf(const char *path, int *nwf)
{
    if (!path || !nwf) return TDW_EINVAL;
    File F;
    F.f = fopen(path, "rb");
    if (!F.f) return TDW_EINVAL;
    if (fseek(F.f, 40, SEEK_SET) || fread(nwf, sizeof(int), 1, F.f) != 1) return TDW_EINVAL;
    return *nwf > 0 ? 0 : TDW_EINVAL;
}

extern "C" int tdw_checkpoint_save(tdw_ctx *c, const char *path, tdw_state *s, double t0, const double *h_fbeta,
                                   const double *h_kt, const tdw_checkpoint_scalars *sc)
{
    if (!c || !path || !s || !h_fbeta || !sc) return TDW_EINVAL;
    if (c->dim <= 2 && !h_kt) return TDW_EINVAL;
    if (c->p.precision != TDW_FP64) return TDW_EINVAL;   // the file holds FP64 history entries
    int rc = tdw_finalize(c, s);
    if (rc) return rc;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return TDW_EINVAL;
    const int n = s->n;
    const Layout L = layout(c->dim, c->p.scheme, n, c->N);
    File F;
    F.f = fopen(path, "wb");
    if (!F.f) return TDW_EINVAL;
    // header
    char hdr[HEADER];
    long o = 0;
    auto put = [&](const void *v, size_t b) { memcpy(hdr + o, v, b); o += (long)b; };
    put(&t0, 8);
    put(&sc->mu_a, 8);
    put(&sc->mu_b, 8);
    put(&sc->ec, 8);
    put(&sc->kF, 8);
    put(&n, 4);
    put(&sc->eF, 8);
    put(&sc->Effg, 8);
    put(&sc->beta, 8);
    if (fwrite(hdr, 1, HEADER, F.f) != (size_t)HEADER) return TDW_EINVAL;
    // per-state sections: nwf values, then the padding slot (zeros)
    std::vector<double> sec((size_t)n + 1, 0.0);
    auto put_sec = [&](long off, const double *v) -> int {
        memcpy(sec.data(), v, sizeof(double) * n);
        sec[n] = 0.0;
        if (fseek(F.f, off, SEEK_SET) || fwrite(sec.data(), sizeof(double), n + 1, F.f) != (size_t)n + 1)
            return TDW_EINVAL;
        return 0;
    };
    if ((rc = put_sec(L.fbeta, h_fbeta))) return rc;
    if (L.kz >= 0 && (rc = put_sec(L.kz, h_kt))) return rc;
    if (L.ky >= 0 && (rc = put_sec(L.ky, h_kt + n))) return rc;
    Stage stg;
    if ((rc = stg.alloc((size_t)64 << 20))) return rc;
    const tdw_complex *arr[6] = {s->wf, s->f[0], s->f[1], s->f[2], s->f[3], s->f[4]};
    for (int k = 0; k < L.narr; k++) {
        const long off = L.arrays + (long)(k * L.arr_span);
        if ((rc = dev_to_file(F.f, off, arr[k], L.arr_bytes, stg.p, stg.cap, c->stream))) return rc;
        // zero padding slot, so the file has its full size
        const size_t padb = L.arr_span - L.arr_bytes;
        std::vector<char> z(padb < ((size_t)1 << 20) ? padb : ((size_t)1 << 20), 0);
        for (size_t done = 0; done < padb;) {
            const size_t m = padb - done < z.size() ? padb - done : z.size();
            if (fwrite(z.data(), 1, m, F.f) != m) return TDW_EINVAL;
            done += m;
        }
    }
    if ((rc = dev_to_file(F.f, L.pots, s->pots, sizeof(double) * 12 * c->N, stg.p, stg.cap, c->stream))) return rc;
    return 0;
}

extern "C" int tdw_checkpoint_load(tdw_ctx *c, const char *path, tdw_state *s, double *t0, double *h_fbeta,
                                   double *h_kt, tdw_checkpoint_scalars *sc)
{
    if (!c || !path || !s || !t0 || !h_fbeta || !sc) return TDW_EINVAL;
    if (c->dim <= 2 && !h_kt) return TDW_EINVAL;
    if (c->p.precision != TDW_FP64) return TDW_EINVAL;
    File F;
    F.f = fopen(path, "rb");
    if (!F.f) return TDW_EINVAL;
    char hdr[HEADER];
    if (fread(hdr, 1, HEADER, F.f) != (size_t)HEADER) return TDW_EINVAL;
    long o = 0;
    auto get = [&](void *v, size_t b) { memcpy(v, hdr + o, b); o += (long)b; };
    int nwf = 0;
    get(t0, 8);
    get(&sc->mu_a, 8);
    get(&sc->mu_b, 8);
    get(&sc->ec, 8);
    get(&sc->kF, 8);
    get(&nwf, 4);
    get(&sc->eF, 8);
    get(&sc->Effg, 8);
    get(&sc->beta, 8);
    if (nwf != s->n || nwf > c->nwf) return TDW_ESIZE;
    const int n = nwf;
    const Layout L = layout(c->dim, c->p.scheme, n, c->N);
    auto get_sec = [&](long off, double *v) -> int {
        if (fseek(F.f, off, SEEK_SET) || fread(v, sizeof(double), n, F.f) != (size_t)n) return TDW_EINVAL;
        return 0;
    };
    int rc;
    if ((rc = get_sec(L.fbeta, h_fbeta))) return rc;
    if (L.kz >= 0 && (rc = get_sec(L.kz, h_kt))) return rc;
    if (L.ky >= 0 && (rc = get_sec(L.ky, h_kt + n))) return rc;
    Stage stg;
    if ((rc = stg.alloc((size_t)64 << 20))) return rc;
    tdw_complex *arr[6] = {s->wf, s->f[0], s->f[1], s->f[2], s->f[3], s->f[4]};
    for (int k = 0; k < L.narr; k++) {
        const long off = L.arrays + (long)(k * L.arr_span);
        if ((rc = file_to_dev(F.f, off, arr[k], L.arr_bytes, stg.p, stg.cap, c->stream))) return rc;
    }
    if ((rc = file_to_dev(F.f, L.pots, s->pots, sizeof(double) * 12 * c->N, stg.p, stg.cap, c->stream))) return rc;
    // the loaded arrays are the exact state of the reference step
    c->pending = 0;
    return 0;
}
