// tdw_stio.cpp -- static-solver (st-wslda-3d) initial states for the td run
// (host).
//
// Restates the inittype 3 start of the td driver (pca.c:628-725) and its
// readers: check_comptibility (pca_io.h:382-417), read_checkpoint_info_pca
// (pca_io.h:340-373), scan_s3dpca_info_files (pca_io.h:1519-1556) and
// read_s3dpca_wf (pca_io.h:1562-1633).  Files in <prefix>/ (the new names
// first, the old s3dpca.* names when wf.info is absent):
//   wf.info | s3dpca.info           int dim (old: the precision 8), int nwf,
//                                   int nx, ny, nz, double dx, dy, dz, kF,
//                                   mu_a, mu_b, ec, beta
//   wf.%04d.info | s3dpca.%04d.info int dim, int nwf of that file
//   wf.%04d_un.wdat | s3dpca.%04d.wfu   u of the file's states (N complex each)
//   wf.%04d_vn.wdat | s3dpca.%04d.wfv   v
//   wf.%04d_en.dat  | s3dpca.%04d.en    their energies (double)
//   wf.pud | s3dpca.pud              12*N doubles of potentials
// Return codes are the reference's: negative for the wave-function reader,
// WSLDA_ERR_* (10002, 10006, 10007, 10030-10033) for the info files.
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include "tdw_internal.hpp"
#include "wslda_td.h"

namespace {

const int ERR_CANNOT_OPEN_FILE = 10002;       // wslda_errors.h:25
const int ERR_S3DPCA_INFO_FILES = 10006;      // :29
const int ERR_S3DPCA_INFO_MISSING = 10007;    // :30
const int ERR_INITSTATE_FOR_1D = 10030;       // :53
const int ERR_INITSTATE_FOR_2D = 10031;       // :54
const int ERR_INITSTATE_FOR_3D = 10032;       // :55
const int ERR_INITSTATE_NOT_FOUND = 10033;    // :56
const int OLD_FORMAT_TAG = 8;                 // PCA_PRECISION, pca_io.h:22

bool exists(const char *path)
{
    struct stat st;
    return stat(path, &st) == 0;
}

}  // namespace

extern "C" int check_comptibility(char *prefix, int requested_dim, char *suffix)
{
    char file_name[1024];
    snprintf(file_name, sizeof(file_name), "%s/wf.info", prefix);
    if (exists(file_name)) {
        strcpy(suffix, "wf");
    } else {
        snprintf(file_name, sizeof(file_name), "%s/s%ddpca.info", prefix, requested_dim);
        if (!exists(file_name)) return ERR_INITSTATE_NOT_FOUND;
        sprintf(suffix, "s%ddpca", requested_dim);
    }
    FILE *f = fopen(file_name, "rb");
    if (!f) return ERR_CANNOT_OPEN_FILE;
    int header_dim = 0;
    size_t got = fread(&header_dim, sizeof(int), 1, f);
    fclose(f);
    if (got != 1) return ERR_CANNOT_OPEN_FILE;
    if (header_dim == OLD_FORMAT_TAG) return 0;
    if (header_dim != requested_dim) {
        if (header_dim == 1) return ERR_INITSTATE_FOR_1D;
        if (header_dim == 2) return ERR_INITSTATE_FOR_2D;
        if (header_dim == 3) return ERR_INITSTATE_FOR_3D;
    }
    return 0;
}

extern "C" int read_checkpoint_info_pca(const char *file_name, int *nwf, int *nx, int *ny, int *nz, double *dx,
                                        double *dy, double *dz, double *kF, double *mu, double *ec, double *beta)
{
    FILE *f = fopen(file_name, "rb");
    if (!f) return ERR_CANNOT_OPEN_FILE;
    int prec = 0, werr = 0;
    // the reference reports the first field that could not be read
    if (fread(&prec, sizeof(int), 1, f) != 1) werr = -2;
    if (!werr && fread(nwf, sizeof(int), 1, f) != 1) werr = -4;
    if (!werr && fread(nx, sizeof(int), 1, f) != 1) werr = -5;
    if (!werr && fread(ny, sizeof(int), 1, f) != 1) werr = -6;
    if (!werr && fread(nz, sizeof(int), 1, f) != 1) werr = -7;
    if (!werr && fread(dx, sizeof(double), 1, f) != 1) werr = -8;
    if (!werr && fread(dy, sizeof(double), 1, f) != 1) werr = -9;
    if (!werr && fread(dz, sizeof(double), 1, f) != 1) werr = -10;
    if (!werr && fread(kF, sizeof(double), 1, f) != 1) werr = -14;
    if (!werr && fread(&mu[0], sizeof(double), 1, f) != 1) werr = -11;
    if (!werr && fread(&mu[1], sizeof(double), 1, f) != 1) werr = -11;
    if (!werr && fread(ec, sizeof(double), 1, f) != 1) werr = -12;
    if (!werr && fread(beta, sizeof(double), 1, f) != 1) werr = -13;
    fclose(f);
    return werr;
}

extern "C" int scan_s3dpca_info_files(const char *prefix, int number_of_files, int *nwf, int *nwf_per_file)
{
    char file_name[1024];
    int tnwf = 0;
    for (int ikz = 0; ikz < number_of_files; ikz++) {
        snprintf(file_name, sizeof(file_name), "%s/wf.%04d.info", prefix, ikz);
        if (!exists(file_name)) snprintf(file_name, sizeof(file_name), "%s/s3dpca.%04d.info", prefix, ikz);
        FILE *f = fopen(file_name, "rb");
        if (!f) return ERR_S3DPCA_INFO_MISSING;
        int info_dim = 0, i = 0;
        size_t g1 = fread(&info_dim, sizeof(int), 1, f);
        size_t g2 = fread(&i, sizeof(int), 1, f);
        fclose(f);
        if (g1 != 1 || g2 != 1) return ERR_S3DPCA_INFO_MISSING;
        // old-format files hold the precision here: accepted as they are
        if (info_dim != 3) {
            if (info_dim == 1) return ERR_INITSTATE_FOR_1D;
            if (info_dim == 2) return ERR_INITSTATE_FOR_2D;
        }
        nwf_per_file[ikz] = i;
        tnwf += i;
    }
    if (tnwf != *nwf) return ERR_S3DPCA_INFO_FILES;
    return 0;
}

namespace {

// read_s3dpca_wf for a lattice of N points
int read_s3d_wf(const char *prefix, int number_of_files, const int *nwf_per_file, int mylidx, int myuidx, size_t N,
                tdw_complex *h_wavefun, double *h_fbetaEn)
{
    char file_name[1024], fu_name[1024], fv_name[1024], fe_name[1024];
    const int nwfip = myuidx - mylidx;
    snprintf(file_name, sizeof(file_name), "%s/wf.%04d.info", prefix, 0);
    const bool new_format = exists(file_name);
    int iwf = 0;
    for (int ikz = 0; ikz < number_of_files; ikz++) {
        FILE *fu = nullptr, *fv = nullptr, *fe = nullptr;
        int err = 0;
        for (int ii = 0; ii < nwf_per_file[ikz] && !err; ii++, iwf++) {
            if (iwf < mylidx || iwf >= myuidx) continue;
            if (!fu) {
                if (new_format) {
                    snprintf(fu_name, sizeof(fu_name), "%s/wf.%04d_un.wdat", prefix, ikz);
                    snprintf(fv_name, sizeof(fv_name), "%s/wf.%04d_vn.wdat", prefix, ikz);
                    snprintf(fe_name, sizeof(fe_name), "%s/wf.%04d_en.dat", prefix, ikz);
                } else {
                    snprintf(fu_name, sizeof(fu_name), "%s/s3dpca.%04d.wfu", prefix, ikz);
                    snprintf(fv_name, sizeof(fv_name), "%s/s3dpca.%04d.wfv", prefix, ikz);
                    snprintf(fe_name, sizeof(fe_name), "%s/s3dpca.%04d.en", prefix, ikz);
                }
                fu = fopen(fu_name, "rb");
                fv = fopen(fv_name, "rb");
                fe = fopen(fe_name, "rb");
                if (!fu) err = -1;
                else if (!fv) err = -2;
                else if (!fe) err = -4;
                else if (fseek(fu, (long)(sizeof(tdw_complex) * N * ii), SEEK_SET) != 0) err = -11;
                else if (fseek(fv, (long)(sizeof(tdw_complex) * N * ii), SEEK_SET) != 0) err = -12;
                else if (fseek(fe, (long)(sizeof(double) * ii), SEEK_SET) != 0) err = -14;
                if (err) break;
            }
            const size_t slot = (size_t)(iwf - mylidx);
            if (fread(h_wavefun + N * slot, sizeof(tdw_complex) * N, 1, fu) != 1) err = -21;
            else if (fread(h_wavefun + N * slot + N * (size_t)nwfip, sizeof(tdw_complex) * N, 1, fv) != 1) err = -22;
            else if (fread(h_fbetaEn + slot, sizeof(double), 1, fe) != 1) err = -23;
        }
        if (fu) fclose(fu);
        if (fv) fclose(fv);
        if (fe) fclose(fe);
        if (err) return err;
    }
    return 0;
}

}  // namespace

// read_s3dpca_wf (pca_io.h:1562): N from the bound context (NXYZ)
extern "C" int read_s3dpca_wf(const char *prefix, int number_of_files, int *nwf_per_file, int mylidx, int myuidx,
                              tdw_complex *h_wavefun, double *h_fbetaEn)
{
    tdw_ctx *c = tdw_bound();
    if (!c) return TDW_ENOCTX;
    return read_s3d_wf(prefix, number_of_files, nwf_per_file, mylidx, myuidx, c->N, h_wavefun, h_fbetaEn);
}

extern "C" int tdw_st3d_info(const char *prefix, int iogroups, tdw_st_info *info, int *nwf_per_file)
{
    if (!prefix || !info || !nwf_per_file || iogroups <= 0) return TDW_EINVAL;
    char suffix[64], pre[1024], file_name[1200];
    snprintf(pre, sizeof(pre), "%s", prefix);
    int e = check_comptibility(pre, 3, suffix);
    if (e) return e;
    snprintf(file_name, sizeof(file_name), "%s/%s.info", prefix, suffix);
    double mu[2];
    e = read_checkpoint_info_pca(file_name, &info->nwf, &info->nx, &info->ny, &info->nz, &info->dx, &info->dy,
                                 &info->dz, &info->kF, mu, &info->ec, &info->beta);
    if (e) return e;
    info->mu_a = mu[0];
    info->mu_b = mu[1];
    info->new_format = strcmp(suffix, "wf") == 0;
    return scan_s3dpca_info_files(prefix, iogroups, &info->nwf, nwf_per_file);
}

extern "C" int tdw_st3d_load(const tdw_params *p, const char *prefix, int iogroups, const int *nwf_per_file,
                             const tdw_st_info *info, int lidx, int uidx, int spinsymmetry, tdw_complex *h_wf,
                             double *h_fbeta, double *h_pots, tdw_checkpoint_scalars *sc)
{
    if (!p || !prefix || !nwf_per_file || !info || !h_wf || !h_fbeta || lidx < 0 || uidx <= lidx ||
        uidx > info->nwf)
        return TDW_EINVAL;
    // the lattice of the files must be the run's (pca.c:642-650)
    if (p->dim != 3 || info->nx != p->nx || info->ny != p->ny || info->nz != p->nz || info->dx != p->dx ||
        info->dy != p->dy || info->dz != p->dz)
        return TDW_EINVAL;
    const size_t N = (size_t)p->nx * p->ny * p->nz;
    const int nwfip = uidx - lidx;
    int e = read_s3d_wf(prefix, iogroups, nwf_per_file, lidx, uidx, N, h_wf, h_fbeta);
    if (e) return e;
    if (h_pots) {
        char file_name[1200];
        snprintf(file_name, sizeof(file_name), "%s/%s.pud", prefix, info->new_format ? "wf" : "s3dpca");
        FILE *f = fopen(file_name, "rb");
        if (!f) return ERR_CANNOT_OPEN_FILE;
        const size_t got = fread(h_pots, sizeof(double), 12 * N, f);
        fclose(f);
        if (got != 12 * N) return TDW_EIO;
    }
    // particle numbers and the energy unit (pca.c:705-719, over the loaded
    // states: all of them on one GPU); the energies become the occupations
    // f(E) of the run
    double Na = 0.0, Nb = 0.0;
    for (int a = 0; a < nwfip; a++) {
        const double fb = tdw_fbeta(h_fbeta[a], info->beta);
        const tdw_complex *u = h_wf + N * (size_t)a, *v = h_wf + N * ((size_t)nwfip + a);
        for (size_t i = 0; i < N; i++) {
            Na += (pow(u[i].re, 2) + pow(u[i].im, 2)) * fb;
            Nb += (pow(v[i].re, 2) + pow(v[i].im, 2)) * (1.0 - fb);
        }
        h_fbeta[a] = fb;
    }
    if (sc) {
        const double eF = 0.5 * info->kF * info->kF;
        sc->mu_a = info->mu_a;
        sc->mu_b = info->mu_b;
        sc->ec = info->ec;
        sc->kF = info->kF;
        sc->eF = eF;
        sc->beta = info->beta;
        sc->Effg = 0.6 * (spinsymmetry == 1 ? Na + Nb : Na) * eF;
    }
    return 0;
}
