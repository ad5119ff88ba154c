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

// ---------------------------------------------------------------------------
// st-wslda-1d starts (inittype 1): the transverse modes of the 1D solver
// (wslda_wavevectors.h:25-100, 284-390) and its readers
// scan_stwslda1d_info_files (pca_io.h:845-912) and read_stwslda1d_wf
// (pca_io.h:1177-1372), for a 1D td run (codedim 1: each state evolved once
// with its mode's degeneracy), a 2D one (codedim 2: each state expanded over
// the mode's ky with kz >= 0 as u(x) e^{i ky y} / sqrt(LY), kz kept as the
// transverse momentum) or a 3D one (codedim 3: each state expanded over the
// mode's (ky, kz) as u(x) e^{i(ky y + kz z)} / sqrt(LY LZ)).
// Files <prefix>/wf.%04d.{info,_un.wdat,_vn.wdat,_kkyz.dat,_en.dat} or the
// old s1dpca.%04d.{info,wfu,wfv,kkyz,en}, one set per transverse mode.
// ---------------------------------------------------------------------------
namespace {

const double CMP_EPS = 1.0e-12;   // wslda_wavevectors.h:12
const int ERR_SCAN_INFO_FILES_SUM_FAILED = 10026;   // wslda_errors.h:49

double kgrid(int i, int n, double L)
{
    return (i < n / 2) ? 2. * M_PI / L * (double)i : 2. * M_PI / L * (double)(i - n);
}

// unique |k_perp|^2 classes of the (ky, kz) grid (Nyquist rows skipped
// unless cubic cutoff), first representative and degeneracy
int kmodes(const tdw_params *p, tdw_kmode *out, int max)
{
    const double LY = p->dy * p->ny, LZ = p->dz * p->nz;
    const bool cubic = p->regularization == TDW_CUBIC_CUTOFF;
    std::vector<double> k2s;
    int cnt = 0;
    for (int iy = 0; iy < p->ny; iy++)
        for (int iz = 0; iz < p->nz; iz++) {
            if (!cubic && p->ny > 1 && iy == p->ny / 2) continue;
            if (!cubic && p->nz > 1 && iz == p->nz / 2) continue;
            const double ky = kgrid(iy, p->ny, LY), kz = kgrid(iz, p->nz, LZ), k2 = ky * ky + kz * kz;
            int i = 0;
            for (; i < cnt; i++)
                if (fabs(k2s[i] - k2) < CMP_EPS) break;
            if (i < cnt) {
                if (out && i < max) out[i].weight++;
            } else {
                if (out && cnt < max) out[cnt] = tdw_kmode{ky, kz, 1};
                k2s.push_back(k2);
                cnt++;
            }
        }
    return cnt;
}

// members (ky, kz) of the mode with |k|^2 = ky^2 + kz^2 in grid order
// (wslda_kmodes_1d_get_modes, wslda_wavevectors.h:346-383)
void mode_members(const tdw_params *p, double ky0, double kz0, std::vector<double> &mky, std::vector<double> &mkz)
{
    const double LY = p->dy * p->ny, LZ = p->dz * p->nz, k2in = ky0 * ky0 + kz0 * kz0;
    const bool cubic = p->regularization == TDW_CUBIC_CUTOFF;
    for (int iy = 0; iy < p->ny; iy++)
        for (int iz = 0; iz < p->nz; iz++) {
            if (!cubic && p->ny > 1 && iy == p->ny / 2) continue;
            if (!cubic && p->nz > 1 && iz == p->nz / 2) continue;
            const double ky = kgrid(iy, p->ny, LY), kz = kgrid(iz, p->nz, LZ);
            if (fabs(ky * ky + kz * kz - k2in) < CMP_EPS) {
                mky.push_back(ky);
                mkz.push_back(kz);
            }
        }
}

// the (ky, kz) pairs a 2D run takes from one mode: the members grouped by kz
// in order of first appearance, kz >= 0 only (kz = -pi/dz too with cubic
// cutoff) (wslda_kmodes_1d_to_2d / _getcnt2d, wslda_wavevectors.h:390-470)
int modes_2d(const tdw_params *p, double ky0, double kz0, std::vector<double> &oky, std::vector<double> &okz)
{
    std::vector<double> mky, mkz, gz;
    std::vector<std::vector<double>> gy;
    mode_members(p, ky0, kz0, mky, mkz);
    for (size_t i = 0; i < mky.size(); i++) {
        size_t j = 0;
        for (; j < gz.size(); j++)
            if (fabs(mkz[i] - gz[j]) < CMP_EPS) break;
        if (j == gz.size()) {
            gz.push_back(mkz[i]);
            gy.push_back({});
        }
        gy[j].push_back(mky[i]);
    }
    const bool cubic = p->regularization == TDW_CUBIC_CUTOFF;
    for (size_t j = 0; j < gz.size(); j++) {
        if (!(gz[j] > -CMP_EPS || (cubic && fabs(gz[j] + M_PI / p->dz) < 1.0e-12))) continue;
        for (double y : gy[j]) {
            oky.push_back(y);
            okz.push_back(gz[j]);
        }
    }
    return (int)oky.size();
}

}  // namespace

extern "C" int tdw_st1d_kmodes(const tdw_params *p, tdw_kmode *kvecs, int max, int *cnt)
{
    if (!p || !cnt) return TDW_EINVAL;
    *cnt = kmodes(p, kvecs, kvecs ? max : 0);
    return 0;
}

extern "C" int scan_stwslda1d_info_files(const char *prefix, int codedim, int kvecs_to_consder, tdw_kmode *kvecs,
                                         int *nwf, int *nwf_per_kyz)
{
    char file_name[1024];
    int tnwf = 0;
    for (int ikz = 0; ikz < kvecs_to_consder; ikz++) {
        snprintf(file_name, sizeof(file_name), "%s/wf.%04d.info", prefix, ikz);
        if (!exists(file_name)) snprintf(file_name, sizeof(file_name), "%s/s1dpca.%04d.info", prefix, ikz);
        FILE *f = fopen(file_name, "rb");
        if (!f) return ERR_CANNOT_OPEN_FILE;
        int info_dim = 0, i = 0;
        size_t g1 = fread(&info_dim, sizeof(int), 1, f);
        size_t g2 = fread(&i, sizeof(int), 1, f);
        fclose(f);
        if (g1 != 1 || g2 != 1) return ERR_CANNOT_OPEN_FILE;
        if (info_dim == 2) return ERR_INITSTATE_FOR_2D;
        if (info_dim == 3) return ERR_INITSTATE_FOR_3D;
        nwf_per_kyz[ikz] = i;
        tnwf += i * kvecs[ikz].weight;
    }
    if (tnwf != *nwf) return ERR_SCAN_INFO_FILES_SUM_FAILED;
    tnwf = 0;
    for (int ikz = 0; ikz < kvecs_to_consder; ikz++) {
        int mult = 1;
        if (codedim == 3) mult = kvecs[ikz].weight;
        if (codedim == 2) {
            // the mode table of the bound context's transverse grid
            tdw_ctx *c = tdw_bound();
            if (!c) return TDW_ENOCTX;
            std::vector<double> oky, okz;
            mult = modes_2d(&c->p, kvecs[ikz].ky, kvecs[ikz].kz, oky, okz);
        }
        tnwf += nwf_per_kyz[ikz] * mult;
    }
    *nwf = tnwf;
    return 0;
}

// read_stwslda1d_wf: the lattice (NX and, for codedim 3, NY, NZ, DY, DZ and
// the cutoff) from the bound context
extern "C" int read_stwslda1d_wf(const char *prefix, int codedim, int kvecs_to_consder, tdw_kmode *kvecs,
                                 int *nwf_per_kyz, int mylidx, int myuidx, tdw_complex *h_wavefun,
                                 double *h_fbetaEn, double *h_kkyz, int *h_cnt)
{
    tdw_ctx *c = tdw_bound();
    if (!c) return TDW_ENOCTX;
    if (codedim < 1 || codedim > 3) return TDW_EINVAL;
    if (codedim != c->dim) return TDW_EINVAL;
    const tdw_params *p = &c->p;
    const int NX = p->nx;
    const size_t N = c->N;
    const int nwfip = myuidx - mylidx;
    std::vector<tdw_complex> u(NX), v(NX);
    std::vector<tdw_kmode> table(p->ny * p->nz + 1);
    const int ntab = kmodes(p, table.data(), (int)table.size());
    char name[1024], fu_n[1024], fv_n[1024], fk_n[1024], fe_n[1024];
    snprintf(name, sizeof(name), "%s/wf.%04d.info", prefix, 0);
    const bool new_format = exists(name);
    const double LY = p->dy * p->ny, LZ = p->dz * p->nz;
    const bool cubic = p->regularization == TDW_CUBIC_CUTOFF;
    int iwf = 0;
    for (int ikz = 0; ikz < kvecs_to_consder; ikz++) {
        // the (ky, kz) members of this mode (wslda_kmodes_1d_get_modes)
        std::vector<double> mky, mkz;
        if (codedim == 3) {
            mode_members(p, kvecs[ikz].ky, kvecs[ikz].kz, mky, mkz);
            if ((int)mky.size() != kvecs[ikz].weight) return -122;
        } else if (codedim == 2) {
            modes_2d(p, kvecs[ikz].ky, kvecs[ikz].kz, mky, mkz);
        }
        const int dcoeff = codedim == 1 ? 1 : (int)mky.size();
        FILE *fu = nullptr, *fv = nullptr, *fk = nullptr, *fe = nullptr;
        int err = 0;
        for (int dd = 0; dd < dcoeff && !err; dd++) {
            if (fu && dd >= 1) {
                if (fseek(fu, 0, SEEK_SET) || fseek(fv, 0, SEEK_SET) || fseek(fk, 0, SEEK_SET) || fseek(fe, 0, SEEK_SET))
                    err = -111;
            }
            for (int ii = 0; ii < nwf_per_kyz[ikz] && !err; ii++, iwf++) {
                if (iwf < mylidx || iwf >= myuidx) continue;
                if (!fu) {
                    if (new_format) {
                        snprintf(fu_n, sizeof(fu_n), "%s/wf.%04d_un.wdat", prefix, ikz);
                        snprintf(fv_n, sizeof(fv_n), "%s/wf.%04d_vn.wdat", prefix, ikz);
                        snprintf(fk_n, sizeof(fk_n), "%s/wf.%04d_kkyz.dat", prefix, ikz);
                        snprintf(fe_n, sizeof(fe_n), "%s/wf.%04d_en.dat", prefix, ikz);
                    } else {
                        snprintf(fu_n, sizeof(fu_n), "%s/s1dpca.%04d.wfu", prefix, ikz);
                        snprintf(fv_n, sizeof(fv_n), "%s/s1dpca.%04d.wfv", prefix, ikz);
                        snprintf(fk_n, sizeof(fk_n), "%s/s1dpca.%04d.kkyz", prefix, ikz);
                        snprintf(fe_n, sizeof(fe_n), "%s/s1dpca.%04d.en", prefix, ikz);
                    }
                    fu = fopen(fu_n, "rb");
                    fv = fopen(fv_n, "rb");
                    fk = fopen(fk_n, "rb");
                    fe = fopen(fe_n, "rb");
                    if (!fu) err = -1;
                    else if (!fv) err = -2;
                    else if (!fk) err = -3;
                    else if (!fe) err = -4;
                    else if (fseek(fu, (long)(sizeof(tdw_complex) * NX * ii), SEEK_SET)) err = -11;
                    else if (fseek(fv, (long)(sizeof(tdw_complex) * NX * ii), SEEK_SET)) err = -12;
                    else if (fseek(fk, (long)(sizeof(double) * ii * 2), SEEK_SET)) err = -13;
                    else if (fseek(fe, (long)(sizeof(double) * ii), SEEK_SET)) err = -14;
                    if (err) break;
                }
                double en, ky, kz;
                if (fread(u.data(), sizeof(tdw_complex) * NX, 1, fu) != 1) { err = -21; break; }
                if (fread(v.data(), sizeof(tdw_complex) * NX, 1, fv) != 1) { err = -22; break; }
                if (fread(&en, sizeof(double), 1, fe) != 1) { err = -23; break; }
                if (fread(&ky, sizeof(double), 1, fk) != 1) { err = -23; break; }
                if (fread(&kz, sizeof(double), 1, fk) != 1) { err = -24; break; }
                const size_t slot = (size_t)(iwf - mylidx);
                tdw_complex *U = h_wavefun + N * slot, *V = h_wavefun + N * slot + N * (size_t)nwfip;
                if (codedim == 1) {
                    for (int ix = 0; ix < NX; ix++) { U[ix] = u[ix]; V[ix] = v[ix]; }
                    h_fbetaEn[slot] = en;
                    h_kkyz[slot] = ky;
                    h_kkyz[slot + nwfip] = kz;
                    int w = 0;
                    for (int t = 0; t < ntab; t++)
                        if (fabs(table[t].ky - ky) < CMP_EPS && fabs(table[t].kz - kz) < CMP_EPS) { w = table[t].weight; break; }
                    h_cnt[slot] = w;
                } else if (codedim == 2) {
                    if (ky != kvecs[ikz].ky) { err = -44; break; }
                    if (kz != kvecs[ikz].kz) { err = -45; break; }
                    const double sky = mky[dd], norm = 1.0 / sqrt(LY);
                    size_t ixy = 0;
                    for (int ix = 0; ix < NX; ix++)
                        for (int iy = 0; iy < p->ny; iy++, ixy++) {
                            const double ph = sky * iy * p->dy;
                            const double cr = cos(ph) * norm, ci = sin(ph) * norm;
                            U[ixy].re = u[ix].re * cr - u[ix].im * ci;
                            U[ixy].im = u[ix].re * ci + u[ix].im * cr;
                            V[ixy].re = v[ix].re * cr - v[ix].im * ci;
                            V[ixy].im = v[ix].re * ci + v[ix].im * cr;
                        }
                    h_fbetaEn[slot] = en;
                    h_kkyz[slot] = mkz[dd];
                } else {
                    if (ky != kvecs[ikz].ky) { err = -54; break; }
                    if (kz != kvecs[ikz].kz) { err = -55; break; }
                    const double sky = mky[dd], skz = mkz[dd], norm = 1.0 / sqrt(LY * LZ);
                    size_t ixyz = 0;
                    for (int ix = 0; ix < NX; ix++)
                        for (int iy = 0; iy < p->ny; iy++)
                            for (int iz = 0; iz < p->nz; iz++, ixyz++) {
                                const double ph = sky * iy * p->dy + skz * iz * p->dz;
                                const double cr = cos(ph) * norm, ci = sin(ph) * norm;
                                U[ixyz].re = u[ix].re * cr - u[ix].im * ci;
                                U[ixyz].im = u[ix].re * ci + u[ix].im * cr;
                                V[ixyz].re = v[ix].re * cr - v[ix].im * ci;
                                V[ixyz].im = v[ix].re * ci + v[ix].im * cr;
                            }
                    h_fbetaEn[slot] = en;
                }
            }
        }
        if (fu) fclose(fu);
        if (fv) fclose(fv);
        if (fk) fclose(fk);
        if (fe) fclose(fe);
        if (err) return err;
    }
    return 0;
}
