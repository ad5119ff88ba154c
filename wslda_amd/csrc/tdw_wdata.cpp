// tdw_wdata.cpp -- W-data measurement output of the td codes.
//
// A data set is a metadata file <prefix>.wtxt and one binary file
// <prefix>_<var>.wdat per variable, to which every measurement ("cycle")
// appends one datablock (lib/wdata/doc/README.md).  What is written follows
// the td drivers (pca.c:1021-1046, 1623-1630; cpca.c:889; ccpca.c:764):
//   create_wdata_metadata   wslda_writevars.c:109-253  variables from writevar
//   clear_files             :258-266                   refuse / clear old files
//   write_wdata_metadata_file :268-288                 the .wtxt, rewritten after
//                                                       every cycle
//   write_measurments       :298-452                   one datablock per variable
// and the W-data library calls they use (lib/wdata/c/wdata.c):
//   wdata_print_metadata :495-584 (text layout of the .wtxt), wdata_setconst
//   :1444-1462, wdata_add_datablock_wdat :1016-1035 (append), wdata_write_cycle_d
//   :1180-1203 (FP32 conversion for writeprec f), wdata_clear_database
//   :1521-1559.
//
// The GPU side is one copy of the 12*N densities and 12*N potentials to
// pinned host memory per cycle, plus V_ext evaluated on the device for the
// V - V_ext and V_ext variables.  The built-in external-potential model has
// no pairing field and no velocity field, so delta_ext is the pairing field
// itself (delta_ext(delta) = delta, the default callback of
// problem-definition.h) and velocity_ext is zero.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <string>
#include <vector>

#include "tdw_internal.hpp"
#include "tdw_hooks.hpp"

namespace {

constexpr int VMAJ = 0, VMIN = 3, VPATCH = 0;   // W-data format version written (wdata.h:18-20)

// where a variable's datablock comes from
enum Src {
    S_RHO_A, S_RHO_B, S_J_A, S_J_B, S_TAU_A, S_TAU_B, S_NU, S_DELTA, S_V_A, S_V_B, S_VEXT_A, S_VEXT_B,
    S_VELEXT_A, S_VELEXT_B, S_DELTAEXT, S_ALPHA_A, S_ALPHA_B, S_A_A, S_A_B
};

struct Var {
    std::string name, type, unit, format;
    Src src;
};

struct Const {
    std::string name, unit;
    double value;
};

}  // namespace

struct tdw_wdata {
    tdw_ctx *c = nullptr;
    std::string prefix, codename;
    int nx = 0, ny = 0, nz = 0, datadim = 0, cycles = 0, inittype = 0, overwrite = 1;
    double dx = 0, dy = 0, dz = 0, t0 = 0, dt = 0;
    char prec = 'd';
    std::vector<Var> vars;
    std::vector<std::pair<std::string, std::string>> links;
    std::vector<Const> consts;
    double *h = nullptr;   // pinned: 12N densities, 12N potentials, 2N V_ext
    ~tdw_wdata()
    {
        if (h) hipHostFree(h);
    }
};

namespace {

// wdata_variable_process_var (wslda_writevars.c:89-107): the type string of
// the chosen precision
std::string typed(char prec, char kind)
{
    if (prec == 'f') return kind == 'r' ? "real4" : kind == 'c' ? "complex8" : "vector4";
    return kind == 'r' ? "real" : kind == 'c' ? "complex" : "vector";
}

// one writevar key -> variables / links (wslda_writevars.c:133-248)
void add_key(tdw_wdata *w, const std::string &key, int spinsym)
{
    auto pair = [&](const char *a, const char *b, char kind, Src sa, Src sb) {
        w->vars.push_back({a, typed(w->prec, kind), "none", "wdat", sa});
        if (!spinsym) w->vars.push_back({b, typed(w->prec, kind), "none", "wdat", sb});
        else w->links.push_back({b, a});
    };
    auto one = [&](const char *a, char kind, Src s) { w->vars.push_back({a, typed(w->prec, kind), "none", "wdat", s}); };
    if (key == "density" || key == "rho") pair("rho_a", "rho_b", 'r', S_RHO_A, S_RHO_B);
    else if (key == "current" || key == "j") pair("j_a", "j_b", 'v', S_J_A, S_J_B);
    else if (key == "tau") pair("tau_a", "tau_b", 'r', S_TAU_A, S_TAU_B);
    else if (key == "V") pair("V_a", "V_b", 'r', S_V_A, S_V_B);
    else if (key == "V_ext") pair("V_ext_a", "V_ext_b", 'r', S_VEXT_A, S_VEXT_B);
    else if (key == "velocity_ext") pair("velocity_ext_a", "velocity_ext_b", 'v', S_VELEXT_A, S_VELEXT_B);
    else if (key == "delta") one("delta", 'c', S_DELTA);
    else if (key == "delta_ext") one("delta_ext", 'c', S_DELTAEXT);
    else if (key == "nu") one("nu", 'c', S_NU);
    else if (key == "alpha") pair("alpha_a", "alpha_b", 'r', S_ALPHA_A, S_ALPHA_B);
    else if (key == "A") pair("A_a", "A_b", 'v', S_A_A, S_A_B);
    // unknown keys add nothing, as in create_wdata_metadata
}

// writevar tokens with the parser's "all" / "default" expansion and
// de-duplication (pca_utils.c:495-535); no keys -> "rho delta j" (:561-566)
std::vector<std::string> writevar_keys(const char *list)
{
    std::vector<std::string> out;
    auto push = [&](const std::string &k) {
        for (auto &o : out)
            if (o == k) return;
        out.push_back(k);
    };
    std::string s = list ? list : "";
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && (s[i] == ' ' || s[i] == ',' || s[i] == '\t' || s[i] == '\n')) i++;
        size_t j = i;
        while (j < s.size() && !(s[j] == ' ' || s[j] == ',' || s[j] == '\t' || s[j] == '\n')) j++;
        if (j > i) {
            const std::string k = s.substr(i, j - i);
            if (k[0] == '#') break;
            if (k == "all")
                for (const char *a : {"rho", "delta", "j", "nu", "tau", "V", "V_ext", "delta_ext", "velocity_ext",
                                      "alpha", "A"})
                    push(a);
            else if (k == "default")
                for (const char *a : {"rho", "delta", "j"}) push(a);
            else push(k);
        }
        i = j;
    }
    if (out.empty())
        for (const char *a : {"rho", "delta", "j"}) out.push_back(a);
    return out;
}

std::string file_of(const tdw_wdata *w, const Var &v) { return w->prefix + "_" + v.name + "." + v.format; }

bool exists(const std::string &f)
{
    FILE *p = fopen(f.c_str(), "rb");
    if (!p) return false;
    fclose(p);
    return true;
}

// wdata_print_metadata (wdata.c:495-564), preceded by the td codes' own
// header line (wslda_writevars.c:282)
int write_wtxt(const tdw_wdata *w)
{
    const std::string fn = w->prefix + ".wtxt";
    FILE *f = fopen(fn.c_str(), "w");
    if (!f) return TDW_EIO;
    char stamp[20];
    time_t now = time(nullptr);
    struct tm tmv;
    localtime_r(&now, &tmv);
    strftime(stamp, sizeof stamp, "%x-%X", &tmv);
    fprintf(f, "# Generated by %s, engine version %s [%s]\n\n", w->codename.c_str(), "wslda-mi355x", stamp);
    fprintf(f, "# W-DATA %d.%d.%d\n", VMAJ, VMIN, VPATCH);
    fprintf(f, "# More information about the WData format see: "
               "https://gitlab.fizyka.pw.edu.pl/wtools/wdata/-/wikis/home\n");
    fprintf(f, "nx %24d   # lattice\n", w->nx);
    fprintf(f, "ny %24d   # lattice\n", w->ny);
    fprintf(f, "nz %24d   # lattice\n", w->nz);
    fprintf(f, "dx %24g   # spacing\n", w->dx);
    fprintf(f, "dy %24g   # spacing\n", w->dy);
    fprintf(f, "dz %24g   # spacing\n", w->dz);
    fprintf(f, "datadim %19d   # dimension of block size: 1=nx, 2=nx*ny, 3=nx*ny*nz\n", w->datadim);
    if (w->prefix.size() > 20) fprintf(f, "prefix %s   # prefix for files belonging to this data set`\n", w->prefix.c_str());
    else fprintf(f, "prefix %20s   # prefix for files belonging to this data set`\n", w->prefix.c_str());
    fprintf(f, "cycles %20d   # number of cycles (measurements)\n", w->cycles);
    fprintf(f, "t0 %24g   # time value for the first cycle\n", w->t0);
    fprintf(f, "dt %24g   # time interval between cycles\n", w->dt);
    if (!w->vars.empty()) {
        fprintf(f, "\n# variables\n");
        fprintf(f, "# tag                  name                    type                    unit                  format\n");
        for (auto &v : w->vars)
            fprintf(f, "var%24s%24s%24s%24s\n", v.name.c_str(), v.type.c_str(), v.unit.c_str(), v.format.c_str());
    }
    if (!w->links.empty()) {
        fprintf(f, "\n# links\n");
        fprintf(f, "# tag                  name                 link-to\n");
        for (auto &l : w->links) fprintf(f, "link%23s%24s\n", l.first.c_str(), l.second.c_str());
    }
    if (!w->consts.empty()) {
        fprintf(f, "\n# consts\n");
        fprintf(f, "# tag                  name                   value                    unit\n");
        for (auto &k : w->consts) fprintf(f, "const%22s%24g%24s\n", k.name.c_str(), k.value, k.unit.c_str());
    }
    fprintf(f, "\n");
    return fclose(f) == 0 ? 0 : TDW_EIO;
}

// append one datablock (wdata_add_datablock_wdat, wdata_write_cycle_d)
int append(const tdw_wdata *w, const Var &v, const double *d, size_t count)
{
    FILE *f = fopen(file_of(w, v).c_str(), "ab");
    if (!f) return TDW_EIO;
    size_t put;
    if (w->prec == 'f') {
        std::vector<float> t(count);
        for (size_t i = 0; i < count; i++) t[i] = (float)d[i];
        put = fwrite(t.data(), sizeof(float) * count, 1, f);
    } else {
        put = fwrite(d, sizeof(double) * count, 1, f);
    }
    const int rc = fclose(f);
    return (put == 1 && rc == 0) ? 0 : TDW_EIO;
}

}  // namespace

extern "C" int tdw_wdata_create(tdw_ctx *c, const tdw_wdata_params *p, tdw_wdata **out)
{
    if (!c || !p || !out || !p->outprefix || !p->outprefix[0]) return TDW_EINVAL;
    if (p->writeprec != 'd' && p->writeprec != 'f' && p->writeprec != 'D' && p->writeprec != 'F') return TDW_EINVAL;
    tdw_wdata *w = new tdw_wdata;
    w->c = c;
    w->prefix = p->outprefix;
    w->codename = p->codename ? p->codename : (c->dim == 1 ? "td-wslda-1d" : c->dim == 2 ? "td-wslda-2d" : "td-wslda-3d");
    // lattice of the code (pca.c:1023, cpca.c:889, ccpca.c:764: datadim = code dimension)
    w->nx = c->p.nx;
    w->ny = c->p.ny;
    w->nz = c->p.nz;
    w->dx = c->p.dx;
    w->dy = c->p.dy;
    w->dz = c->p.dz;
    w->datadim = c->dim;
    w->t0 = p->t0;
    w->dt = p->dt;
    w->prec = (p->writeprec == 'f' || p->writeprec == 'F') ? 'f' : 'd';
    w->inittype = p->inittype;
    w->overwrite = p->overwrite;
    for (auto &k : writevar_keys(p->writevar)) add_key(w, k, p->spinsymmetry);
    if (hipHostMalloc((void **)&w->h, sizeof(double) * 26 * c->N, 0) != hipSuccess) {
        w->h = nullptr;
        delete w;
        return TDW_ENOMEM;
    }
    *out = w;
    return 0;
}

extern "C" int tdw_wdata_setconst(tdw_wdata *w, const char *name, double value, const char *unit)
{
    if (!w || !name || !name[0]) return TDW_EINVAL;
    for (auto &k : w->consts)
        if (k.name == name) {
            k.value = value;
            if (unit) k.unit = unit;
            return 0;
        }
    w->consts.push_back({name, unit ? unit : "none", value});
    return 0;
}

extern "C" int tdw_wdata_begin(tdw_wdata *w)
{
    if (!w) return TDW_EINVAL;
    // clear_files: with overwrite = 0 an existing variable file is an error
    if (!w->overwrite)
        for (auto &v : w->vars)
            if (exists(file_of(w, v))) return TDW_EEXIST;
    for (auto &v : w->vars) remove(file_of(w, v).c_str());
    w->cycles = 0;
    return write_wtxt(w);
}

extern "C" int tdw_wdata_write(tdw_wdata *w, int it, const double *d_dens, const double *d_pots)
{
    if (!w || !d_dens || !d_pots) return TDW_EINVAL;
    tdw_ctx *c = w->c;
    const size_t N = c->N;
    double *hd = w->h, *hp = w->h + 12 * N, *hv = w->h + 24 * N;
    bool need_vext = false;
    for (auto &v : w->vars)
        need_vext |= v.src == S_V_A || v.src == S_V_B || v.src == S_VEXT_A || v.src == S_VEXT_B;
    if (need_vext) TDW_TRY(tdw_vext_eval(c, it, c->d_vext, c->d_vext + N));
    TDW_HIP(hipMemcpyAsync(hd, d_dens, sizeof(double) * 12 * N, hipMemcpyDeviceToHost, c->stream));
    TDW_HIP(hipMemcpyAsync(hp, d_pots, sizeof(double) * 12 * N, hipMemcpyDeviceToHost, c->stream));
    if (need_vext) TDW_HIP(hipMemcpyAsync(hv, c->d_vext, sizeof(double) * 2 * N, hipMemcpyDeviceToHost, c->stream));
    TDW_HIP(hipStreamSynchronize(c->stream));
    // V written as the mean-field part V - V_ext, except for the first cycle
    // of a run started from a static solution (wslda_writevars.c:341-353)
    const bool keep_v = it == 0 && w->inittype >= 1 && w->inittype <= 3;
    std::vector<double> tmp;
    for (auto &v : w->vars) {
        const double *src = nullptr;
        size_t cnt = N;
        switch (v.src) {
        case S_RHO_A: src = hd + 2 * N; break;
        case S_RHO_B: src = hd + 7 * N; break;
        case S_TAU_A: src = hd + 3 * N; break;
        case S_TAU_B: src = hd + 8 * N; break;
        case S_J_A: src = hd + 4 * N; cnt = 3 * N; break;
        case S_J_B: src = hd + 9 * N; cnt = 3 * N; break;
        case S_NU: src = hd; cnt = 2 * N; break;
        case S_DELTA: src = hp + 2 * N; cnt = 2 * N; break;
        case S_DELTAEXT:
            // delta_ext(delta) of the problem; zero without ENABLE_DELTA_EXT
            // (wslda_writevars.c:415-425, pca_kernels.cu:1120-1147)
            tmp.assign(2 * N, 0.0);
            if (c->proj && (c->proj->flags & TDW_PROJ_DELTA_EXT)) {
                TDW_TRY(c->proj->delta_ext(c->stream, it, (const double2 *)(d_pots + 2 * N), c->d_dext));
                TDW_HIP(hipMemcpyAsync(tmp.data(), c->d_dext, sizeof(double) * 2 * N, hipMemcpyDeviceToHost,
                                       c->stream));
                TDW_HIP(hipStreamSynchronize(c->stream));
            }
            src = tmp.data();
            cnt = 2 * N;
            break;
        case S_ALPHA_A: src = hp + 4 * N; break;
        case S_ALPHA_B: src = hp + 5 * N; break;
        case S_A_A: src = hp + 6 * N; cnt = 3 * N; break;
        case S_A_B: src = hp + 9 * N; cnt = 3 * N; break;
        case S_VEXT_A: src = hv; break;
        case S_VEXT_B: src = hv + N; break;
        case S_V_A:
        case S_V_B: {
            const double *V = hp + (v.src == S_V_A ? 0 : N), *E = hv + (v.src == S_V_A ? 0 : N);
            tmp.assign(V, V + N);
            if (!keep_v)
                for (size_t i = 0; i < N; i++) tmp[i] = V[i] - E[i];
            src = tmp.data();
            break;
        }
        case S_VELEXT_A:
        case S_VELEXT_B:
            // velocity_ext() of the problem (wslda_writevars.c:426-441)
            tmp.assign(3 * N, 0.0);
            if (c->proj && c->proj->velocity_ext) {
                TDW_TRY(c->proj->velocity_ext(c->stream, it, v.src == S_VELEXT_B ? 1 : 0, c->d_vel));
                TDW_HIP(hipMemcpyAsync(tmp.data(), c->d_vel, sizeof(double) * 3 * N, hipMemcpyDeviceToHost,
                                       c->stream));
                TDW_HIP(hipStreamSynchronize(c->stream));
            }
            src = tmp.data();
            cnt = 3 * N;
            break;
        }
        TDW_TRY(append(w, v, src, cnt));
    }
    w->cycles++;
    return write_wtxt(w);
}

extern "C" int tdw_wdata_cycles(const tdw_wdata *w) { return w ? w->cycles : TDW_EINVAL; }

extern "C" int tdw_wdata_destroy(tdw_wdata *w)
{
    delete w;
    return 0;
}

// ---------------------------------------------------------------------------
// Drop-in: the device halves of get_delta_ext / get_velocity_ext
// (pca_kernels.cu:1120-1180; 1D twins ccpca_kernels.cu:1087-1141) on the
// bound context.
// ---------------------------------------------------------------------------
// The reference evaluates delta_ext(delta) (zero without ENABLE_DELTA_EXT)
// and velocity_ext() of the problem; the built-in problem has neither, so both
// are zero there.
extern "C" int get_delta_ext(int datadim, int it, void *deltain, void *data)
{
    (void)datadim;
    tdw_ctx *c = tdw_bound();
    if (!c) return TDW_ENOCTX;
    if (!deltain || !data) return TDW_EINVAL;
    const size_t N = c->N;
    if (c->proj && (c->proj->flags & TDW_PROJ_DELTA_EXT)) {
        TDW_TRY(c->proj->delta_ext(c->stream, it, (const double2 *)deltain, c->d_dext));
        TDW_HIP(hipStreamSynchronize(c->stream));
        return (int)hipMemcpy(data, c->d_dext, sizeof(double2) * N, hipMemcpyDeviceToHost);
    }
    TDW_HIP(hipStreamSynchronize(c->stream));
    memset(data, 0, sizeof(double) * 2 * N);
    return 0;
}

extern "C" int get_velocity_ext(int datadim, int spin, int it, double *data)
{
    (void)datadim;
    tdw_ctx *c = tdw_bound();
    if (!c) return TDW_ENOCTX;
    if (!data) return TDW_EINVAL;
    const size_t N = c->N;
    if (c->proj && c->proj->velocity_ext) {
        TDW_TRY(c->proj->velocity_ext(c->stream, it, spin ? 1 : 0, c->d_vel));
        TDW_HIP(hipStreamSynchronize(c->stream));
        return (int)hipMemcpy(data, c->d_vel, sizeof(double) * 3 * N, hipMemcpyDeviceToHost);
    }
    memset(data, 0, sizeof(double) * 3 * N);
    return 0;
}
