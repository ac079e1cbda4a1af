// Restart files in the reference's binary format (cSaturnModel::save_state / load_state,
// planet/FileIO_Sat.cpp:513-660): five int32 (magic, im, jm, km, iteration), then whole arrays,
// each as im*jm rows of km doubles, which for the dense [i][j][k] host arrays is one contiguous
// block per array. The array lists are restart_arrays() (:513-527) and restart_arrays_carried()
// (:538-554, S_precip_ch4 appended in SAT3).
#include "restart.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>

namespace atsat {

namespace {

constexpr int32_t kMagic1 = 0x53415431, kMagic2 = 0x53415432, kMagic3 = 0x53415433;   // "SAT1/2/3"

RestartItem F(int f){ return { f, -1 }; }
RestartItem N(int f){ return { ATSAT_F_N0 + f, -1 }; }
RestartItem X(int x){ return { -1, x }; }

struct FileCloser { void operator()(std::FILE* f) const { if(f) std::fclose(f); } };
using File = std::unique_ptr<std::FILE, FileCloser>;

size_t cells(const atsat_grid& g){ return (size_t)g.im * g.jm * g.km; }

}  // namespace

std::vector<RestartItem> restart_items(int format){
    // restart_arrays(), FileIO_Sat.cpp:513-527
    std::vector<RestartItem> v = {
        F(ATSAT_T), F(ATSAT_U), F(ATSAT_V), F(ATSAT_W),
        N(ATSAT_T), N(ATSAT_U), N(ATSAT_V), N(ATSAT_W),
        F(ATSAT_H2O), F(ATSAT_H2O_CLOUD), F(ATSAT_H2O_ICE),
        N(ATSAT_H2O), N(ATSAT_H2O_CLOUD), N(ATSAT_H2O_ICE),
        F(ATSAT_H2S), N(ATSAT_H2S),
        F(ATSAT_NH3), F(ATSAT_NH3_CLOUD), F(ATSAT_NH3_ICE),
        N(ATSAT_NH3), N(ATSAT_NH3_CLOUD), N(ATSAT_NH3_ICE),
        F(ATSAT_CH4), F(ATSAT_CH4_CLOUD), F(ATSAT_CH4_ICE),
        N(ATSAT_CH4), N(ATSAT_CH4_CLOUD), N(ATSAT_CH4_ICE),
        F(ATSAT_NH4SH), N(ATSAT_NH4SH),
        F(ATSAT_P_DYN), X(ATSAT_RX_P_DYNN), F(ATSAT_P_STAT),
        F(ATSAT_TKE), F(ATSAT_DIS), N(ATSAT_TKE), N(ATSAT_DIS), F(ATSAT_NUE) };
    if(format >= 2){
        // restart_arrays_carried(), FileIO_Sat.cpp:538-554
        const RestartItem c[] = {
            F(ATSAT_TKE_SOURCE), F(ATSAT_DIS_SOURCE),
            F(ATSAT_DIFFLUX_H2S), F(ATSAT_DIFFLUX_NH3), F(ATSAT_DIFFLUX_NH4SH),
            X(ATSAT_RX_J_H2S), X(ATSAT_RX_J_NH3), X(ATSAT_RX_J_NH4SH),
            F(ATSAT_MASSFLUX_H2S), F(ATSAT_MASSFLUX_NH3), F(ATSAT_MASSFLUX_NH4SH),
            F(ATSAT_THERMALMASSFLUX), F(ATSAT_FLUXLIM_NH4SH),
            F(ATSAT_RHS_U), F(ATSAT_RHS_V),
            F(ATSAT_RHO_MIX), F(ATSAT_Q_PRECIP),
            F(ATSAT_S_PRECIP_H2O), F(ATSAT_S_PRECIP_H2O_CLOUD), F(ATSAT_S_PRECIP_H2O_ICE),
            F(ATSAT_S_PRECIP_NH3), F(ATSAT_S_PRECIP_NH3_ICE),
            F(ATSAT_S_PRECIP_NH3_CLOUD), F(ATSAT_S_PRECIP_CH4_CLOUD), F(ATSAT_S_PRECIP_CH4_ICE),
            X(ATSAT_RX_W_H2S), X(ATSAT_RX_W_NH3), X(ATSAT_RX_W_NH4SH),
            X(ATSAT_RX_BUOYANCY_FORCE), X(ATSAT_RX_PRES_GRAD_FORCE), X(ATSAT_RX_Q_SENSIBLE),
            X(ATSAT_RX_JT_H2S), X(ATSAT_RX_JT_NH3), X(ATSAT_RX_JT_NH4SH) };
        v.insert(v.end(), std::begin(c), std::end(c));
        if(format >= 3) v.push_back(F(ATSAT_S_PRECIP_CH4));
    }
    return v;
}

int restart_write(const char* path, const atsat_grid& g, int iter, int format,
                  const std::function<int(const RestartItem&, double*, const double**)>& get, std::string& err){
    if(format != 1 && format != 3){ err = "restart format must be 1 or 3"; return ATSAT_E_ARG; }
    File f(std::fopen(path, "wb"));
    if(!f){ err = std::string("cannot open ") + path + " for writing"; return ATSAT_E_IO; }
    const size_t n = cells(g);
    std::vector<double> scratch(n);
    const int32_t hdr[5] = { format == 3 ? kMagic3 : kMagic1, g.im, g.jm, g.km, iter };
    if(std::fwrite(hdr, sizeof(hdr), 1, f.get()) != 1){ err = std::string("write failed: ") + path; return ATSAT_E_IO; }
    for(const RestartItem& it : restart_items(format)){
        const double* a = nullptr;
        const int rc = get(it, scratch.data(), &a);
        if(rc) return rc;
        if(std::fwrite(a, sizeof(double), n, f.get()) != n){ err = std::string("write failed (disk full?): ") + path; return ATSAT_E_IO; }
    }
    if(std::fflush(f.get()) != 0){ err = std::string("write failed: ") + path; return ATSAT_E_IO; }
    return ATSAT_OK;
}

int restart_read(const char* path, const atsat_grid& g, int* iter, int* format,
                 const std::function<double*(const RestartItem&, double*)>& dest,
                 const std::function<int(const RestartItem&, const double*)>& after, std::string& err){
    File f(std::fopen(path, "rb"));
    if(!f){ err = std::string("no file ") + path; return ATSAT_E_IO; }
    int32_t hdr[5];
    if(std::fread(hdr, sizeof(hdr), 1, f.get()) != 1){ err = std::string("bad header in ") + path; return ATSAT_E_IO; }
    const int fmt = hdr[0] == kMagic3 ? 3 : hdr[0] == kMagic2 ? 2 : hdr[0] == kMagic1 ? 1 : 0;
    if(!fmt || hdr[1] != g.im || hdr[2] != g.jm || hdr[3] != g.km){
        err = std::string("bad header / grid mismatch in ") + path;
        return ATSAT_E_IO;
    }
    const size_t n = cells(g);
    std::vector<double> scratch(n);
    for(const RestartItem& it : restart_items(fmt)){
        double* d = dest(it, scratch.data());
        if(!d){
            if(std::fseek(f.get(), (long)(n * sizeof(double)), SEEK_CUR) != 0){ err = std::string("truncated file ") + path; return ATSAT_E_IO; }
            continue;
        }
        if(std::fread(d, sizeof(double), n, f.get()) != n){ err = std::string("truncated file ") + path; return ATSAT_E_IO; }
        const int rc = after(it, d);
        if(rc) return rc;
    }
    // fseek past the end does not fail: a truncated last array that was skipped shows here
    const long pos = std::ftell(f.get());
    if(std::fseek(f.get(), 0, SEEK_END) != 0 || std::ftell(f.get()) < pos){ err = std::string("truncated file ") + path; return ATSAT_E_IO; }
    if(iter) *iter = hdr[4];
    if(format) *format = fmt;
    return ATSAT_OK;
}

}  // namespace atsat

using namespace atsat;

extern "C" int atsat_restart_write_host(const char* path, const atsat_grid* g, int iter, const atsat_host_state* s,
                                        const atsat_restart_extra* x, int format){
    std::string err;
    auto done = [&](int rc){ if(rc) set_last_error(err); return rc; };
    if(!path || !g || !s){ err = "null argument"; return done(ATSAT_E_ARG); }
    for(const RestartItem& it : restart_items(format == 3 ? 3 : 1))
        if(it.field >= 0 && !s->f[it.field]){ err = "a field the restart format lists is missing from the host state"; return done(ATSAT_E_ARG); }
    bool zeroed = false;
    return done(restart_write(path, *g, iter, format, [&](const RestartItem& it, double* scratch, const double** a) -> int {
        if(it.field >= 0) *a = s->f[it.field];
        else if(x && x->a[it.extra]) *a = x->a[it.extra];
        else if(it.extra == ATSAT_RX_P_DYNN) *a = s->f[ATSAT_P_DYN];        // restoreVar: p_dynn = p_dyn
        else {
            if(!zeroed){ std::memset(scratch, 0, sizeof(double) * (size_t)g->im * g->jm * g->km); zeroed = true; }
            *a = scratch;
        }
        return 0;
    }, err));
}

extern "C" int atsat_restart_read_host(const char* path, const atsat_grid* g, atsat_host_state* s,
                                       atsat_restart_extra* x, int* iter, int* format){
    std::string err;
    auto done = [&](int rc){ if(rc) set_last_error(err); return rc; };
    if(!path || !g || !s){ err = "null argument"; return done(ATSAT_E_ARG); }
    return done(restart_read(path, *g, iter, format,
        [&](const RestartItem& it, double*) -> double* {
            if(it.field >= 0) return s->f[it.field];
            return x ? x->a[it.extra] : nullptr;
        },
        [](const RestartItem&, const double*) { return 0; }, err));
}
