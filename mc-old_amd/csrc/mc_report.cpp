// mc_report.cpp — output.h5 in the reference's layout (Simulator::report,
// report.cpp:9-50; Estimator::report, Estimator.cpp:368-422), written with the
// HDF5 C library loaded at run time, so the engine does not depend on HDF5:
//
//   /summary/{Ncycle, Nsample, Npassive, Ntrack} (unsigned long long),
//   /summary/mode (string), /summary/survival_roulette/{wr, ws} (double);
//   per estimator /<name>/ with attribute "indexing" = "[surface]" or
//   "[cell]" (then "[energy]"), one dataset per filter holding its grid with
//   attribute "unit" ("id#" or "eV"), and /<name>/<score>/{mean, uncertainty}
//   shaped by the filter sizes (attachments, then energy bins).
//
// The library is the first of $MC_HDF5_LIBRARY, libhdf5.so (the loader's
// search path) and /opt/conda/lib/libhdf5.so that loads.
#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mc_api.h"
#include "mc_problem.hpp"

namespace mc {

namespace {

typedef int64_t hid_t;             // H5Ipublic.h (HDF5 1.10)
typedef int herr_t;
typedef unsigned long long hsize_t;
constexpr hid_t kDefault = 0;      // H5P_DEFAULT, H5S_ALL
constexpr unsigned kAccTrunc = 0x0002u;
constexpr int kScalar = 0;         // H5S_SCALAR
constexpr size_t kVariable = (size_t)-1;

struct H5 {
    void* so = nullptr;
    herr_t (*open)();
    hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*gcreate)(hid_t, const char*, hid_t, hid_t, hid_t);
    herr_t (*gclose)(hid_t);
    hid_t (*screate)(int);
    hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
    herr_t (*sclose)(hid_t);
    hid_t (*dcreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    herr_t (*dclose)(hid_t);
    hid_t (*acreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t);
    herr_t (*awrite)(hid_t, hid_t, const void*);
    herr_t (*aclose)(hid_t);
    hid_t (*tcopy)(hid_t);
    herr_t (*tset_size)(hid_t, size_t);
    herr_t (*tclose)(hid_t);
    hid_t *t_ullong, *t_double, *t_c_s1;

    template <class F>
    bool sym(F& f, const char* name)
    {
        f = (F)dlsym(so, name);
        return f != nullptr;
    }
    bool load()
    {
        std::vector<std::string> names;
        if (const char* e = std::getenv("MC_HDF5_LIBRARY")) names.push_back(e);
        names.push_back("libhdf5.so");
        names.push_back("/opt/conda/lib/libhdf5.so");
        for (const auto& n : names)
            if ((so = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
        if (!so) return false;
        bool ok = sym(open, "H5open") && sym(fcreate, "H5Fcreate") && sym(fclose, "H5Fclose") &&
                  sym(gcreate, "H5Gcreate2") && sym(gclose, "H5Gclose") && sym(screate, "H5Screate") &&
                  sym(screate_simple, "H5Screate_simple") && sym(sclose, "H5Sclose") &&
                  sym(dcreate, "H5Dcreate2") && sym(dwrite, "H5Dwrite") && sym(dclose, "H5Dclose") &&
                  sym(acreate, "H5Acreate2") && sym(awrite, "H5Awrite") && sym(aclose, "H5Aclose") &&
                  sym(tcopy, "H5Tcopy") && sym(tset_size, "H5Tset_size") && sym(tclose, "H5Tclose") &&
                  sym(t_ullong, "H5T_NATIVE_ULLONG_g") && sym(t_double, "H5T_NATIVE_DOUBLE_g") &&
                  sym(t_c_s1, "H5T_C_S1_g");
        // the predefined type ids are set by the library's initialisation
        return ok && open() >= 0;
    }
};

H5* hdf5()
{
    static std::once_flag once;
    static H5 h;
    static bool ok = false;
    std::call_once(once, [] { ok = h.load(); });
    return ok ? &h : nullptr;
}

// One HDF5 file being written; every id it opens is closed by its end.
class Writer {
public:
    Writer(H5& h, const std::string& path) : L(h)
    {
        file = check(L.fcreate(path.c_str(), kAccTrunc, kDefault, kDefault), "create " + path);
        str = check(L.tcopy(*L.t_c_s1), "string type");
        check(L.tset_size(str, kVariable), "variable-length string");
        scalar = check(L.screate(kScalar), "scalar space");
    }
    ~Writer()
    {
        for (auto it = open_ids.rbegin(); it != open_ids.rend(); ++it) it->second(it->first);
        if (scalar >= 0) L.sclose(scalar);
        if (str >= 0) L.tclose(str);
        if (file >= 0) L.fclose(file);
    }
    hid_t group(hid_t at, const std::string& name)
    {
        const hid_t g = check(L.gcreate(at, name.c_str(), kDefault, kDefault, kDefault), "group " + name);
        open_ids.emplace_back(g, L.gclose);
        return g;
    }
    void scalar_ull(hid_t at, const char* name, unsigned long long v) { write(at, name, *L.t_ullong, scalar, &v); }
    void scalar_double(hid_t at, const char* name, double v) { write(at, name, *L.t_double, scalar, &v); }
    void scalar_string(hid_t at, const char* name, const std::string& v)
    {
        const char* c = v.c_str();
        write(at, name, str, scalar, &c);
    }
    hid_t array(hid_t at, const std::string& name, const std::vector<hsize_t>& dims, const double* v)
    {
        const hid_t sp = check(L.screate_simple((int)dims.size(), dims.data(), nullptr), "space " + name);
        open_ids.emplace_back(sp, L.sclose);
        return write(at, name.c_str(), *L.t_double, sp, v);
    }
    void attribute(hid_t at, const char* name, const std::string& v)
    {
        const hid_t a = check(L.acreate(at, name, str, scalar, kDefault, kDefault), std::string("attribute ") + name);
        open_ids.emplace_back(a, L.aclose);
        const char* c = v.c_str();
        check(L.awrite(a, str, &c), std::string("write attribute ") + name);
    }
    hid_t root() const { return file; }

private:
    hid_t write(hid_t at, const char* name, hid_t type, hid_t space, const void* v)
    {
        const hid_t d = check(L.dcreate(at, name, type, space, kDefault, kDefault, kDefault), std::string("dataset ") + name);
        open_ids.emplace_back(d, L.dclose);
        check(L.dwrite(d, type, kDefault, kDefault, kDefault, v), std::string("write ") + name);
        return d;
    }
    template <class T>
    T check(T id, const std::string& what)
    {
        if (id < 0) throw Error(MC_ERR_IO, "HDF5: cannot " + what);
        return id;
    }
    H5& L;
    hid_t file = -1, str = -1, scalar = -1;
    std::vector<std::pair<hid_t, herr_t (*)(hid_t)>> open_ids;
};

}  // namespace

bool hdf5_available() { return hdf5() != nullptr; }

void report_h5(const Problem& P, const mc_tally_out& t, uint64_t n_histories, const std::string& path)
{
    H5* h = hdf5();
    if (!h) throw Error(MC_ERR_UNSUPPORTED, "no HDF5 library could be loaded (set MC_HDF5_LIBRARY)");
    Writer w(*h, path);
    // report.cpp:23-39
    const hid_t summary = w.group(w.root(), "/summary");
    w.scalar_ull(summary, "Ncycle", 1ull);
    w.scalar_ull(summary, "Nsample", (unsigned long long)n_histories);
    w.scalar_ull(summary, "Npassive", 0ull);
    w.scalar_ull(summary, "Ntrack", (unsigned long long)t.flights);
    w.scalar_string(summary, "mode", "fixed source");
    const hid_t roulette = w.group(w.root(), "/summary/survival_roulette");
    w.scalar_double(roulette, "wr", P.wr);
    w.scalar_double(roulette, "ws", P.ws);
    // Estimator::report, Estimator.cpp:368-422
    for (size_t e = 0; e < P.estimators.size(); ++e) {
        const mc_estimator_desc& E = P.estimators[e];
        const hid_t g = w.group(w.root(), "/" + P.estimator_names[e]);
        const bool surf = E.attach_kind == MC_ATTACH_SURFACE;
        std::vector<hsize_t> dims{(hsize_t)E.n_attach};
        std::string indexing = surf ? "[surface]" : "[cell]";
        const std::vector<double>& ids = P.estimator_attach_ids[e];
        const hid_t f0 = w.array(g, surf ? "surface" : "cell", {(hsize_t)ids.size()}, ids.data());
        w.attribute(f0, "unit", "id#");
        if (E.egrid_count) {
            indexing += "[energy]";
            const std::vector<double>& eg = P.estimator_egrid[e];
            const hid_t f1 = w.array(g, "energy", {(hsize_t)eg.size()}, eg.data());
            w.attribute(f1, "unit", "eV");
            dims.push_back((hsize_t)(eg.size() - 1));
        }
        w.attribute(g, "indexing", indexing);
        const int per = E.n_bins / E.n_scores;
        for (int s = 0; s < E.n_scores; ++s) {
            const hid_t sg = w.group(g, P.estimator_scores[e][s]);
            w.array(sg, "mean", dims, t.mean + E.bin_offset + s * per);
            w.array(sg, "uncertainty", dims, t.uncer + E.bin_offset + s * per);
        }
    }
}

}  // namespace mc
