// mc_problem.cpp — deck loader in the reference's XML schema.
//
// Follows Simulator::Simulator (src/simulator/setup.cpp:30-1069) element by
// element.  Where the reference exits or crashes, this loader raises an
// mc::Error with a status code instead.  Extensions (all optional):
//   <nuclide><scatter xs=".."/></nuclide>      constant scattering XS (one-speed)
//   <description ... seed=".."/>               LCG seed (default 1, Random.cpp:105)
//   <parameters><parameter cell=".." nuclide=".." reaction="scatter|capture|total"/>
//                                              gradient parameters (theta scales)
#include "mc_problem.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>

#include "mc_physics.hpp"
#include "mc_xml.hpp"

namespace mc {

int ref_binary_search(double x, const double* v, int n) { return binary_search(x, v, n); }

int Problem::search_cell(const double p[3]) const
{
    const Vec3 q = {p[0], p[1], p[2]};
    for (size_t c = 0; c < cells.size(); ++c) {
        bool inside = true;
        for (int k = 0; k < cells[c].surf_count; ++k) {
            const int j = cells[c].surf_begin + k;
            const mc_surface_desc& S = surfaces[cell_surf[j]];
            if (surface_eval(S.type, S.c, q) * (double)cell_sense[j] < 0) { inside = false; break; }
        }
        if (inside) return (int)c;
    }
    return -1;
}

void Problem::finalize()
{
    desc = mc_problem_desc{};
    desc.n_surfaces = (int32_t)surfaces.size();
    desc.n_cells = (int32_t)cells.size();
    desc.n_materials = (int32_t)materials.size();
    desc.n_nuclides = (int32_t)nuclides.size();
    desc.n_sources = (int32_t)sources.size();
    desc.n_estimators = (int32_t)estimators.size();
    desc.n_params = (int32_t)params.size();
    desc.n_bins = n_bins;
    desc.surfaces = surfaces.data();
    desc.cells = cells.data();
    desc.cell_surf = cell_surf.data();
    desc.cell_sense = cell_sense.data();
    desc.materials = materials.data();
    desc.mat_nuclide = mat_nuclide.data();
    desc.mat_density = mat_density.data();
    desc.nuclides = nuclides.data();
    desc.xs_energy = xs_energy.data();
    desc.xs_scatter = xs_scatter.data();
    desc.xs_capture = xs_capture.data();
    desc.xs_total = xs_total.data();
    desc.sources = sources.data();
    desc.source_cdf = source_cdf.data();
    desc.estimators = estimators.data();
    desc.att_estimator = att_estimator.data();
    desc.att_index = att_index.data();
    desc.egrid = egrid.data();
    desc.params = params.data();
    desc.mat_param = mat_param.data();
    desc.wr = wr;
    desc.ws = ws;
    desc.n_histories = n_histories;
    desc.seed = seed;
    desc.comb_bank_max = comb_bank_max;
    desc.comb_teeth = comb_teeth;
}

static void need_finite(const std::vector<double>& E, const std::string& path)
{
    if (E.empty()) throw Error(MC_ERR_INPUT, "no cross-section rows in " + path);
    for (double e : E)
        if (!std::isfinite(e)) throw Error(MC_ERR_INPUT, "non-finite energy in " + path);
}

void load_xs_file(const std::string& path, double& A, std::vector<double>& E,
                  std::vector<double>& s, std::vector<double>& c, std::vector<double>& f)
{
    std::ifstream in(path);
    if (!in) throw Error(MC_ERR_IO, "cannot read cross-section file " + path);
    std::string line;
    std::vector<std::vector<double>> rows;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::vector<double> row;
        std::string tok;
        while (ls >> tok) {
            char* end = nullptr;
            const double v = std::strtod(tok.c_str(), &end);
            if (end == tok.c_str()) throw Error(MC_ERR_INPUT, "non-numeric token '" + tok + "' in " + path);
            row.push_back(v);
        }
        if (!row.empty()) rows.push_back(row);
    }
    // line 1: A; lines 2-4: Watt parameters (xs_library/README.txt:1-7)
    if (rows.size() < 5 || rows[0].size() != 1) throw Error(MC_ERR_INPUT, "bad header in " + path);
    A = rows[0][0];
    for (int k = 1; k <= 3; ++k)
        if (rows[k].size() != 2) throw Error(MC_ERR_INPUT, "bad Watt parameters in " + path);
    E.clear(); s.clear(); c.clear(); f.clear();
    for (size_t k = 4; k < rows.size(); ++k) {
        const auto& r = rows[k];
        if (r.size() != 5 && r.size() != 6)
            throw Error(MC_ERR_INPUT, "row " + std::to_string(k + 1) + " of " + path + " has " +
                                          std::to_string(r.size()) + " columns (expected 5 or 6)");
        E.push_back(r[0]);
        s.push_back(r[1]);
        c.push_back(r[2]);
        f.push_back(r[3]);
    }
    // Energies are taken as they stand, as the reference's reader does
    // (setup.cpp:346-376): its tables repeat an energy where a cross section
    // jumps (xs_library/008016.txt, 006000.txt, 002003.txt), and
    // binary_search's "largest i with E_i < x" (Algorithm.cpp:46-64) never
    // picks such a zero-width interval; a file whose energies decrease
    // somewhere (xs_library/005011.txt, by rounding) is searched exactly as
    // the reference searches it.
    need_finite(E, path);
}

namespace {

template <class T>
int find_name(const std::vector<std::string>& names, const T& n)
{
    for (size_t i = 0; i < names.size(); ++i)
        if (names[i] == n) return (int)i;
    return -1;
}

struct DistDouble {
    int kind;        // MC_ENERGY_DELTA / MC_ENERGY_UNIFORM
    double a, b;
};
struct DistPoint {
    int kind;        // MC_DIR_DELTA / MC_DIR_ISOTROPIC
    double x, y, z;
};

void need(bool ok, mc_status st, const std::string& msg)
{
    if (!ok) throw Error(st, msg);
}

}  // namespace

void load_problem_xml(Problem& P, const std::string& xml, const std::string& xs_dir)
{
    std::vector<XmlNode> roots;
    std::string err;
    if (!xml_parse(xml, roots, err)) throw Error(MC_ERR_INPUT, "XML: " + err);
    auto top = [&](const char* n) -> const XmlNode* {
        for (const auto& r : roots)
            if (r.name == n) return &r;
        return nullptr;
    };

    // ---------------------------------------------------------- simulation
    const XmlNode* sim = top("simulation");
    need(sim && sim->child("description"), MC_ERR_INPUT, "missing <simulation><description>");
    const XmlNode* descr = sim->child("description");
    P.name = descr->value("name");
    P.n_histories = (uint64_t)descr->as_double("samples");       // setup.cpp:50
    if (descr->has("seed")) P.seed = (uint64_t)std::strtoull(descr->value("seed").c_str(), nullptr, 10);
    need(P.seed > 0 && (P.seed & 1) == 1, MC_ERR_INPUT, "seed must be a positive odd integer");
    need(!sim->child("ksearch"), MC_ERR_UNSUPPORTED, "k-eigenvalue search is out of scope (non-multiplying media only)");
    need(!sim->child("tdmc"), MC_ERR_UNSUPPORTED, "time-dependent MC is out of scope");
    need(!sim->child("trmm"), MC_ERR_UNSUPPORTED, "TRMM is out of scope");
    if (const XmlNode* pc = top("population_control")) {                 // setup.cpp:57-67
        if (const XmlNode* comb = pc->child("particle_comb")) {
            P.comb_bank_max = comb->as_int("bank_max");
            P.comb_teeth = comb->as_int("teeth");
            need(P.comb_bank_max >= 1 && P.comb_teeth >= 1, MC_ERR_INPUT,
                 "<particle_comb> needs bank_max >= 1 and teeth >= 1");
        }
    }

    // ------------------------------------------------------- distributions
    std::map<std::string, DistDouble> ddist;
    std::map<std::string, DistPoint> pdist;
    if (const XmlNode* dists = top("distributions")) {              // setup.cpp:174-305
        for (const auto& d : dists->children) {
            const std::string name = d.value("name");
            const std::string data = d.value("datatype");
            if (data == "double") {
                if (d.name == "delta") ddist[name] = {MC_ENERGY_DELTA, d.as_double("val"), 0.0};
                else if (d.name == "uniform") ddist[name] = {MC_ENERGY_UNIFORM, d.as_double("a"), d.as_double("b")};
                else if (d.name == "watt") throw Error(MC_ERR_UNSUPPORTED, "Watt fission spectrum is out of scope");
                else throw Error(MC_ERR_INPUT, "unsupported double distribution " + d.name);
            } else if (data == "point") {
                if (d.name == "delta") pdist[name] = {MC_DIR_DELTA, d.as_double("x"), d.as_double("y"), d.as_double("z")};
                else if (d.name == "isotropic") pdist[name] = {MC_DIR_ISOTROPIC, 0, 0, 0};
                else if (d.name == "independentXYZ") throw Error(MC_ERR_UNSUPPORTED, "independentXYZ direction is not supported");
                else throw Error(MC_ERR_INPUT, "unsupported point distribution " + d.name);
            } else {
                throw Error(MC_ERR_INPUT, "unsupported distribution datatype '" + data + "'");
            }
        }
    }

    // ------------------------------------------------------------ nuclides
    if (const XmlNode* nucs = top("nuclides")) {                     // setup.cpp:310-467
        for (const XmlNode* n : nucs->children_named("nuclide")) {
            mc_nuclide_desc N{};
            const std::string name = n->value("name");
            if (n->has("ZAID")) {
                const std::string path = xs_dir + "/" + n->value("ZAID") + ".txt";
                std::vector<double> E, s, c, f;
                load_xs_file(path, N.A, E, s, c, f);
                for (double v : f)
                    need(v == 0.0, MC_ERR_UNSUPPORTED, "nuclide " + name + " is fissile (non-multiplying media only)");
                N.kind = MC_NUC_TABLE;
                N.grid_begin = (int32_t)P.xs_energy.size();
                N.grid_count = (int32_t)E.size();
                for (size_t k = 0; k < E.size(); ++k) {
                    P.xs_energy.push_back(E[k]);
                    P.xs_scatter.push_back(s[k]);
                    P.xs_capture.push_back(c[k]);
                    P.xs_total.push_back(s[k] + c[k] + f[k]);       // setup.cpp:372
                }
            } else {
                N.A = n->has("A") ? n->as_double("A") : MC_MASS_INFINITE;  // n_A = MAX_float, setup.cpp:313
                if (N.A >= 3.0e38) N.A = MC_MASS_INFINITE;
                N.kind = MC_NUC_CONSTANT;
                for (const auto& r : n->children) {
                    need(r.has("xs"), MC_ERR_INPUT, "unknown XS type in nuclide " + name);
                    if (r.name == "capture") N.xs_capture = r.as_double("xs");
                    else if (r.name == "scatter") N.xs_scatter = r.as_double("xs");
                    else throw Error(MC_ERR_INPUT, "user-defined nuclide " + name + " supports capture and scatter only");
                }
                N.xs_total = N.xs_scatter + N.xs_capture + 0.0;
            }
            need(N.A > 0.0, MC_ERR_INPUT, "nuclide " + name + " has non-positive mass");
            // the reference scatters every nuclide off a free gas at 293.6 K
            // (Reaction.cpp:27-118); a deck may ask for the target-at-rest
            // limit (T -> 0) instead
            N.kinematics = MC_KIN_FREE_GAS;
            if (n->has("kinematics")) {
                const std::string k = n->value("kinematics");
                need(k == "free_gas" || k == "target_at_rest", MC_ERR_INPUT,
                     "unknown kinematics '" + k + "' of nuclide " + name);
                N.kinematics = k == "free_gas" ? MC_KIN_FREE_GAS : MC_KIN_TARGET_AT_REST;
            }
            P.nuclides.push_back(N);
            P.nuclide_names.push_back(name);
        }
    }

    // ----------------------------------------------------------- materials
    if (const XmlNode* mats = top("materials")) {                    // setup.cpp:474-492
        for (const XmlNode* m : mats->children_named("material")) {
            mc_material_desc M{};
            M.nuc_begin = (int32_t)P.mat_nuclide.size();
            for (const XmlNode* n : m->children_named("nuclide")) {
                const int k = find_name(P.nuclide_names, n->value("name"));
                need(k >= 0, MC_ERR_INPUT, "unknown nuclide '" + n->value("name") + "' in material " + m->value("name"));
                P.mat_nuclide.push_back(k);
                P.mat_density.push_back(n->as_double("density"));
            }
            M.nuc_count = (int32_t)P.mat_nuclide.size() - M.nuc_begin;
            P.materials.push_back(M);
            P.material_names.push_back(m->value("name"));
        }
    }

    // ------------------------------------------------------------ surfaces
    if (const XmlNode* surfs = top("surfaces")) {                    // setup.cpp:499-584
        for (const auto& s : surfs->children) {
            mc_surface_desc S{};
            std::string bc = "transmission";
            if (s.has("bc")) bc = s.value("bc");
            if (bc == "transmission") S.bc = MC_BC_TRANSMISSION;
            else if (bc == "reflective") S.bc = MC_BC_REFLECTIVE;
            else if (bc == "vacuum") S.bc = MC_BC_VACUUM;
            else throw Error(MC_ERR_INPUT, "unknown boundary condition '" + bc + "'");
            if (s.name == "plane_x") { S.type = MC_SURF_PLANE_X; S.c[0] = s.as_double("x"); }
            else if (s.name == "plane_y") { S.type = MC_SURF_PLANE_Y; S.c[0] = s.as_double("y"); }
            else if (s.name == "plane_z") { S.type = MC_SURF_PLANE_Z; S.c[0] = s.as_double("z"); }
            else if (s.name == "plane") {
                S.type = MC_SURF_PLANE;
                const double a = s.as_double("a"), b = s.as_double("b"), c = s.as_double("c");
                S.c[0] = a; S.c[1] = b; S.c[2] = c; S.c[3] = s.as_double("d");
                const double L = 2.0 / (a * a + b * b + c * c);       // Geometry.cpp:19-22
                S.c[4] = L * a; S.c[5] = L * b; S.c[6] = L * c;
            } else if (s.name == "sphere") {
                S.type = MC_SURF_SPHERE;
                S.c[0] = s.as_double("x"); S.c[1] = s.as_double("y"); S.c[2] = s.as_double("z");
                S.c[3] = s.as_double("r"); S.c[4] = S.c[3] * S.c[3];
            } else if (s.name == "cylinder_x") {
                S.type = MC_SURF_CYL_X;
                S.c[0] = s.as_double("y"); S.c[1] = s.as_double("z");
                S.c[2] = s.as_double("r"); S.c[4] = S.c[2] * S.c[2];
            } else if (s.name == "cylinder_z") {
                S.type = MC_SURF_CYL_Z;
                S.c[0] = s.as_double("x"); S.c[1] = s.as_double("y");
                S.c[2] = s.as_double("r"); S.c[4] = S.c[2] * S.c[2];
            } else {
                throw Error(MC_ERR_INPUT, "unknown surface type " + s.name);
            }
            P.surfaces.push_back(S);
            P.surface_names.push_back(s.value("name"));
        }
    }

    // --------------------------------------------------------------- cells
    if (const XmlNode* cells = top("cells")) {                       // setup.cpp:591-629
        for (const XmlNode* c : cells->children_named("cell")) {
            mc_cell_desc C{};
            std::string name = "Cell " + std::to_string(P.cells.size() + 1);
            if (c->has("name")) name = c->value("name");
            C.importance = c->has("importance") ? c->as_double("importance") : 1.0;
            C.material = -1;
            if (c->has("material")) {
                C.material = find_name(P.material_names, c->value("material"));
                need(C.material >= 0, MC_ERR_INPUT, "unknown material in cell " + name);
            }
            C.surf_begin = (int32_t)P.cell_surf.size();
            for (const XmlNode* s : c->children_named("surface")) {
                const int k = find_name(P.surface_names, s->value("name"));
                need(k >= 0, MC_ERR_INPUT, "unknown surface '" + s->value("name") + "' in cell " + name);
                P.cell_surf.push_back(k);
                P.cell_sense.push_back(s->as_int("sense"));
            }
            C.surf_count = (int32_t)P.cell_surf.size() - C.surf_begin;
            need(C.importance >= 0.0, MC_ERR_INPUT, "negative importance in cell " + name);
            P.cells.push_back(C);
            P.cell_names.push_back(name);
        }
    }

    // ---------------------------------------------------------- estimators
    // Attachments are collected per geometry in deck order (Surface::estimators,
    // Cell::estimators_TL / estimators_C), then laid out in CSR form.
    std::vector<std::vector<std::pair<int, int>>> surf_att(P.surfaces.size()), tl_att(P.cells.size()),
        col_att(P.cells.size());
    if (const XmlNode* ests = top("estimators")) {                  // setup.cpp:637-805
        for (const XmlNode* e : ests->children_named("estimator")) {
            mc_estimator_desc E{};
            const int eid = (int)P.estimators.size();
            std::string name = "Estimator #" + std::to_string(eid + 1);
            if (e->has("name")) name = e->value("name");
            std::string type = "TL";
            if (e->has("type")) type = e->value("type");
            int kern;
            if (type == "TL") kern = MC_KERNEL_TRACK_LENGTH;
            else if (type == "C") kern = MC_KERNEL_COLLISION;
            else throw Error(MC_ERR_INPUT, "unsupported score type in estimator " + name);
            need(e->has("scores"), MC_ERR_INPUT, "there is no score in estimator " + name);
            std::istringstream iss(e->value("scores"));
            std::vector<std::string> snames;
            for (std::string s; iss >> s;) {
                need(E.n_scores < MC_MAX_SCORES, MC_ERR_CAPACITY, "too many scores in estimator " + name);
                int sc;
                if (s == "flux") sc = MC_SCORE_FLUX;
                else if (s == "absorption") sc = MC_SCORE_ABSORPTION;
                else if (s == "scatter") sc = MC_SCORE_SCATTER;
                else if (s == "capture") sc = MC_SCORE_CAPTURE;
                else if (s == "total") sc = MC_SCORE_TOTAL;
                else if (s == "cross") { kern = MC_KERNEL_NEUTRON; sc = MC_SCORE_FLUX; }   // setup.cpp:688-690
                else if (s == "fission" || s == "nu-fission")
                    throw Error(MC_ERR_UNSUPPORTED, "fission scores are out of scope (non-multiplying media)");
                else throw Error(MC_ERR_INPUT, "unsupported score type " + s + " in estimator " + name);
                E.score[E.n_scores] = sc;
                E.score_kernel[E.n_scores] = kern;
                E.n_scores++;
                snames.push_back(s);
            }
            need(E.n_scores > 0, MC_ERR_INPUT, "there is no score in estimator " + name);
            need(!e->child("tdmc"), MC_ERR_UNSUPPORTED, "TDMC filters are out of scope");
            std::vector<double> grid;
            const auto surfs = e->children_named("surface");
            const auto cells = e->children_named("cell");
            need(!(surfs.size() && cells.size()), MC_ERR_UNSUPPORTED,
                 "estimator " + name + " attached to both surfaces and cells");
            std::vector<std::pair<int, bool>> attached;   // (geometry, is_surface)
            for (const XmlNode* s : surfs) {
                const int k = find_name(P.surface_names, s->value("name"));
                need(k >= 0, MC_ERR_INPUT, "unknown surface label " + s->value("name") + " in estimator " + name);
                attached.emplace_back(k, true);
                grid.push_back((double)k);
            }
            for (const XmlNode* c : cells) {
                const int k = find_name(P.cell_names, c->value("name"));
                need(k >= 0, MC_ERR_INPUT, "unknown cell label " + c->value("name") + " in estimator " + name);
                attached.emplace_back(k, false);
                grid.push_back((double)k);
            }
            need(!attached.empty(), MC_ERR_INPUT, "estimator " + name + " needs to be attached somewhere");
            E.attach_kind = surfs.empty() ? MC_ATTACH_CELL : MC_ATTACH_SURFACE;
            E.n_attach = (int32_t)grid.size();
            for (const auto& a : attached) {
                // FilterSurface/FilterCell::idx_l: binary_search(ID, grid) + 1 (Estimator.cpp:140,156)
                const int fi = binary_search((double)a.first, grid.data(), (int)grid.size()) + 1;
                need(fi >= 0 && fi < (int)grid.size(), MC_ERR_INPUT,
                     "attachment index out of range in estimator " + name);
                if (a.second) surf_att[a.first].emplace_back(eid, fi);
                else if (kern == MC_KERNEL_COLLISION || type == "C") col_att[a.first].emplace_back(eid, fi);
                else tl_att[a.first].emplace_back(eid, fi);
            }
            std::vector<double> eg;
            for (const XmlNode* f : e->children_named("filter")) {
                const std::string ft = f->value("type");
                std::vector<double> fg;
                if (f->has("grid")) {
                    std::istringstream g(f->value("grid"));
                    for (double v; g >> v;) fg.push_back(v);
                } else if (f->has("grid_linear")) {                    // setup.cpp:753-763
                    std::istringstream g(f->value("grid_linear"));
                    double a, step, b;
                    g >> a >> step >> b;
                    need(step > 0, MC_ERR_INPUT, "grid_linear step must be positive");
                    fg.push_back(a);
                    while (fg.back() < b) fg.push_back(fg.back() + step);
                    fg.pop_back();
                    fg.push_back(b);
                } else if (f->has("grid_lethargy")) {                  // setup.cpp:764-776
                    std::istringstream g(f->value("grid_lethargy"));
                    double a, b, N;
                    g >> a >> b >> N;
                    need(N >= 1 && a > 0 && b > a, MC_ERR_INPUT, "bad grid_lethargy in estimator " + name);
                    const double step = std::log(b / a) / N;
                    fg.push_back(0.0);
                    while ((double)fg.size() != N + 1) fg.push_back(fg.back() + step);
                    std::reverse(fg.begin(), fg.end());
                    for (auto& v : fg) v = b * std::exp(-v);
                } else {
                    throw Error(MC_ERR_INPUT, "need filter grid for estimator " + name);
                }
                if (ft == "energy") {
                    need(eg.empty(), MC_ERR_UNSUPPORTED, "more than one energy filter in estimator " + name);
                    need(fg.size() >= 2, MC_ERR_INPUT, "energy filter needs at least two edges");
                    eg = fg;
                } else if (ft == "time") {
                    throw Error(MC_ERR_UNSUPPORTED, "time filters are out of scope (no time-dependent tallies)");
                } else {
                    throw Error(MC_ERR_INPUT, "unknown filter type for estimator " + name);
                }
            }
            E.egrid_begin = (int32_t)P.egrid.size();
            E.egrid_count = (int32_t)eg.size();
            P.egrid.insert(P.egrid.end(), eg.begin(), eg.end());
            const int NE = eg.empty() ? 1 : (int)eg.size() - 1;
            E.bin_offset = P.n_bins;
            E.n_bins = E.n_scores * E.n_attach * NE;
            P.n_bins += E.n_bins;
            P.estimators.push_back(E);
            P.estimator_names.push_back(name);
            P.estimator_scores.push_back(snames);
            P.estimator_attach_ids.push_back(grid);
            P.estimator_egrid.push_back(eg);
        }
    }
    for (size_t s = 0; s < P.surfaces.size(); ++s) {
        P.surfaces[s].est_begin = (int32_t)P.att_estimator.size();
        for (const auto& a : surf_att[s]) { P.att_estimator.push_back(a.first); P.att_index.push_back(a.second); }
        P.surfaces[s].est_count = (int32_t)P.att_estimator.size() - P.surfaces[s].est_begin;
    }
    for (size_t c = 0; c < P.cells.size(); ++c) {
        P.cells[c].tl_begin = (int32_t)P.att_estimator.size();
        for (const auto& a : tl_att[c]) { P.att_estimator.push_back(a.first); P.att_index.push_back(a.second); }
        P.cells[c].tl_count = (int32_t)P.att_estimator.size() - P.cells[c].tl_begin;
        P.cells[c].col_begin = (int32_t)P.att_estimator.size();
        for (const auto& a : col_att[c]) { P.att_estimator.push_back(a.first); P.att_index.push_back(a.second); }
        P.cells[c].col_count = (int32_t)P.att_estimator.size() - P.cells[c].col_begin;
    }

    // ---------------------------------------------------------- parameters
    if (const XmlNode* prm = top("parameters")) {
        std::vector<int> clone_of_cell(P.cells.size(), -1);
        for (const XmlNode* q : prm->children_named("parameter")) {
            const int c = find_name(P.cell_names, q->value("cell"));
            need(c >= 0, MC_ERR_INPUT, "unknown cell '" + q->value("cell") + "' in parameter");
            need(P.cells[c].material >= 0, MC_ERR_INPUT, "parameter on void cell " + P.cell_names[c]);
            if (clone_of_cell[c] < 0) {      // per-cell material instance
                const mc_material_desc base = P.materials[P.cells[c].material];
                mc_material_desc M{};
                M.nuc_begin = (int32_t)P.mat_nuclide.size();
                M.nuc_count = base.nuc_count;
                for (int k = 0; k < base.nuc_count; ++k) {
                    P.mat_nuclide.push_back(P.mat_nuclide[base.nuc_begin + k]);
                    P.mat_density.push_back(P.mat_density[base.nuc_begin + k]);
                }
                P.materials.push_back(M);
                P.material_names.push_back(P.material_names[P.cells[c].material] + "@" + P.cell_names[c]);
                clone_of_cell[c] = (int)P.materials.size() - 1;
                P.cells[c].material = clone_of_cell[c];
            }
            const int m = clone_of_cell[c];
            int slot = -1;
            for (int k = 0; k < P.materials[m].nuc_count; ++k)
                if (P.nuclide_names[P.mat_nuclide[P.materials[m].nuc_begin + k]] == q->value("nuclide")) { slot = k; break; }
            need(slot >= 0, MC_ERR_INPUT, "nuclide '" + q->value("nuclide") + "' not in the material of cell " + P.cell_names[c]);
            const std::string r = q->has("reaction") ? q->value("reaction") : std::string("total");
            mc_param_desc D{};
            D.cell = c;
            D.material = m;
            D.nuc_slot = slot;
            D.reactions = r == "scatter" ? MC_REACTION_SCATTER : r == "capture" ? MC_REACTION_CAPTURE
                        : r == "total" ? (MC_REACTION_SCATTER | MC_REACTION_CAPTURE) : 0;
            need(D.reactions != 0, MC_ERR_INPUT, "parameter reaction must be scatter, capture or total");
            // one parameter per (cell, nuclide, reaction): a second one would
            // replace the first's scaled density while both claimed its gradient
            for (const mc_param_desc& O : P.params)
                need(!(O.material == m && O.nuc_slot == slot && (O.reactions & D.reactions)), MC_ERR_INPUT,
                     "two parameters scale the " + r + " cross section of nuclide '" + q->value("nuclide") +
                         "' in cell " + P.cell_names[c]);
            P.params.push_back(D);
        }
    }
    for (size_t m = 0; m < P.materials.size(); ++m) {
        P.materials[m].param_begin = (int32_t)P.mat_param.size();
        for (size_t k = 0; k < P.params.size(); ++k)
            if (P.params[k].material == (int)m) P.mat_param.push_back((int32_t)k);
        P.materials[m].param_count = (int32_t)P.mat_param.size() - P.materials[m].param_begin;
    }

    // ------------------------------------------------------------- sources
    if (const XmlNode* srcs = top("sources")) {                     // setup.cpp:1020-1066
        for (const auto& s : srcs->children) {
            mc_source_desc S{};
            S.probability = s.has("probability") ? s.as_double("probability") : 1.0;
            S.dir_kind = MC_DIR_ISOTROPIC;                           // setup.cpp:1027-1030
            S.energy_kind = MC_ENERGY_DELTA;
            S.energy[0] = 2e6;
            if (s.has("direction")) {
                auto it = pdist.find(s.value("direction"));
                need(it != pdist.end(), MC_ERR_INPUT, "unknown direction distribution in source");
                S.dir_kind = it->second.kind;
                S.dir[0] = it->second.x; S.dir[1] = it->second.y; S.dir[2] = it->second.z;
            }
            if (s.has("energy")) {
                auto it = ddist.find(s.value("energy"));
                need(it != ddist.end(), MC_ERR_INPUT, "unknown energy distribution in source");
                S.energy_kind = it->second.kind;
                S.energy[0] = it->second.a;
                S.energy[1] = it->second.b;
            }
            need(s.name == "point", s.name == "disk_z" ? MC_ERR_UNSUPPORTED : MC_ERR_INPUT,
                 "unknown source type: " + s.name);
            S.pos[0] = s.as_double("x"); S.pos[1] = s.as_double("y"); S.pos[2] = s.as_double("z");
            S.cell = P.search_cell(S.pos);
            need(S.cell >= 0, MC_ERR_LOST_PARTICLE, "source point is outside every cell");
            P.sources.push_back(S);
        }
    }
    need(!P.sources.empty(), MC_ERR_INPUT, "the deck defines no source");
    // SourceBank::set_up, Source.cpp:65-72: equal probabilities by repeated addition
    P.source_cdf.assign(P.sources.size() + 1, 0.0);
    const double dp = 1.0 / (double)P.sources.size();
    for (size_t i = 1; i < P.source_cdf.size(); ++i) P.source_cdf[i] = P.source_cdf[i - 1] + dp;
    need(P.n_histories > 1, MC_ERR_INPUT, "samples must be at least 2");
    P.finalize();
}

namespace {

template <class T>
std::vector<T> copy_n(const T* p, int64_t n, const char* what)
{
    if (n > 0 && !p) throw Error(MC_ERR_INPUT, std::string("mc_problem_desc.") + what + " is NULL");
    return n > 0 ? std::vector<T>(p, p + n) : std::vector<T>();
}

void in_range(int64_t v, int64_t lo, int64_t hi, const std::string& what)
{
    if (v < lo || v >= hi)
        throw Error(MC_ERR_INPUT, what + " = " + std::to_string(v) + " is outside [" + std::to_string(lo) + ", " +
                                      std::to_string(hi) + ")");
}

}  // namespace

void load_problem_desc(Problem& P, const mc_problem_desc& d)
{
    need(d.n_surfaces >= 0 && d.n_cells > 0 && d.n_materials >= 0 && d.n_nuclides >= 0 && d.n_sources > 0 &&
             d.n_estimators >= 0 && d.n_params >= 0 && d.n_bins >= 0,
         MC_ERR_INPUT, "mc_problem_desc: counts must be non-negative, with at least one cell and one source");
    P.surfaces = copy_n(d.surfaces, d.n_surfaces, "surfaces");
    P.cells = copy_n(d.cells, d.n_cells, "cells");
    P.materials = copy_n(d.materials, d.n_materials, "materials");
    P.nuclides = copy_n(d.nuclides, d.n_nuclides, "nuclides");
    P.sources = copy_n(d.sources, d.n_sources, "sources");
    P.estimators = copy_n(d.estimators, d.n_estimators, "estimators");
    P.params = copy_n(d.params, d.n_params, "params");
    // lengths of the indexed arrays follow from the records' ranges
    int64_t ncs = 0, nmn = 0, nxs = 0, natt = 0, neg = 0, nmp = 0;
    for (const auto& C : P.cells) {
        need(C.surf_begin >= 0 && C.surf_count >= 0 && C.tl_begin >= 0 && C.tl_count >= 0 && C.col_begin >= 0 &&
                 C.col_count >= 0,
             MC_ERR_INPUT, "mc_problem_desc: negative cell range");
        ncs = std::max<int64_t>(ncs, (int64_t)C.surf_begin + C.surf_count);
        natt = std::max<int64_t>(natt, std::max<int64_t>((int64_t)C.tl_begin + C.tl_count,
                                                         (int64_t)C.col_begin + C.col_count));
    }
    for (const auto& S : P.surfaces) {
        need(S.est_begin >= 0 && S.est_count >= 0, MC_ERR_INPUT, "mc_problem_desc: negative surface range");
        natt = std::max<int64_t>(natt, (int64_t)S.est_begin + S.est_count);
    }
    for (const auto& M : P.materials) {
        need(M.nuc_begin >= 0 && M.nuc_count >= 0 && M.param_begin >= 0 && M.param_count >= 0, MC_ERR_INPUT,
             "mc_problem_desc: negative material range");
        nmn = std::max<int64_t>(nmn, (int64_t)M.nuc_begin + M.nuc_count);
        nmp = std::max<int64_t>(nmp, (int64_t)M.param_begin + M.param_count);
    }
    for (const auto& N : P.nuclides) {
        if (N.kind != MC_NUC_TABLE) continue;
        need(N.grid_begin >= 0 && N.grid_count > 0, MC_ERR_INPUT, "mc_problem_desc: empty cross-section table");
        nxs = std::max<int64_t>(nxs, (int64_t)N.grid_begin + N.grid_count);
    }
    for (const auto& E : P.estimators) {
        need(E.egrid_begin >= 0 && E.egrid_count >= 0, MC_ERR_INPUT, "mc_problem_desc: negative energy-grid range");
        neg = std::max<int64_t>(neg, (int64_t)E.egrid_begin + E.egrid_count);
    }
    P.cell_surf = copy_n(d.cell_surf, ncs, "cell_surf");
    P.cell_sense = copy_n(d.cell_sense, ncs, "cell_sense");
    P.mat_nuclide = copy_n(d.mat_nuclide, nmn, "mat_nuclide");
    P.mat_density = copy_n(d.mat_density, nmn, "mat_density");
    P.xs_energy = copy_n(d.xs_energy, nxs, "xs_energy");
    P.xs_scatter = copy_n(d.xs_scatter, nxs, "xs_scatter");
    P.xs_capture = copy_n(d.xs_capture, nxs, "xs_capture");
    P.xs_total = copy_n(d.xs_total, nxs, "xs_total");
    P.att_estimator = copy_n(d.att_estimator, natt, "att_estimator");
    P.att_index = copy_n(d.att_index, natt, "att_index");
    P.egrid = copy_n(d.egrid, neg, "egrid");
    P.mat_param = copy_n(d.mat_param, nmp, "mat_param");
    // every index in range
    for (size_t k = 0; k < P.surfaces.size(); ++k) {
        const auto& S = P.surfaces[k];
        const bool known = S.type == MC_SURF_PLANE_X || S.type == MC_SURF_PLANE_Y || S.type == MC_SURF_PLANE_Z ||
                           S.type == MC_SURF_PLANE || S.type == MC_SURF_SPHERE || S.type == MC_SURF_CYL_X ||
                           S.type == MC_SURF_CYL_Z;
        need(known, MC_ERR_INPUT, "surface " + std::to_string(k) + ": unknown type " + std::to_string(S.type));
        need(S.bc >= -1 && S.bc <= 1, MC_ERR_INPUT, "surface " + std::to_string(k) + ": unknown boundary condition");
    }
    for (int64_t j = 0; j < ncs; ++j) {
        in_range(P.cell_surf[j], 0, (int64_t)P.surfaces.size(), "cell_surf[" + std::to_string(j) + "]");
        need(P.cell_sense[j] == 1 || P.cell_sense[j] == -1, MC_ERR_INPUT,
             "cell_sense[" + std::to_string(j) + "] must be +1 or -1");
    }
    for (size_t c = 0; c < P.cells.size(); ++c) {
        if (P.cells[c].material != -1)
            in_range(P.cells[c].material, 0, (int64_t)P.materials.size(), "cells[" + std::to_string(c) + "].material");
        need(P.cells[c].importance >= 0.0, MC_ERR_INPUT, "cells[" + std::to_string(c) + "]: negative importance");
    }
    for (int64_t j = 0; j < nmn; ++j) {
        in_range(P.mat_nuclide[j], 0, (int64_t)P.nuclides.size(), "mat_nuclide[" + std::to_string(j) + "]");
        need(P.mat_density[j] >= 0.0, MC_ERR_INPUT, "mat_density[" + std::to_string(j) + "] is negative");
    }
    for (size_t k = 0; k < P.nuclides.size(); ++k) {
        const auto& N = P.nuclides[k];
        need(N.kind == MC_NUC_CONSTANT || N.kind == MC_NUC_TABLE, MC_ERR_INPUT,
             "nuclides[" + std::to_string(k) + "]: unknown kind");
        need(N.A > 0.0, MC_ERR_INPUT, "nuclides[" + std::to_string(k) + "]: non-positive mass");
        need(N.kinematics == MC_KIN_TARGET_AT_REST || N.kinematics == MC_KIN_FREE_GAS, MC_ERR_INPUT,
             "nuclides[" + std::to_string(k) + "]: unknown kinematics");
    }
    if (nxs > 0) need_finite(P.xs_energy, "mc_problem_desc.xs_energy");
    int32_t bins = 0;
    for (size_t e = 0; e < P.estimators.size(); ++e) {
        const auto& E = P.estimators[e];
        const std::string en = "estimators[" + std::to_string(e) + "]";
        need(E.attach_kind == MC_ATTACH_SURFACE || E.attach_kind == MC_ATTACH_CELL, MC_ERR_INPUT,
             en + ": unknown attachment kind");
        need(E.n_scores > 0 && E.n_scores <= MC_MAX_SCORES && E.n_attach > 0, MC_ERR_INPUT,
             en + ": needs 1-8 scores and an attachment");
        for (int s = 0; s < E.n_scores; ++s) {
            need(E.score[s] >= MC_SCORE_FLUX && E.score[s] <= MC_SCORE_TOTAL, MC_ERR_INPUT, en + ": unknown score");
            need(E.score_kernel[s] >= MC_KERNEL_TRACK_LENGTH && E.score_kernel[s] <= MC_KERNEL_NEUTRON, MC_ERR_INPUT,
                 en + ": unknown score kernel");
        }
        need(E.egrid_count == 0 || E.egrid_count >= 2, MC_ERR_INPUT, en + ": an energy filter needs two edges");
        const int NE = E.egrid_count ? E.egrid_count - 1 : 1;
        need(E.bin_offset == bins && E.n_bins == E.n_scores * E.n_attach * NE, MC_ERR_INPUT,
             en + ": bins must follow the [score][attach][energy] layout, estimators in order");
        bins += E.n_bins;
    }
    need(bins == d.n_bins, MC_ERR_INPUT, "mc_problem_desc.n_bins differs from the estimators' bins");
    P.n_bins = d.n_bins;
    for (int64_t j = 0; j < natt; ++j) {
        in_range(P.att_estimator[j], 0, (int64_t)P.estimators.size(), "att_estimator[" + std::to_string(j) + "]");
        in_range(P.att_index[j], 0, P.estimators[P.att_estimator[j]].n_attach, "att_index[" + std::to_string(j) + "]");
    }
    for (size_t k = 0; k < P.params.size(); ++k) {
        const auto& Q = P.params[k];
        const std::string pn = "params[" + std::to_string(k) + "]";
        in_range(Q.cell, 0, (int64_t)P.cells.size(), pn + ".cell");
        in_range(Q.material, 0, (int64_t)P.materials.size(), pn + ".material");
        in_range(Q.nuc_slot, 0, P.materials[Q.material].nuc_count, pn + ".nuc_slot");
        need(Q.reactions >= 1 && Q.reactions <= 3, MC_ERR_INPUT, pn + ": unknown reactions mask");
        for (size_t o = 0; o < k; ++o)
            need(!(P.params[o].material == Q.material && P.params[o].nuc_slot == Q.nuc_slot &&
                   (P.params[o].reactions & Q.reactions)),
                 MC_ERR_INPUT, pn + " scales a reaction another parameter already scales");
    }
    for (int64_t j = 0; j < nmp; ++j) in_range(P.mat_param[j], 0, (int64_t)P.params.size(), "mat_param[" + std::to_string(j) + "]");
    for (size_t m = 0; m < P.materials.size(); ++m)
        for (int q = 0; q < P.materials[m].param_count; ++q)
            need(P.params[P.mat_param[P.materials[m].param_begin + q]].material == (int)m, MC_ERR_INPUT,
                 "mat_param of material " + std::to_string(m) + " lists a parameter of another material");
    // sources: the cell of the point (general.cpp:26-34) where not given
    for (size_t k = 0; k < P.sources.size(); ++k) {
        auto& S = P.sources[k];
        need(S.dir_kind == MC_DIR_ISOTROPIC || S.dir_kind == MC_DIR_DELTA, MC_ERR_INPUT,
             "sources[" + std::to_string(k) + "]: unknown direction kind");
        need(S.energy_kind == MC_ENERGY_DELTA || S.energy_kind == MC_ENERGY_UNIFORM, MC_ERR_INPUT,
             "sources[" + std::to_string(k) + "]: unknown energy kind");
        if (S.cell < 0) S.cell = P.search_cell(S.pos);
        need(S.cell >= 0, MC_ERR_LOST_PARTICLE, "sources[" + std::to_string(k) + "] is outside every cell");
        in_range(S.cell, 0, (int64_t)P.cells.size(), "sources[" + std::to_string(k) + "].cell");
    }
    if (d.source_cdf) {
        P.source_cdf.assign(d.source_cdf, d.source_cdf + d.n_sources + 1);
    } else {
        // SourceBank::set_up, Source.cpp:65-72
        P.source_cdf.assign(P.sources.size() + 1, 0.0);
        const double dp = 1.0 / (double)P.sources.size();
        for (size_t i = 1; i < P.source_cdf.size(); ++i) P.source_cdf[i] = P.source_cdf[i - 1] + dp;
    }
    P.wr = d.wr;
    P.ws = d.ws;
    need(P.wr >= 0.0 && P.ws > 0.0, MC_ERR_INPUT, "mc_problem_desc: roulette needs wr >= 0 and ws > 0");
    P.n_histories = d.n_histories;
    need(P.n_histories > 1, MC_ERR_INPUT, "mc_problem_desc.n_histories must be at least 2");
    P.seed = d.seed;
    need(P.seed > 0 && (P.seed & 1) == 1, MC_ERR_INPUT, "mc_problem_desc.seed must be a positive odd integer");
    P.comb_bank_max = d.comb_bank_max;
    P.comb_teeth = d.comb_teeth;
    need((P.comb_bank_max == 0 && P.comb_teeth == 0) || (P.comb_bank_max >= 1 && P.comb_teeth >= 1), MC_ERR_INPUT,
         "mc_problem_desc: comb_bank_max and comb_teeth are both 0 (no comb) or both at least 1");
    // names by position; report metadata from the records
    P.name = "problem";
    for (size_t k = 0; k < P.surfaces.size(); ++k) P.surface_names.push_back("surface " + std::to_string(k));
    for (size_t k = 0; k < P.cells.size(); ++k) P.cell_names.push_back("cell " + std::to_string(k));
    for (size_t k = 0; k < P.materials.size(); ++k) P.material_names.push_back("material " + std::to_string(k));
    for (size_t k = 0; k < P.nuclides.size(); ++k) P.nuclide_names.push_back("nuclide " + std::to_string(k));
    static const char* kScore[] = {"", "flux", "absorption", "scatter", "capture", "total"};
    for (size_t e = 0; e < P.estimators.size(); ++e) {
        const auto& E = P.estimators[e];
        P.estimator_names.push_back("Estimator #" + std::to_string(e + 1));
        std::vector<std::string> sn;
        for (int s = 0; s < E.n_scores; ++s)
            sn.push_back(E.score_kernel[s] == MC_KERNEL_NEUTRON ? "cross" : kScore[E.score[s]]);
        P.estimator_scores.push_back(sn);
        // the filter's geometry IDs, in filter-bin order
        std::vector<double> ids(E.n_attach, -1.0);
        auto note = [&](int32_t begin, int32_t count, size_t geom) {
            for (int32_t a = begin; a < begin + count; ++a)
                if (P.att_estimator[a] == (int32_t)e) ids[P.att_index[a]] = (double)geom;
        };
        for (size_t k = 0; k < P.surfaces.size(); ++k) note(P.surfaces[k].est_begin, P.surfaces[k].est_count, k);
        for (size_t c = 0; c < P.cells.size(); ++c) {
            note(P.cells[c].tl_begin, P.cells[c].tl_count, c);
            note(P.cells[c].col_begin, P.cells[c].col_count, c);
        }
        P.estimator_attach_ids.push_back(ids);
        P.estimator_egrid.emplace_back(P.egrid.begin() + E.egrid_begin, P.egrid.begin() + E.egrid_begin + E.egrid_count);
    }
    P.finalize();
}

}  // namespace mc
