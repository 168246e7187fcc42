// simulator.cpp — mcold::Simulator over the C ABI (see simulator.hpp).
#include "simulator.hpp"

#include "../csrc/mc_physics.hpp"

namespace mcold {

namespace {
void check(mc_status s)
{
    if (s != MC_OK) throw Error(s, mc_last_error());
}
}  // namespace

double Surface::eval(const Point& p) const
{
    return mc::surface_eval(s_desc.type, s_desc.c, mc::Vec3{p.x, p.y, p.z});
}

double Surface::distance(const Particle& P) const
{
    const Point p = P.pos(), u = P.dir();
    return mc::surface_distance(s_desc.type, s_desc.c, mc::Vec3{p.x, p.y, p.z}, mc::Vec3{u.x, u.y, u.z});
}

void Surface::reflect(Particle& P) const
{
    const Point u = P.dir();
    const mc::Vec3 r = mc::surface_reflect(s_desc.type, s_desc.c, mc::Vec3{u.x, u.y, u.z});
    P.set_direction(Point(r.x, r.y, r.z));
}

bool Cell::test_point(const Point& p) const
{
    for (const auto& s : c_surfaces)
        if (s.first->eval(p) * (double)s.second < 0) return false;
    return true;
}

std::vector<double> Material::SigmaT(const std::vector<double>& E) const
{
    std::vector<double> out(E.size(), 0.0);
    if (E.empty()) return out;
    std::vector<int32_t> mat(E.size(), (int32_t)m_index);
    check(mc_xs_lookup(m_owner->device_engine(), E.data(), mat.data(), out.data(), E.size()));
    return out;
}

double Material::SigmaT(double E) const { return SigmaT(std::vector<double>{E})[0]; }

Simulator::Simulator(const std::string io_dir, const std::string xs_dir, const mc_engine_cfg* cfg)
{
    check(mc_problem_from_xml(io_dir.c_str(), xs_dir.c_str(), &problem));
    if (cfg) {
        cfg_copy = *cfg;
        engine_cfg = &cfg_copy;
    }
    const mc_problem_desc* d = nullptr;
    check(mc_problem_get_desc(problem, &d));
    Nsample = d->n_histories;
    // read-only views of the object graph, in deck order
    for (int n = 0; n < d->n_nuclides; ++n)
        Nuclides.push_back(std::make_shared<Nuclide>(mc_problem_object_name(problem, MC_OBJ_NUCLIDE, n),
                                                     d->nuclides[n]));
    for (int m = 0; m < d->n_materials; ++m) {
        const mc_material_desc& M = d->materials[m];
        std::vector<std::pair<std::shared_ptr<Nuclide>, double>> nd;
        for (int k = 0; k < M.nuc_count; ++k)
            nd.emplace_back(Nuclides[d->mat_nuclide[M.nuc_begin + k]], d->mat_density[M.nuc_begin + k]);
        Materials.push_back(std::make_shared<Material>(mc_problem_object_name(problem, MC_OBJ_MATERIAL, m), m,
                                                       this, std::move(nd)));
    }
    for (int s = 0; s < d->n_surfaces; ++s)
        Surfaces.push_back(std::make_shared<Surface>(mc_problem_object_name(problem, MC_OBJ_SURFACE, s), s,
                                                     d->surfaces[s]));
    for (int c = 0; c < d->n_cells; ++c) {
        const mc_cell_desc& C = d->cells[c];
        std::vector<std::pair<std::shared_ptr<Surface>, int>> sv;
        for (int k = 0; k < C.surf_count; ++k)
            sv.emplace_back(Surfaces[d->cell_surf[C.surf_begin + k]], d->cell_sense[C.surf_begin + k]);
        Cells.push_back(std::make_shared<Cell>(mc_problem_object_name(problem, MC_OBJ_CELL, c), c, C.importance,
                                               C.material >= 0 ? Materials[C.material] : nullptr, std::move(sv)));
    }
}

mc_engine* Simulator::device_engine()
{
    if (!engine) check(mc_engine_create(&problem, 1, engine_cfg, &engine));
    return engine;
}

std::shared_ptr<Cell> Simulator::search_cell(const Point& p) const
{
    for (const auto& c : Cells)
        if (c->test_point(p)) return c;
    return nullptr;
}

Simulator::~Simulator()
{
    if (engine) mc_engine_destroy(engine);
    if (problem) mc_problem_destroy(problem);
}

void Simulator::start()
{
    const mc_problem_desc* d = nullptr;
    check(mc_problem_get_desc(problem, &d));
    const size_t nb = (size_t)(d->n_bins > 0 ? d->n_bins : 1);
    sum.assign(nb, 0.0);
    sq.assign(nb, 0.0);
    mean.assign(nb, 0.0);
    uncer.assign(nb, 0.0);
    mc_tally_out out{};
    out.sum = sum.data();
    out.sq = sq.data();
    out.mean = mean.data();
    out.uncer = uncer.data();
    check(mc_forward(device_engine(), nullptr, &out));
    Ntrack = out.flights;
    device_seconds = out.seconds;
    Estimators.clear();
    for (int e = 0; e < d->n_estimators; ++e) {
        const mc_estimator_desc& E = d->estimators[e];
        std::vector<Tally> t(E.n_bins);
        for (int b = 0; b < E.n_bins; ++b) {
            t[b].mean = mean[E.bin_offset + b];
            t[b].uncer = uncer[E.bin_offset + b];
        }
        const char* name = mc_problem_estimator_name(problem, e);
        Estimators.push_back(std::make_shared<Estimator>(name ? name : "", std::move(t)));
    }
    started = true;
}

void Simulator::report(const std::string io_dir)
{
    if (!started) throw Error(MC_ERR_ARGUMENT, "start() has not run");
    std::string dir = io_dir;
    if (!dir.empty() && dir.back() != '/') dir += '/';
    mc_tally_out out{};
    out.sum = sum.data();
    out.sq = sq.data();
    out.mean = mean.data();
    out.uncer = uncer.data();
    out.flights = Ntrack;
    out.seconds = device_seconds;
    // report.cpp:20-21 writes <io_dir>/output.h5; without an HDF5 library the
    // same content goes to output.txt
    const mc_status st = mc_report_h5(problem, &out, Nsample, (dir + "output.h5").c_str());
    if (st == MC_ERR_UNSUPPORTED) {
        check(mc_report_text(problem, &out, Nsample, (dir + "output.txt").c_str()));
        report_path = dir + "output.txt";
    } else {
        check(st);
        report_path = dir + "output.h5";
    }
}

}  // namespace mcold
