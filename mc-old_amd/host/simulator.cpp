// simulator.cpp — mcold::Simulator over the C ABI (see simulator.hpp).
#include "simulator.hpp"

namespace mcold {

namespace {
void check(mc_status s)
{
    if (s != MC_OK) throw Error(s, mc_last_error());
}
}  // namespace

Simulator::Simulator(const std::string io_dir, const std::string xs_dir, const mc_engine_cfg* cfg)
{
    check(mc_problem_from_xml(io_dir.c_str(), xs_dir.c_str(), &problem));
    const mc_problem_desc* d = nullptr;
    check(mc_problem_get_desc(problem, &d));
    Nsample = d->n_histories;
    mc_status st = mc_engine_create(&problem, 1, cfg, &engine);
    if (st != MC_OK) {
        const std::string msg = mc_last_error();
        mc_problem_destroy(problem);
        problem = nullptr;
        throw Error(st, msg);
    }
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
    check(mc_forward(engine, nullptr, &out));
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
    check(mc_report_text(problem, &out, Nsample, (dir + "output.txt").c_str()));
}

}  // namespace mcold
