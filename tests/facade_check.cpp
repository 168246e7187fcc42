// facade_check.cpp — exercises the read-only object views of the C++
// Simulator facade (mc-old_amd/host/simulator.hpp) on a deck and prints them
// one per line for tests/test_facade.py.  With a third argument it also looks
// up every material's Sigma_t at that energy (needs the GPU).
#include <cstdio>
#include <cstdlib>

#include "simulator.hpp"

int main(int argc, char** argv)
{
    if (argc < 3) return 2;
    try {
        mcold::Simulator S(argv[1], argv[2]);
        std::printf("counts %zu %zu %zu %zu %llu\n", S.Surfaces.size(), S.Cells.size(), S.Materials.size(),
                    S.Nuclides.size(), S.Nsample);
        for (const auto& s : S.Surfaces) std::printf("surface %s %d %d\n", s->name().c_str(), s->type(), s->bc());
        for (const auto& n : S.Nuclides)
            std::printf("nuclide %s %.17g %d %d\n", n->name().c_str(), n->A(), n->grid_points(), (int)n->free_gas());
        for (const auto& m : S.Materials) {
            std::printf("material %s", m->name().c_str());
            for (const auto& nd : m->nuclides()) std::printf(" %s:%.17g", nd.first->name().c_str(), nd.second);
            std::printf("\n");
        }
        for (const auto& c : S.Cells) {
            std::printf("cell %s %.17g %s", c->name().c_str(), c->importance(),
                        c->material() ? c->material()->name().c_str() : "void");
            for (const auto& sv : c->surfaces()) std::printf(" %s:%+d", sv.first->name().c_str(), sv.second);
            std::printf("\n");
        }
        const mcold::Point p(1.5, 1.5, 0.0);
        const auto c = S.search_cell(p);
        std::printf("search %s\n", c ? c->name().c_str() : "none");
        mcold::Particle P(p, mcold::Point(0.6, 0.8, 0.0));
        for (const auto& s : S.Surfaces) {
            mcold::Particle Q = P;
            s->reflect(Q);
            std::printf("geom %s %.17g %.17g %.17g %.17g %.17g\n", s->name().c_str(), s->eval(p), s->distance(P),
                        Q.dir().x, Q.dir().y, Q.dir().z);
        }
        if (argc > 3) {
            const double E = std::atof(argv[3]);
            for (const auto& m : S.Materials) std::printf("sigt %s %.17g\n", m->name().c_str(), m->SigmaT(E));
        }
    } catch (const mcold::Error& e) {
        std::fprintf(stderr, "[ERROR %d] %s\n", (int)e.status, e.what());
        return 1;
    }
    return 0;
}
