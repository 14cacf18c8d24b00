// A caller of the hip platform through nothing but the reference's own C++ interfaces
// (AbstractFactory, Molecules, PropagatorAnalyzer, PropagatorComputation): what a
// C++ user of Langevin_FTS writes, with `new HipFactory(...)` where PlatformSelector
// would hand out the factory.
//
//   hip_factory_check <spec> <fields in> <concentrations out>
//
// spec: lines "model continuous|discrete", "ds x", "nx a b c", "lx a b c",
// "bond NAME length", "aggregate 0|1", "polymer volume_fraction" followed by
// "block NAME contour_length v u" lines. Fields: the monomers' w in name order, grid
// size doubles each (binary). Prints "Q p value" per polymer and "stress s0 s1 s2";
// writes the total concentration of every monomer in name order (binary).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "HipFactory.h"

int main(int argc, char** argv) {
    if (argc != 4) {
        std::cerr << "usage: " << argv[0] << " <spec> <fields in> <concentrations out>" << std::endl;
        return 2;
    }
    try {
        std::ifstream spec(argv[1]);
        std::string model, line;
        double ds = 0.0;
        std::vector<int> nx;
        std::vector<double> lx;
        std::map<std::string, double> bonds;
        bool aggregate = true;
        std::vector<std::pair<double, std::vector<BlockInput>>> polymers;
        while (std::getline(spec, line)) {
            std::istringstream in(line);
            std::string key;
            if (!(in >> key)) continue;
            if (key == "model") in >> model;
            else if (key == "ds") in >> ds;
            else if (key == "nx") { int v; while (in >> v) nx.push_back(v); }
            else if (key == "lx") { double v; while (in >> v) lx.push_back(v); }
            else if (key == "bond") { std::string m; double b; in >> m >> b; bonds[m] = b; }
            else if (key == "aggregate") { int a; in >> a; aggregate = a != 0; }
            else if (key == "polymer") { double vf; in >> vf; polymers.push_back({vf, {}}); }
            else if (key == "block") {
                BlockInput b;
                in >> b.monomer_type >> b.contour_length >> b.v >> b.u;
                polymers.back().second.push_back(b);
            }
        }
        AbstractFactory* factory = new HipFactory(model, false);
        factory->display_info();
        ComputationBox* cb = factory->create_computation_box(nx, lx, {});
        Molecules* molecules = factory->create_molecules_information(model, ds, bonds);
        for (auto& p : polymers) molecules->add_polymer(p.first, p.second);
        PropagatorAnalyzer* analyzer = factory->create_propagator_analyzer(molecules, aggregate);
        PropagatorComputation* solver = factory->create_pseudospectral_solver(cb, molecules, analyzer);

        const int M = cb->get_n_grid();
        std::vector<std::vector<double>> w(bonds.size(), std::vector<double>(M));
        FILE* f = std::fopen(argv[2], "rb");
        if (!f) throw std::runtime_error("cannot read the fields");
        for (auto& x : w)
            if (std::fread(x.data(), sizeof(double), M, f) != (size_t)M) throw std::runtime_error("short fields file");
        std::fclose(f);
        std::map<std::string, const double*> w_block;
        int i = 0;
        for (const auto& b : bonds) w_block[b.first] = w[i++].data();

        solver->compute_statistics(w_block);
        solver->compute_stress();
        for (int p = 0; p < molecules->get_n_polymer_types(); p++)
            std::printf("Q %d %.17g\n", p, solver->get_total_partition(p));
        std::vector<double> stress = solver->get_stress();
        std::printf("stress");
        for (double s : stress) std::printf(" %.17g", s);
        std::printf("\n");
        std::printf("partition_check %d\n", solver->check_total_partition() ? 1 : 0);
        f = std::fopen(argv[3], "wb");
        std::vector<double> phi(M);
        for (const auto& b : bonds) {
            solver->get_total_concentration(b.first, phi.data());
            std::fwrite(phi.data(), sizeof(double), M, f);
        }
        std::fclose(f);
        delete solver;
        delete analyzer;
        delete molecules;
        delete cb;
        delete factory;
    } catch (std::exception& e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
