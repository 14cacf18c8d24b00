// Checker-only driver around the reference's own common-layer sources.
//
// This file is test infrastructure. It is linked (by oracle/Makefile.ref) with
// the reference's unmodified sources read in place from the reference tree:
//   src/common/{Polymer,Molecules,PropagatorCode,PropagatorAnalyzer,Scheduler,
//               Pseudo,ComputationBox,AndersonMixing,CircularBuffer,
//               FiniteDifference}.cpp
//   src/platforms/cpu/CpuAndersonMixing.cpp
// and writes JSON dumps that the tests compare the product against:
//   mode plan    : Polymer keys, PropagatorAnalyzer propagators/blocks,
//                  Scheduler schedules for several stream counts
//   mode pseudo  : Pseudo::get_boltz_bond / get_weighted_fourier_basis tables
//   mode simpson : SimpsonRule::get_coeff
//   mode am      : CpuAndersonMixing trajectory on a fixed synthetic map
//   mode deps    : PropagatorCode::get_deps_from_key on given keys
//   mode laplacian : FiniteDifference::get_laplacian_matrix tridiagonal tables
//                    per axis for given boundary conditions
// Nothing here is shipped or measured. Input is a line-based spec on stdin.
#include <cmath>
#include <cstdio>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "AndersonMixing.h"
#include "ComputationBox.h"
#include "CpuAndersonMixing.h"
#include "FiniteDifference.h"
#include "Molecules.h"
#include "Polymer.h"
#include "PropagatorAnalyzer.h"
#include "PropagatorCode.h"
#include "Pseudo.h"
#include "Scheduler.h"
#include "SimpsonRule.h"

// ComputationBox has pure-virtual device methods; a host-only concrete box.
class HostBox : public ComputationBox {
public:
    HostBox(std::vector<int> nx, std::vector<double> lx, std::vector<std::string> bc = {})
        : ComputationBox(nx, lx, bc, nullptr) {}
    double integral_device(const double*) override { return 0; }
    double inner_product_device(const double*, const double*) override { return 0; }
    double inner_product_inverse_weight_device(const double*, const double*, const double*) override { return 0; }
    double multi_inner_product_device(int, const double*, const double*) override { return 0; }
    void zero_mean_device(double*) override {}
};

static void put_double(double v) { std::printf("%.17g", v); }

static void put_vec(const std::vector<double>& v) {
    std::printf("[");
    for (size_t i = 0; i < v.size(); i++) {
        if (i) std::printf(",");
        put_double(v[i]);
    }
    std::printf("]");
}

static void put_arr(const double* v, int n) {
    std::printf("[");
    for (int i = 0; i < n; i++) {
        if (i) std::printf(",");
        put_double(v[i]);
    }
    std::printf("]");
}

struct Spec {
    std::string mode;
    std::string model = "continuous";
    double ds = 0.1;
    std::map<std::string, double> bonds;
    struct P {
        double vf;
        std::vector<BlockInput> blocks;
        std::map<int, std::string> qinit;
    };
    std::vector<P> polymers;
    std::vector<std::pair<double, std::string>> solvents;
    int aggregate = 0;
    std::vector<int> streams;
    std::vector<int> nx;
    std::vector<double> lx;
    std::vector<int> simpson;
    std::vector<std::string> keys;
    std::vector<std::string> bc;
    // am
    int n_var = 64, max_hist = 20, n_iter = 40;
    double start_error = 1e-1, mix_min = 0.1, mix_init = 0.1;
};

static Spec read_spec() {
    Spec s;
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream is(line);
        std::string tok;
        if (!(is >> tok)) continue;
        if (tok == "mode") is >> s.mode;
        else if (tok == "model") is >> s.model;
        else if (tok == "ds") is >> s.ds;
        else if (tok == "bond") { std::string m; double b; is >> m >> b; s.bonds[m] = b; }
        else if (tok == "polymer") { Spec::P p; is >> p.vf; s.polymers.push_back(p); }
        else if (tok == "block") {
            BlockInput b; is >> b.monomer_type >> b.contour_length >> b.v >> b.u;
            s.polymers.back().blocks.push_back(b);
        }
        else if (tok == "qinit") { int node; std::string g; is >> node >> g; s.polymers.back().qinit[node] = g; }
        else if (tok == "solvent") { double vf; std::string m; is >> vf >> m; s.solvents.push_back({vf, m}); }
        else if (tok == "aggregate") is >> s.aggregate;
        else if (tok == "streams") { int n; while (is >> n) s.streams.push_back(n); }
        else if (tok == "nx") { int n; while (is >> n) s.nx.push_back(n); }
        else if (tok == "lx") { double l; while (is >> l) s.lx.push_back(l); }
        else if (tok == "simpson") { int n; while (is >> n) s.simpson.push_back(n); }
        else if (tok == "key") { std::string k; is >> k; s.keys.push_back(k); }
        else if (tok == "bc") { std::string b; while (is >> b) s.bc.push_back(b); }
        else if (tok == "am") is >> s.n_var >> s.max_hist >> s.start_error >> s.mix_min >> s.mix_init >> s.n_iter;
    }
    return s;
}

static void dump_plan(const Spec& s) {
    Molecules mol(s.model, s.ds, s.bonds);
    for (const auto& p : s.polymers) mol.add_polymer(p.vf, p.blocks, p.qinit);
    for (const auto& sv : s.solvents) mol.add_solvent(sv.first, sv.second);
    PropagatorAnalyzer an(&mol, s.aggregate != 0);

    std::printf("{\"model\":\"%s\",\"aggregate\":%d,\"polymers\":[", mol.get_model_name().c_str(), s.aggregate);
    for (int p = 0; p < mol.get_n_polymer_types(); p++) {
        Polymer& pc = mol.get_polymer(p);
        if (p) std::printf(",");
        std::printf("{\"alpha\":"); put_double(pc.get_alpha());
        std::printf(",\"volume_fraction\":"); put_double(pc.get_volume_fraction());
        std::printf(",\"n_segment_total\":%d,\"blocks\":[", pc.get_n_segment_total());
        auto& blocks = pc.get_blocks();
        for (size_t b = 0; b < blocks.size(); b++) {
            if (b) std::printf(",");
            std::printf("{\"monomer\":\"%s\",\"n_segment\":%d,\"v\":%d,\"u\":%d,\"key_vu\":\"%s\",\"key_uv\":\"%s\"}",
                        blocks[b].monomer_type.c_str(), blocks[b].n_segment, blocks[b].v, blocks[b].u,
                        pc.get_propagator_key(blocks[b].v, blocks[b].u).c_str(),
                        pc.get_propagator_key(blocks[b].u, blocks[b].v).c_str());
        }
        std::printf("]}");
    }
    std::printf("],\"propagators\":[");
    bool first = true;
    for (const auto& item : an.get_computation_propagators()) {
        if (!first) std::printf(",");
        first = false;
        std::printf("{\"key\":\"%s\",\"max_n_segment\":%d,\"monomer\":\"%s\",\"height\":%d,\"deps\":[",
                    item.first.c_str(), item.second.max_n_segment, item.second.monomer_type.c_str(),
                    item.second.height);
        for (size_t d = 0; d < item.second.deps.size(); d++) {
            if (d) std::printf(",");
            std::printf("[\"%s\",%d,%d]", std::get<0>(item.second.deps[d]).c_str(),
                        std::get<1>(item.second.deps[d]), std::get<2>(item.second.deps[d]));
        }
        std::printf("],\"junction_ends\":[");
        bool f2 = true;
        for (int n : item.second.junction_ends) {
            if (!f2) std::printf(",");
            f2 = false;
            std::printf("%d", n);
        }
        std::printf("]}");
    }
    std::printf("],\"blocks\":[");
    first = true;
    for (const auto& item : an.get_computation_blocks()) {
        if (!first) std::printf(",");
        first = false;
        const auto& blk = item.second;
        std::printf("{\"polymer\":%d,\"left\":\"%s\",\"right\":\"%s\",\"monomer\":\"%s\",\"n_left\":%d,"
                    "\"n_right\":%d,\"n_repeated\":%d,\"v_u\":[",
                    std::get<0>(item.first), std::get<1>(item.first).c_str(), std::get<2>(item.first).c_str(),
                    blk.monomer_type.c_str(), blk.n_segment_left, blk.n_segment_right, blk.n_repeated);
        for (size_t i = 0; i < blk.v_u.size(); i++) {
            if (i) std::printf(",");
            std::printf("[%d,%d]", std::get<0>(blk.v_u[i]), std::get<1>(blk.v_u[i]));
        }
        std::printf("]}");
    }
    std::printf("],\"schedules\":{");
    for (size_t si = 0; si < s.streams.size(); si++) {
        if (si) std::printf(",");
        Scheduler sc(an.get_computation_propagators(), s.streams[si]);
        std::printf("\"%d\":[", s.streams[si]);
        auto& sched = sc.get_schedule();
        for (size_t t = 0; t < sched.size(); t++) {
            if (t) std::printf(",");
            std::printf("[");
            for (size_t j = 0; j < sched[t].size(); j++) {
                if (j) std::printf(",");
                std::printf("[\"%s\",%d,%d]", std::get<0>(sched[t][j]).c_str(), std::get<1>(sched[t][j]),
                            std::get<2>(sched[t][j]));
            }
            std::printf("]");
        }
        std::printf("]");
    }
    std::printf("}}\n");
}

static void dump_pseudo(const Spec& s) {
    HostBox cb(s.nx, s.lx);
    const int MC = Pseudo::get_n_complex_grid(s.nx);
    std::vector<double> bb(MC), bbh(MC), fx(MC), fy(MC), fz(MC);
    std::printf("{\"n_complex\":%d,\"volume\":", MC);
    put_double(cb.get_volume());
    std::printf(",\"dx\":");
    put_vec(cb.get_dx());
    std::printf(",\"monomers\":{");
    bool first = true;
    for (const auto& b : s.bonds) {
        double bsq = b.second * b.second;
        Pseudo::get_boltz_bond(cb.get_boundary_conditions(), bb.data(), bsq, s.nx, cb.get_dx(), s.ds);
        Pseudo::get_boltz_bond(cb.get_boundary_conditions(), bbh.data(), bsq / 2, s.nx, cb.get_dx(), s.ds);
        if (!first) std::printf(",");
        first = false;
        std::printf("\"%s\":{\"boltz_bond\":", b.first.c_str());
        put_vec(bb);
        std::printf(",\"boltz_bond_half\":");
        put_vec(bbh);
        std::printf("}");
    }
    Pseudo::get_weighted_fourier_basis(cb.get_boundary_conditions(), fx.data(), fy.data(), fz.data(), s.nx,
                                       cb.get_dx());
    std::printf("},\"basis_x\":");
    put_vec(fx);
    std::printf(",\"basis_y\":");
    put_vec(fy);
    std::printf(",\"basis_z\":");
    put_vec(fz);
    std::printf("}\n");
}

// one tridiagonal table per axis and bond: lower, diagonal, upper coefficients
static void dump_laplacian(const Spec& s) {
    HostBox cb(s.nx, s.lx, s.bc);
    const int dim = (int)s.nx.size();
    int nmax = 1;
    for (int n : s.nx) nmax = n > nmax ? n : nmax;
    std::vector<double> t(9 * (size_t)nmax, 0.0);
    std::printf("{\"monomers\":{");
    bool first = true;
    for (const auto& b : s.bonds) {
        double* p = t.data();
        FiniteDifference::get_laplacian_matrix(cb.get_boundary_conditions(), s.nx, cb.get_dx(), p, p + nmax,
                                               p + 2 * nmax, p + 3 * nmax, p + 4 * nmax, p + 5 * nmax,
                                               p + 6 * nmax, p + 7 * nmax, p + 8 * nmax, b.second * b.second, s.ds);
        if (!first) std::printf(",");
        first = false;
        std::printf("\"%s\":[", b.first.c_str());
        for (int d = 0; d < dim; d++) {
            if (d) std::printf(",");
            std::printf("{\"l\":");
            put_arr(p + (3 * d) * nmax, s.nx[d]);
            std::printf(",\"d\":");
            put_arr(p + (3 * d + 1) * nmax, s.nx[d]);
            std::printf(",\"h\":");
            put_arr(p + (3 * d + 2) * nmax, s.nx[d]);
            std::printf("}");
        }
        std::printf("]");
    }
    std::printf("}}\n");
}

static void dump_simpson(const Spec& s) {
    std::printf("{");
    for (size_t i = 0; i < s.simpson.size(); i++) {
        if (i) std::printf(",");
        std::printf("\"%d\":", s.simpson[i]);
        put_vec(SimpsonRule::get_coeff(s.simpson[i]));
    }
    std::printf("}\n");
}

static void dump_deps(const Spec& s) {
    std::printf("{");
    for (size_t i = 0; i < s.keys.size(); i++) {
        if (i) std::printf(",");
        auto deps = PropagatorCode::get_deps_from_key(s.keys[i]);
        std::printf("\"%s\":{\"monomer\":\"%s\",\"height\":%d,\"deps\":[", s.keys[i].c_str(),
                    PropagatorCode::get_monomer_type_from_key(s.keys[i]).c_str(),
                    PropagatorCode::get_height_from_key(s.keys[i]));
        for (size_t d = 0; d < deps.size(); d++) {
            if (d) std::printf(",");
            std::printf("[\"%s\",%d,%d]", std::get<0>(deps[d]).c_str(), std::get<1>(deps[d]), std::get<2>(deps[d]));
        }
        std::printf("]}");
    }
    std::printf("}\n");
}

// Fixed synthetic fixed-point problem for the Anderson-mixing trajectory:
// target map F(w)_i = 0.5*tanh(w_i) + 0.3*cos(0.1*i) + 0.2*w_{(i+1) mod n};
// residual d = F(w) - w; error = sqrt(sum d^2 / (1 + sum w^2)).
static void dump_am(const Spec& s) {
    const int n = s.n_var;
    CpuAndersonMixing am(n, s.max_hist, s.start_error, s.mix_min, s.mix_init);
    std::vector<double> w(n), d(n), wn(n);
    for (int i = 0; i < n; i++) w[i] = std::sin(0.37 * i);
    double err = 1e20, old_err;
    std::printf("{\"n_var\":%d,\"w0\":", n);
    put_vec(w);
    std::printf(",\"steps\":[");
    for (int it = 0; it < s.n_iter; it++) {
        double sd = 0, sw = 0;
        for (int i = 0; i < n; i++) {
            double f = 0.5 * std::tanh(w[i]) + 0.3 * std::cos(0.1 * i) + 0.2 * w[(i + 1) % n];
            d[i] = f - w[i];
            sd += d[i] * d[i];
            sw += w[i] * w[i];
        }
        old_err = err;
        err = std::sqrt(sd / (1.0 + sw));
        am.calculate_new_fields(wn.data(), w.data(), d.data(), old_err, err);
        if (it) std::printf(",");
        std::printf("{\"err\":");
        put_double(err);
        std::printf(",\"w_new\":");
        put_arr(wn.data(), n);
        std::printf("}");
        w = wn;
    }
    std::printf("]}\n");
}

int main() {
    try {
        Spec s = read_spec();
        if (s.mode == "plan") dump_plan(s);
        else if (s.mode == "pseudo") dump_pseudo(s);
        else if (s.mode == "simpson") dump_simpson(s);
        else if (s.mode == "deps") dump_deps(s);
        else if (s.mode == "am") dump_am(s);
        else if (s.mode == "laplacian") dump_laplacian(s);
        else { std::fprintf(stderr, "unknown mode '%s'\n", s.mode.c_str()); return 2; }
    } catch (std::exception& e) {
        std::printf("{\"error\":");
        std::string msg = e.what();
        std::string esc;
        for (char c : msg) {
            if (c == '"' || c == '\\') esc += '\\';
            if (c == '\n') { esc += "\\n"; continue; }
            if (c == '\t') { esc += "\\t"; continue; }
            esc += c;
        }
        std::printf("\"%s\"}\n", esc.c_str());
        return 0;
    }
    return 0;
}
