// Chain model, propagator keys, analyzer and scheduler (see fts_model.h).
//
// The algorithms are the reference's and must agree with its output exactly:
//   Polymer.cpp:14-187          validation, adjacency, cycle/isolation checks
//   PropagatorCode.cpp:19-290   memoised edge codes and key parsing
//   PropagatorAnalyzer.cpp:142-530  deduplication, aggregation, substitution
//   Scheduler.cpp:11-258        lane assignment and time slices
// Map iteration orders, the non-stable std::sort on resolved/start times and the
// in-loop mutations of the block maps are reproduced as they are, because they
// decide key strings, block slicing and the slice boundaries.
#include "fts_model.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <iostream>
#include <sstream>
#include <stack>

namespace lfts {

// ------------------------------------------------------------------ Polymer
Polymer::Polymer(double ds, const std::map<std::string, double>& bond_lengths, double volume_fraction,
                 const std::vector<BlockInput>& block_inputs, const std::map<int, std::string>& chain_end_to_q_init) {
    for (const auto& item : bond_lengths) {
        const std::string& name = item.first;
        if (!std::all_of(name.begin(), name.end(), [](unsigned char c) { return std::isalpha(c) || c == '_'; }))
            LFTS_THROW("\"" + name + "\" is an invalid monomer_type name. Only alphabets and underscore(_) are allowed.");
        if (item.second <= 0) LFTS_THROW("bond_lengths[\"" + name + "\"] must be a positive number.");
    }
    for (size_t i = 0; i < block_inputs.size(); i++) {
        const BlockInput& b = block_inputs[i];
        if (b.contour_length <= 0)
            LFTS_THROW("block_inputs[" + std::to_string(i) + "].contour_lengths (" +
                       std::to_string(b.contour_length) + ") must be a positive number.");
        if (std::abs(std::lround(b.contour_length / ds) - b.contour_length / ds) > 1.e-6)
            LFTS_THROW("block_inputs[" + std::to_string(i) + "].contour_lengths/ds (" +
                       std::to_string(b.contour_length) + "/" + std::to_string(ds) + ") is not an integer.");
        if (bond_lengths.count(b.monomer_type) == 0)
            LFTS_THROW("block_inputs[" + std::to_string(i) + "].monomer_type (\"" + b.monomer_type +
                       "\") is not in bond_lengths.");
    }
    if (block_inputs.empty()) LFTS_THROW("A polymer needs at least one block.");
    q_init_ = chain_end_to_q_init;
    volume_fraction_ = volume_fraction;
    for (const BlockInput& b : block_inputs)
        blocks_.push_back({b.monomer_type, (int)std::lround(b.contour_length / ds), b.contour_length, b.v, b.u});
    alpha_ = 0.0;
    for (const Block& b : blocks_) alpha_ += b.contour_length;

    for (size_t i = 0; i < block_inputs.size(); i++) {
        const int v = block_inputs[i].v, u = block_inputs[i].u;
        adjacent_[v].push_back(u);
        adjacent_[u].push_back(v);
        if (v < 0) LFTS_THROW("v[" + std::to_string(i) + "] (" + std::to_string(v) + ") must be a non-negative integer.");
        if (u < 0) LFTS_THROW("u[" + std::to_string(i) + "] (" + std::to_string(u) + ") must be a non-negative integer.");
        if (v == u)
            LFTS_THROW("v[" + std::to_string(i) + "] and u[" + std::to_string(i) + "] must be different integers.");
    }
    for (size_t i = 0; i < block_inputs.size(); i++) {
        const int v = block_inputs[i].v, u = block_inputs[i].u;
        if (edge_to_block_.count({v, u}) > 0)
            LFTS_THROW("There are duplicated edges. Please check the edge between (" + std::to_string(v) + ", " +
                       std::to_string(u) + ").");
        edge_to_block_[{v, u}] = (int)i;
        edge_to_block_[{u, v}] = (int)i;
    }

    // depth-first search from the first block's v: cycles and unreachable nodes
    std::map<int, bool> visited;
    for (const BlockInput& b : block_inputs) visited[b.v] = false;
    for (const BlockInput& b : block_inputs) visited[b.u] = false;
    std::stack<std::pair<int, int>> todo;
    todo.push({block_inputs[0].v, -1});
    while (!todo.empty()) {
        const int cur = todo.top().first, parent = todo.top().second;
        visited[cur] = true;
        todo.pop();
        const std::vector<int> nodes = adjacent_[cur];
        for (int n : nodes) {
            if (visited[n] && n != parent)
                LFTS_THROW("A cycle is detected, which contains nodes " + std::to_string(n) + " and " +
                           std::to_string(parent) + ". Only acyclic branched polymers are allowed.");
            else if (!visited[n])
                todo.push({n, cur});
        }
    }
    std::set<int> isolated;
    for (const BlockInput& b : block_inputs) {
        if (!visited[b.v]) isolated.insert(b.v);
        if (!visited[b.u]) isolated.insert(b.u);
    }
    if (!isolated.empty()) {
        std::string msg = "There is no route from node " + std::to_string(block_inputs[0].v) + " to nodes: ";
        bool first = true;
        for (int n : isolated) {
            msg += (first ? "" : ", ") + std::to_string(n);
            first = false;
        }
        LFTS_THROW(msg + ".");
    }

    for (const auto& c : propagator_code::generate_codes(*this, q_init_))
        edge_key_[{std::get<0>(c), std::get<1>(c)}] = propagator_code::key_from_code(std::get<2>(c));
}

Block& Polymer::get_block(int v, int u) {
    auto it = edge_to_block_.find({v, u});
    if (it == edge_to_block_.end())
        LFTS_THROW("There is no such edge (" + std::to_string(v) + ", " + std::to_string(u) + ").");
    return blocks_[it->second];
}

int Polymer::get_block_index_from_edge(int v, int u) const {
    auto it = edge_to_block_.find({v, u});
    if (it == edge_to_block_.end())
        LFTS_THROW("There is no such edge (" + std::to_string(v) + ", " + std::to_string(u) + ").");
    return it->second;
}

int Polymer::get_n_segment_total() const {
    int n = 0;
    for (const Block& b : blocks_) n += b.n_segment;
    return n;
}

std::string Polymer::get_propagator_key(int v, int u) const {
    auto it = edge_key_.find({v, u});
    if (it == edge_key_.end())
        LFTS_THROW("There is no such block (v, u): (" + std::to_string(v) + ", " + std::to_string(u) + ").");
    return it->second;
}

// ------------------------------------------------------------------ Molecules
Molecules::Molecules(std::string model_name, double ds, const std::map<std::string, double>& bond_lengths) {
    std::transform(model_name.begin(), model_name.end(), model_name.begin(),
                   [](unsigned char c) { return std::tolower(c); });
    if (model_name != "continuous" && model_name != "discrete")
        LFTS_THROW(model_name + " is an invalid chain model. This must be 'Continuous' or 'Discrete'.");
    model_name_ = model_name;
    ds_ = ds;
    bond_lengths_ = bond_lengths;
}

void Molecules::add_polymer(double volume_fraction, const std::vector<BlockInput>& blocks,
                            const std::map<int, std::string>& chain_end_to_q_init) {
    polymers_.push_back(Polymer(ds_, bond_lengths_, volume_fraction, blocks, chain_end_to_q_init));
}

void Molecules::add_solvent(double volume_fraction, const std::string& monomer_type) {
    solvents_.push_back(std::make_tuple(volume_fraction, monomer_type));
}

Polymer& Molecules::get_polymer(int p) {
    if (p < 0 || p >= (int)polymers_.size())
        LFTS_THROW("Index (" + std::to_string(p) + ") must be in range [0, " + std::to_string(polymers_.size() - 1) + "]");
    return polymers_[p];
}

// ------------------------------------------------------------------ propagator codes
namespace propagator_code {

namespace {

// code of the propagator entering out_node from in_node: sorted codes of the
// propagators arriving at in_node from its other neighbours, then the block.
std::string edge_code(std::map<std::pair<int, int>, std::string>& memo, Polymer& pc,
                      const std::map<int, std::string>& q_init, int in_node, int out_node) {
    auto hit = memo.find({in_node, out_node});
    if (hit != memo.end()) return hit->second;
    std::vector<std::string> subs;
    for (int nb : pc.get_adjacent_nodes()[in_node]) {
        if (nb == out_node) continue;
        auto m = memo.find({nb, in_node});
        std::string sub;
        if (m != memo.end()) {
            sub = m->second;
        } else {
            sub = edge_code(memo, pc, q_init, nb, in_node);
            memo[{nb, in_node}] = sub;
        }
        subs.push_back(sub);
    }
    std::string code;
    if (!subs.empty()) {
        std::sort(subs.begin(), subs.end());
        code = "(";
        for (const auto& s : subs) code += s;
        code += ")";
    } else {
        auto q = q_init.find(in_node);
        if (q != q_init.end()) code = "{" + q->second + "}";
    }
    const Block& b = pc.get_blocks()[pc.get_block_indexes()[{in_node, out_node}]];
    code += b.monomer_type + std::to_string(b.n_segment);
    memo[{in_node, out_node}] = code;
    return code;
}

}  // namespace

std::vector<std::tuple<int, int, std::string>> generate_codes(Polymer& pc, const std::map<int, std::string>& q_init) {
    std::vector<std::tuple<int, int, std::string>> out;
    std::map<std::pair<int, int>, std::string> memo;
    for (const Block& b : pc.get_blocks()) {
        const int v = b.v, u = b.u;
        out.push_back(std::make_tuple(v, u, edge_code(memo, pc, q_init, v, u)));
        out.push_back(std::make_tuple(u, v, edge_code(memo, pc, q_init, u, v)));
    }
    return out;
}

// drop the trailing segment count: everything up to the last letter
std::string key_from_code(const std::string& code) {
    for (int i = (int)code.size() - 1; i >= 0; i--)
        if (std::isalpha((unsigned char)code[i])) return code.substr(0, i + 1);
    return code;
}

// dependencies (key, n_segment, n_repeated) listed at bracket depth one:
// "(K1n1K2n2)X" or "[K1n1:r1,K2n2]X"
std::vector<std::tuple<std::string, int, int>> deps_from_key(const std::string& key) {
    std::vector<std::tuple<std::string, int, int>> deps;
    std::vector<char> open;
    enum { KEY, COUNT, REPEAT } state = KEY;
    size_t start = 1;
    std::string sub;
    int n = 0;
    for (size_t i = 0; i < key.size(); i++) {
        const char c = key[i];
        if (open.size() == 1) {
            const bool digit = std::isdigit((unsigned char)c) != 0;
            if (state == KEY && digit) {
                if (key[start] == ',') start += 1;
                sub = key.substr(start, i - start);
                state = COUNT;
                start = i;
            } else if (state == COUNT && c == ':') {
                n = std::stoi(key.substr(start, i - start));
                state = REPEAT;
                start = i + 1;
            } else if (state == COUNT && !digit) {
                n = std::stoi(key.substr(start, i - start));
                deps.push_back(std::make_tuple(sub, n, 1));
                state = KEY;
                start = i;
            } else if (state == REPEAT && !digit) {
                const int r = std::stoi(key.substr(start, i - start));
                deps.push_back(std::make_tuple(sub, n, r));
                state = KEY;
                start = i;
            }
        }
        if (c == '(' || c == '[') {
            open.push_back(c);
        } else if (c == ')' || c == ']') {
            if (!open.empty() && ((open.back() == '(' && c == ')') || (open.back() == '[' && c == ']')))
                open.pop_back();
            else
                LFTS_THROW("Failed to match parentheses or braces.");
        }
    }
    return deps;
}

std::string monomer_type_from_key(const std::string& key) {
    for (int i = (int)key.size() - 1; i >= 0; i--)
        if (key[i] == ')' || key[i] == ']' || key[i] == '}') return key.substr(i + 1);
    return key;
}

std::string q_input_name_from_key(const std::string& key) {
    if (key.empty() || key[0] != '{')
        LFTS_THROW("There is no related initial condition in key (" + key + ").");
    const size_t close = key.rfind('}');
    return key.substr(1, close - 1);
}

int height_from_key(const std::string& key) {
    int h = 0;
    for (char c : key) {
        if (c == '[' || c == '(') h++;
        else break;
    }
    return h;
}

}  // namespace propagator_code

bool ComparePropagatorKey::operator()(const std::string& a, const std::string& b) const {
    const int ha = propagator_code::height_from_key(a), hb = propagator_code::height_from_key(b);
    if (ha != hb) return ha < hb;
    return a > b;
}

// ------------------------------------------------------------------ analyzer
PropagatorAnalyzer::PropagatorAnalyzer(Molecules* molecules, bool aggregate) {
    if (molecules->get_n_polymer_types() == 0) LFTS_THROW("There is no chain. Add polymers first.");
    aggregate_ = aggregate;
    model_name_ = molecules->get_model_name();
    for (int p = 0; p < molecules->get_n_polymer_types(); p++) add_polymer(molecules->get_polymer(p), p);
}

ComputationEdge& PropagatorAnalyzer::get_computation_propagator(const std::string& key) {
    auto it = propagators_.find(key);
    if (it == propagators_.end()) LFTS_THROW("There is no such key (" + key + ").");
    return it->second;
}

ComputationBlock& PropagatorAnalyzer::get_computation_block(const BlockKey& key) {
    auto it = blocks_.find(key);
    if (it == blocks_.end())
        LFTS_THROW("There is no such key (" + std::to_string(std::get<0>(key)) + ", " + std::get<1>(key) + ", " +
                   std::get<2>(key) + ").");
    return it->second;
}

void PropagatorAnalyzer::add_polymer(Polymer& pc, int polymer_id) {
    std::map<std::string, std::map<std::string, ComputationBlock>> nb;
    std::map<std::tuple<int, int>, std::string> v_u_to_right_key;
    const std::vector<Block> blocks = pc.get_blocks();
    for (const Block& b : blocks) {
        int v = b.v, u = b.u;
        std::string left = pc.get_propagator_key(v, u), right = pc.get_propagator_key(u, v);
        if (left < right) {
            left.swap(right);
            std::swap(v, u);
        }
        ComputationBlock& cb = nb[left][right];
        cb.monomer_type = b.monomer_type;
        cb.n_segment_right = b.n_segment;
        cb.n_segment_left = b.n_segment;
        cb.v_u.push_back(std::make_tuple(v, u));
        cb.n_repeated = (int)cb.v_u.size();
        v_u_to_right_key[std::make_tuple(v, u)] = right;
    }

    // total segment count (PropagatorAnalyzer.cpp:171-189), used by the display
    int total = 0;
    for (const Block& b : blocks) {
        if (model_name_ == "continuous") {
            total += b.n_segment;
        } else {
            total += b.n_segment - 1;
            if (pc.get_adjacent_nodes()[b.v].size() != 1) total++;
            if (pc.get_adjacent_nodes()[b.u].size() != 1) total++;
        }
    }
    total_segment_numbers_.push_back(total);

    if (aggregate_) {
        std::map<std::string, std::vector<std::string>> aggregated;
        // the map gains first-level entries while it is walked; std::map iteration
        // then visits those that sort after the current position
        for (auto it = nb.begin(); it != nb.end(); ++it) {
            const std::string left = it->first;
            const std::map<std::string, ComputationBlock> rights = it->second;
            std::map<std::string, ComputationBlock> set_I = aggregate(rights, model_name_ == "continuous" ? 0 : 1);
            nb[left] = set_I;
            for (const auto& r : set_I)
                if (r.first[0] == '[' && r.second.n_segment_right == r.second.n_segment_left)
                    aggregated[left].push_back(r.first);
            substitute_right_keys(pc, v_u_to_right_key, nb, aggregated, left);
        }
    }

    for (const auto& vi : nb) {
        for (const auto& ui : vi.second) {
            const std::string& kv = vi.first;
            const std::string& ku = ui.first;
            ComputationBlock& cb = blocks_[std::make_tuple(polymer_id, kv, ku)];
            cb.monomer_type = propagator_code::monomer_type_from_key(kv);
            cb.n_segment_right = ui.second.n_segment_right;
            cb.n_segment_left = ui.second.n_segment_left;
            cb.v_u = ui.second.v_u;
            cb.n_repeated = ui.second.n_repeated;
            const bool junction_left = propagator_code::height_from_key(kv) > 0;
            const bool junction_right = propagator_code::height_from_key(ku) > 0;
            update_propagator(kv, ui.second.n_segment_left, junction_right);
            update_propagator(ku, ui.second.n_segment_right, junction_left);
        }
    }
}

std::map<std::string, ComputationBlock> PropagatorAnalyzer::aggregate(std::map<std::string, ComputationBlock> set_I,
                                                                      int minimum_n_segment) {
    std::set<int> n_values;
    for (const auto& it : set_I) n_values.insert(it.second.n_segment_right);
    for (const int n_cur : n_values) {
        std::map<std::string, ComputationBlock, ComparePropagatorKey> set_S;
        for (const auto& it : set_I)
            if (it.second.n_segment_right == n_cur) set_S[it.first] = it.second;
        if (set_S.size() == 1 || n_cur < 2 * minimum_n_segment) continue;
        const std::string monomer = propagator_code::monomer_type_from_key(set_S.begin()->second.monomer_type);
        bool same = true;
        for (auto it = set_S.rbegin(); it != set_S.rend(); ++it)
            if (it->second.monomer_type != monomer) same = false;
        // once all pieces share a monomer the slices are cut to zero segments,
        // and the reduced minimum stays in force for the larger n_cur that follow
        if (same) minimum_n_segment = 0;
        for (const auto& it : set_S) set_I[it.first].n_segment_right = minimum_n_segment;
        const int n_new = n_cur - minimum_n_segment;
        std::vector<std::tuple<int, int>> v_u;
        std::string code = "[";
        bool first = true;
        for (auto it = set_S.rbegin(); it != set_S.rend(); ++it) {
            const std::string& dep = it->first;
            if (!first) code += ",";
            first = false;
            code += dep + std::to_string(set_I[dep].n_segment_right);
            if (set_I[dep].n_repeated > 1) code += ":" + std::to_string(set_I[dep].n_repeated);
            v_u.insert(v_u.end(), set_I[dep].v_u.begin(), set_I[dep].v_u.end());
        }
        code += "]" + monomer;
        ComputationBlock& agg = set_I[code];
        agg.monomer_type = monomer;
        agg.n_segment_right = n_new;
        agg.n_segment_left = n_new;
        agg.v_u = v_u;
        agg.n_repeated = 1;
    }
    return set_I;
}

void PropagatorAnalyzer::substitute_right_keys(Polymer& pc,
                                               std::map<std::tuple<int, int>, std::string>& v_u_to_right_key,
                                               std::map<std::string, std::map<std::string, ComputationBlock>>& nb,
                                               std::map<std::string, std::vector<std::string>>& aggregated,
                                               const std::string& left_key) {
    std::vector<std::string>& keys = aggregated[left_key];
    const size_t n_keys = keys.size();
    for (size_t a = 0; a < n_keys; a++) {
        const std::string agg_key = keys[a];
        ComputationBlock& cb = nb[left_key][agg_key];
        for (const auto& vu : cb.v_u) {
            const int v = std::get<0>(vu), u = std::get<1>(vu);
            std::vector<int> nbrs = pc.get_adjacent_nodes()[v];
            nbrs.erase(std::remove(nbrs.begin(), nbrs.end(), u), nbrs.end());
            for (int j : nbrs) {
                const std::string dep_j = pc.get_propagator_key(j, v);
                std::string new_key = "(" + agg_key + std::to_string(cb.n_segment_right);
                std::vector<std::string> subs;
                for (int k : nbrs)
                    if (k != j) subs.push_back(pc.get_propagator_key(k, v) + std::to_string(pc.get_block(k, v).n_segment));
                std::sort(subs.begin(), subs.end());
                for (const auto& s : subs) new_key += s;
                new_key += ")" + pc.get_block(j, v).monomer_type;

                nb[dep_j].erase(v_u_to_right_key[std::make_tuple(j, v)]);
                auto& rights = nb[dep_j];
                if (rights.find(new_key) == rights.end()) {
                    ComputationBlock& nbk = rights[new_key];
                    nbk.monomer_type = pc.get_block(j, v).monomer_type;
                    nbk.n_segment_right = pc.get_block(j, v).n_segment;
                    nbk.n_segment_left = pc.get_block(j, v).n_segment;
                    nbk.v_u.push_back(std::make_tuple(j, v));
                    nbk.n_repeated = agg_key[0] == '[' ? 1 : cb.n_repeated;
                    aggregated[dep_j].push_back(new_key);
                } else {
                    ComputationBlock& nbk = rights[new_key];
                    nbk.v_u.push_back(std::make_tuple(j, v));
                    if (std::get<1>(nbk.v_u[0]) == v) nbk.n_repeated += cb.n_repeated;
                }
            }
        }
    }
}

void PropagatorAnalyzer::update_propagator(const std::string& key, int n_segment, bool is_junction_end) {
    auto it = propagators_.find(key);
    if (it == propagators_.end()) {
        ComputationEdge& e = propagators_[key];
        e.deps = propagator_code::deps_from_key(key);
        e.monomer_type = propagator_code::monomer_type_from_key(key);
        e.max_n_segment = n_segment;
        e.height = propagator_code::height_from_key(key);
        if (is_junction_end) e.junction_ends.insert(n_segment);
        return;
    }
    if (it->second.max_n_segment < n_segment) it->second.max_n_segment = n_segment;
    if (is_junction_end) it->second.junction_ends.insert(n_segment);
}

void PropagatorAnalyzer::display_blocks() const {
    std::cout << "--------------- Blocks ---------------" << std::endl;
    std::cout << "Polymer id, left key:\n\taggregated, (left, right) is_junction, (left, right) n_segment, right key, "
                 "n_repeat, {v, u} list"
              << std::endl;
    std::tuple<int, std::string> last(-1, "");
    for (const auto& item : blocks_) {
        const std::string& kv = std::get<1>(item.first);
        const std::string& ku = std::get<2>(item.first);
        if (last != std::make_tuple(std::get<0>(item.first), kv)) {
            std::cout << std::endl << std::get<0>(item.first) << ", " << kv << ":" << std::endl;
            last = std::make_tuple(std::get<0>(item.first), kv);
        }
        std::cout << "\t " << (ku.find('[') == std::string::npos ? "X, " : "O, ") << "("
                  << (propagator_code::height_from_key(kv) > 0 ? "O, " : "X, ")
                  << (propagator_code::height_from_key(ku) > 0 ? "O), " : "X), ") << "(" << item.second.n_segment_left
                  << ", " << item.second.n_segment_right << "), " << ku << ", " << item.second.n_repeated;
        for (const auto& vu : item.second.v_u) std::cout << ", {" << std::get<0>(vu) << "," << std::get<1>(vu) << "}";
        std::cout << std::endl;
    }
}

void PropagatorAnalyzer::display_propagators() const {
    int total = 0, reduced = 0;
    for (int t : total_segment_numbers_) total += 2 * t;
    std::cout << "--------------- Propagators ---------------" << std::endl;
    std::cout << "Key:\n\theight, aggregated, max_n_segment, # dependencies, junction_ends" << std::endl;
    for (const auto& item : propagators_) {
        const ComputationEdge& e = item.second;
        if (model_name_ == "continuous") {
            reduced += e.max_n_segment;
        } else {
            reduced += e.max_n_segment - 1 + (int)e.junction_ends.size() + (e.deps.empty() ? 0 : 1);
        }
        std::cout << item.first << ":\n\t " << e.height << ", " << (item.first.find('[') == std::string::npos ? "X, " : "O, ")
                  << e.max_n_segment << ", " << e.deps.size() << ", {";
        bool first = true;
        for (int n : e.junction_ends) {
            std::cout << (first ? "" : ", ") << n;
            first = false;
        }
        std::cout << "}, " << std::endl;
    }
    std::cout << "Total number of modified diffusion equation steps (time complexity) to compute propagators: " << total
              << std::endl;
    std::cout << "Total number of steps after optimizing computation : " << reduced << std::endl;
    double percent = 100 * (1.0 - ((double)reduced) / ((double)total));
    std::cout << "Computational cost reduction (higher is better) : " << std::round(percent * 100) / 100 << " %"
              << std::endl;
}

// ------------------------------------------------------------------ scheduler
Scheduler::Scheduler(const PropagatorMap& props, const int n_streams) {
    if (n_streams < 1) LFTS_THROW("The number of streams must be positive.");
    std::vector<int> finish(n_streams, 0);
    std::vector<std::vector<std::string>> queue(n_streams);
    std::map<std::string, std::tuple<int, int, int>, ComparePropagatorKey> lane_start_finish;
    std::map<std::string, int> resolved;

    // height levels in map order
    std::vector<std::vector<std::string>> levels;
    {
        std::vector<std::string> remaining;
        for (const auto& it : props) remaining.push_back(it.first);
        int h = 0;
        while (!remaining.empty()) {
            std::vector<std::string> level;
            for (const auto& k : remaining)
                if (propagator_code::height_from_key(k) == h) level.push_back(k);
            if (!level.empty()) {
                for (const auto& k : level) remaining.erase(std::remove(remaining.begin(), remaining.end(), k), remaining.end());
                levels.push_back(level);
            }
            h++;
        }
    }

    for (const auto& level : levels) {
        std::vector<std::tuple<std::string, int>> ready;
        for (const auto& key : level) {
            int t_ready = 0;
            for (const auto& d : props.at(key).deps) {
                const int t = std::get<1>(lane_start_finish.at(std::get<0>(d))) + std::max(std::get<1>(d), 1);
                if (t_ready == 0 || t_ready < t) t_ready = t;
            }
            resolved[key] = t_ready;
            ready.push_back(std::make_tuple(key, t_ready));
        }
        // libstdc++'s introsort with this comparison decides the order of ties
        std::sort(ready.begin(), ready.end(),
                  [](const std::tuple<std::string, int>& a, const std::tuple<std::string, int>& b) {
                      return std::get<1>(a) < std::get<1>(b);
                  });
        for (const auto& r : ready) {
            int lane = 0, t_min = finish[0];
            for (int s = 1; s < n_streams; s++)
                if (finish[s] < t_min) {
                    lane = s;
                    t_min = finish[s];
                }
            const std::string& key = std::get<0>(r);
            const int len = std::max(props.at(key).max_n_segment, 1);
            const int start = std::max(finish[lane], resolved[key]);
            lane_start_finish[key] = std::make_tuple(lane, start, start + len);
            finish[lane] = start + len;
            queue[lane].push_back(key);
        }
    }

    for (const auto& it : lane_start_finish) sorted_start_.push_back(std::make_tuple(it.first, std::get<1>(it.second)));
    std::sort(sorted_start_.begin(), sorted_start_.end(),
              [](const std::tuple<std::string, int>& a, const std::tuple<std::string, int>& b) {
                  return std::get<1>(a) < std::get<1>(b);
              });
    std::set<int> stamps;
    for (const auto& s : sorted_start_) {
        const int start = std::get<1>(s);
        stamps.insert(start);
        stamps.insert(start + std::max(props.at(std::get<0>(s)).max_n_segment, 1));
    }
    time_stamp_.assign(stamps.begin(), stamps.end());

    std::vector<size_t> pos(n_streams, 0);
    for (size_t i = 0; i + 1 < time_stamp_.size(); i++) {
        std::vector<Job> slice;
        for (int s = 0; s < n_streams; s++) {
            if (pos[s] < queue[s].size() && time_stamp_[i + 1] > std::get<2>(lane_start_finish.at(queue[s][pos[s]])))
                pos[s]++;
            if (pos[s] < queue[s].size()) {
                const std::string& key = queue[s][pos[s]];
                const auto& lsf = lane_start_finish.at(key);
                if (time_stamp_[i] >= std::get<1>(lsf) && time_stamp_[i + 1] <= std::get<2>(lsf)) {
                    if (props.at(key).max_n_segment == 0)
                        slice.push_back(std::make_tuple(key, 0, 0));
                    else
                        slice.push_back(std::make_tuple(key, time_stamp_[i] - std::get<1>(lsf),
                                                        time_stamp_[i + 1] - std::get<1>(lsf)));
                }
            }
        }
        schedule_.push_back(slice);
    }
}

void Scheduler::display(const PropagatorMap& props) const {
    for (const auto& s : sorted_start_) {
        const std::string& key = std::get<0>(s);
        const int start = std::get<1>(s);
        std::cout << key << ":\n\tmax_n_segment: " << props.at(key).max_n_segment << ", start_time: " << start
                  << ", finish_time: " << start + props.at(key).max_n_segment << std::endl;
    }
    for (size_t i = 0; i < schedule_.size(); i++) {
        std::cout << "time: " << time_stamp_[i] + 1 << "-" << time_stamp_[i + 1] << std::endl;
        for (const auto& j : schedule_[i])
            std::cout << "\t" << std::get<0>(j) << ": " << std::get<1>(j) << ", " << std::get<2>(j) << std::endl;
    }
}

}  // namespace lfts
