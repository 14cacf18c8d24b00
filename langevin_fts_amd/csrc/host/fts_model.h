// Chain model of the propagator path: polymers as trees of blocks, the canonical
// propagator keys, deduplication/aggregation of propagators across a mixture and
// the time-sliced schedule of their computation. Interfaces and results follow
// the reference's common layer (Polymer.h:13-71, Molecules.h:16-69,
// PropagatorCode.h:13-42, PropagatorAnalyzer.h:16-101, Scheduler.h:16-36); the
// analyzer and scheduler output must be identical to the reference's, string for
// string and tuple for tuple.
#pragma once

#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace lfts {

// Error carrying file/line/function like the reference's Exception (Exception.h:7-16).
class Exception : public std::runtime_error {
public:
    Exception(const std::string& msg, const char* file, int line, const char* func)
        : std::runtime_error("File: \"" + std::string(file) + "\", line: " + std::to_string(line) + ", function <" +
                             func + ">\n\t" + msg) {}
    Exception(const std::string& msg, const char* file, const char* func)
        : std::runtime_error("File: \"" + std::string(file) + "\", function <" + func + ">\n\t" + msg) {}
};
#define LFTS_THROW(msg) throw ::lfts::Exception((msg), __FILE__, __LINE__, __func__)
#define LFTS_RETHROW(msg) throw ::lfts::Exception((msg), __FILE__, __func__)

struct BlockInput {
    std::string monomer_type;
    double contour_length;
    int v;
    int u;
};

struct Block {
    std::string monomer_type;
    int n_segment;
    double contour_length;
    int v;
    int u;
};

class Polymer {
public:
    Polymer(double ds, const std::map<std::string, double>& bond_lengths, double volume_fraction,
            const std::vector<BlockInput>& block_inputs, const std::map<int, std::string>& chain_end_to_q_init = {});
    double get_alpha() const { return alpha_; }
    double get_volume_fraction() const { return volume_fraction_; }
    int get_n_blocks() const { return (int)blocks_.size(); }
    std::vector<Block>& get_blocks() { return blocks_; }
    Block& get_block(int v, int u);
    int get_n_segment_total() const;
    int get_n_segment(int idx) const { return blocks_.at(idx).n_segment; }
    int get_block_index_from_edge(int v, int u) const;
    std::map<int, std::vector<int>>& get_adjacent_nodes() { return adjacent_; }
    std::map<std::pair<int, int>, int>& get_block_indexes() { return edge_to_block_; }
    std::string get_propagator_key(int v, int u) const;
    const std::map<int, std::string>& get_chain_end_to_q_init() const { return q_init_; }

private:
    double alpha_ = 0.0;
    double volume_fraction_ = 0.0;
    std::vector<Block> blocks_;
    std::map<int, std::vector<int>> adjacent_;
    std::map<std::pair<int, int>, int> edge_to_block_;
    std::map<std::pair<int, int>, std::string> edge_key_;
    std::map<int, std::string> q_init_;
};

class Molecules {
public:
    Molecules(std::string model_name, double ds, const std::map<std::string, double>& bond_lengths);
    std::string get_model_name() const { return model_name_; }
    double get_ds() const { return ds_; }
    const std::map<std::string, double>& get_bond_lengths() const { return bond_lengths_; }
    void add_polymer(double volume_fraction, const std::vector<BlockInput>& blocks,
                     const std::map<int, std::string>& chain_end_to_q_init = {});
    void add_solvent(double volume_fraction, const std::string& monomer_type);
    int get_n_polymer_types() const { return (int)polymers_.size(); }
    Polymer& get_polymer(int p);
    int get_n_solvent_types() const { return (int)solvents_.size(); }
    std::tuple<double, std::string> get_solvent(int s) const { return solvents_.at(s); }

private:
    std::string model_name_;
    double ds_;
    std::map<std::string, double> bond_lengths_;
    std::vector<Polymer> polymers_;
    std::vector<std::tuple<double, std::string>> solvents_;
};

// Canonical edge codes and key parsing (PropagatorCode.cpp:19-290).
namespace propagator_code {
std::vector<std::tuple<int, int, std::string>> generate_codes(Polymer& pc, const std::map<int, std::string>& q_init);
std::string key_from_code(const std::string& code);
std::vector<std::tuple<std::string, int, int>> deps_from_key(const std::string& key);
std::string monomer_type_from_key(const std::string& key);
std::string q_input_name_from_key(const std::string& key);
int height_from_key(const std::string& key);
}  // namespace propagator_code

// Height ascending, then string descending (PropagatorAnalyzer.cpp:15-27).
struct ComparePropagatorKey {
    bool operator()(const std::string& a, const std::string& b) const;
};

struct ComputationEdge {
    int max_n_segment = 0;
    std::string monomer_type;
    std::vector<std::tuple<std::string, int, int>> deps;
    int height = 0;
    std::set<int> junction_ends;
};

struct ComputationBlock {
    std::string monomer_type;
    int n_segment_left = 0;
    int n_segment_right = 0;
    int n_repeated = 0;
    std::vector<std::tuple<int, int>> v_u;
};

using PropagatorMap = std::map<std::string, ComputationEdge, ComparePropagatorKey>;
using BlockKey = std::tuple<int, std::string, std::string>;
using BlockMap = std::map<BlockKey, ComputationBlock>;

class PropagatorAnalyzer {
public:
    PropagatorAnalyzer(Molecules* molecules, bool aggregate_propagator_computation);
    bool is_aggregated() const { return aggregate_; }
    int get_n_computation_propagator_codes() const { return (int)propagators_.size(); }
    PropagatorMap& get_computation_propagators() { return propagators_; }
    ComputationEdge& get_computation_propagator(const std::string& key);
    BlockMap& get_computation_blocks() { return blocks_; }
    ComputationBlock& get_computation_block(const BlockKey& key);
    std::string get_model_name() const { return model_name_; }
    void display_propagators() const;
    void display_blocks() const;

    // aggregation of the right keys sharing one left key
    static std::map<std::string, ComputationBlock> aggregate(std::map<std::string, ComputationBlock> set_I,
                                                             int minimum_n_segment);

private:
    void add_polymer(Polymer& pc, int polymer_id);
    void substitute_right_keys(Polymer& pc, std::map<std::tuple<int, int>, std::string>& v_u_to_right_key,
                               std::map<std::string, std::map<std::string, ComputationBlock>>& new_blocks,
                               std::map<std::string, std::vector<std::string>>& aggregated,
                               const std::string& left_key);
    void update_propagator(const std::string& key, int n_segment, bool is_junction_end);

    std::string model_name_;
    bool aggregate_;
    BlockMap blocks_;
    PropagatorMap propagators_;
    std::vector<int> total_segment_numbers_;
};

// Greedy list scheduling onto n_streams lanes (Scheduler.cpp:11-190).
class Scheduler {
public:
    using Job = std::tuple<std::string, int, int>;
    Scheduler(const PropagatorMap& propagators, int n_streams);
    const std::vector<std::vector<Job>>& get_schedule() const { return schedule_; }
    void display(const PropagatorMap& propagators) const;

private:
    std::vector<std::vector<Job>> schedule_;
    std::vector<std::tuple<std::string, int>> sorted_start_;
    std::vector<int> time_stamp_;
};

}  // namespace lfts
