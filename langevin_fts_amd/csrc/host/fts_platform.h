// The "hip" platform behind the reference's factory interface
// (AbstractFactory.h:21-62, PlatformSelector.h:12-18, ComputationBox.h:22-76,
// PropagatorComputation.h:21-75, AndersonMixing.h:13-29). Every numerical call
// goes through the C ABI of libfts_hip.so (include/fts_hip.h).
#pragma once

#include <map>
#include <string>
#include <vector>

#include "fts_hip.h"
#include "fts_model.h"

namespace lfts {

// One device context per process (device from LFTS_HIP_DEVICE, default 0).
fts_context* hip_context();
void check_fts(int rc);

class ComputationBox {
public:
    ComputationBox(std::vector<int> nx, std::vector<double> lx, std::vector<std::string> bc, const double* mask);
    ~ComputationBox();
    ComputationBox(const ComputationBox&) = delete;
    ComputationBox& operator=(const ComputationBox&) = delete;

    int get_dim() const { return (int)nx_.size(); }
    std::vector<int> get_nx() const { return nx_; }
    int get_nx(int i) const;
    std::vector<double> get_lx() const { return lx_; }
    double get_lx(int i) const;
    std::vector<double> get_dx() const { return dx_; }
    double get_dx(int i) const;
    double get_dv(int i) const;
    int get_n_grid() const { return n_grid_; }
    double get_volume() const { return volume_; }
    const double* get_mask() const { return mask_.empty() ? nullptr : mask_.data(); }
    bool is_periodic() const;
    std::vector<std::string> get_boundary_conditions() const { return bc_; }
    void set_lx(std::vector<double> new_lx);
    double integral(const double* g) const;
    double inner_product(const double* g, const double* h) const;
    double multi_inner_product(int n_comp, const double* g, const double* h) const;
    void zero_mean(double* g) const;
    fts_grid* handle() const { return grid_; }

private:
    std::vector<int> nx_;
    std::vector<double> lx_, dx_;
    std::vector<std::string> bc_;
    std::vector<double> mask_;
    int n_grid_;
    double volume_;
    fts_grid* grid_ = nullptr;
};

class PropagatorComputation {
public:
    // n_replicas > 1: that many independent copies of the chain set in one solver (replica L-FTS);
    // the concentration and partition getters then report replica 0.
    // realspace: Crank-Nicolson steps under the box's boundary conditions (FP64, continuous chains).
    // reduce_memory: propagator slices in pinned host memory (CudaComputationReduceMemory*)
    PropagatorComputation(ComputationBox* cb, Molecules* molecules, PropagatorAnalyzer* analyzer, int precision,
                          int n_streams, int n_replicas = 1, bool realspace = false, bool reduce_memory = false);
    ~PropagatorComputation();
    PropagatorComputation(const PropagatorComputation&) = delete;
    PropagatorComputation& operator=(const PropagatorComputation&) = delete;

    int get_n_grid() const { return cb_->get_n_grid(); }
    int get_n_blocks(int polymer) { return molecules_->get_polymer(polymer).get_n_blocks(); }
    void update_laplacian_operator();
    void compute_propagators(const std::map<std::string, const double*>& w_input,
                             const std::map<std::string, const double*>& q_init, bool on_device = false);
    void compute_concentrations();
    void compute_statistics(const std::map<std::string, const double*>& w_input,
                            const std::map<std::string, const double*>& q_init) {
        compute_propagators(w_input, q_init);
        compute_concentrations();
    }
    void compute_stress();
    double get_total_partition(int polymer);
    void get_chain_propagator(double* q_out, int polymer, int v, int u, int n);
    void get_total_concentration(const std::string& monomer_type, double* phi, bool on_device = false);
    void get_total_concentration(int polymer, const std::string& monomer_type, double* phi, bool on_device = false);
    void get_total_concentration_gce(double fugacity, int polymer, const std::string& monomer_type, double* phi);
    void get_block_concentration(int polymer, double* phi);
    double get_solvent_partition(int s);
    void get_solvent_concentration(int s, double* phi);
    std::vector<double> get_stress();
    std::vector<double> get_stress_gce(const std::vector<double>& fugacities);
    bool check_total_partition();

    const std::vector<std::string>& monomer_types() const { return monomers_; }
    // blocks (and solvents) whose concentrations make up monomer type `monomer`,
    // each term scaled by `scale` (get_total_concentration without a polymer index)
    void concentration_terms(const std::string& monomer, double scale, std::vector<int>& kinds,
                             std::vector<int>& idx, std::vector<double>& sc) const;
    fts_solver* handle() const { return solver_; }
    // bytes of device memory and of pinned host memory the solver holds
    std::pair<size_t, size_t> memory_usage() const;
    int n_replicas() const { return n_replicas_; }
    int n_blocks_per_replica() const { return (int)block_keys_.size(); }
    Molecules* molecules() const { return molecules_; }
    const std::vector<std::vector<Scheduler::Job>>& schedule() const { return schedule_; }
    int precision() const { return precision_; }
    int n_streams() const { return n_streams_; }

private:
    std::vector<double> dq_dl(int p);
    void sum_blocks(int polymer, const std::string& monomer, bool with_solvents, double scale, double* out,
                    bool on_device);

    ComputationBox* cb_;
    Molecules* molecules_;
    PropagatorAnalyzer* analyzer_;
    int precision_, n_streams_, n_replicas_ = 1;
    bool realspace_ = false;
    std::vector<std::string> monomers_;
    std::map<std::string, int> prop_index_;
    std::vector<std::string> qinit_names_;
    std::vector<BlockKey> block_keys_;
    std::vector<std::vector<Scheduler::Job>> schedule_;
    std::map<std::string, std::vector<double>> last_w_;   // host fields of the last call (check_total_partition)
    fts_solver* solver_ = nullptr;
};

class AndersonMixing {
public:
    // reduce_memory: history rings in pinned host memory (CudaAndersonMixingReduceMemory)
    AndersonMixing(int n_var, int max_hist, double start_error, double mix_min, double mix_init,
                   bool reduce_memory = false);
    ~AndersonMixing();
    AndersonMixing(const AndersonMixing&) = delete;
    AndersonMixing& operator=(const AndersonMixing&) = delete;
    int get_n_var() const { return n_var_; }
    void reset_count();
    void calculate_new_fields(double* w_new, const double* w_current, const double* w_deriv, double old_error_level,
                              double error_level, bool on_device = false);

    fts_am* handle() const { return am_; }

private:
    int n_var_;
    fts_am* am_ = nullptr;
};

// Device-resident saddle-point iteration and Langevin update for the L-FTS driver
// (examples/fts/lfts.py:631-868). The driver hands over the field theory's small
// matrices; the fields stay on the device between calls.
struct FieldTheoryTables {
    std::vector<std::string> monomer_types;                  // S sorted types (lfts.py:280)
    std::map<std::string, std::map<std::string, double>> random_fraction;   // lfts.py:365-397
    std::vector<double> matrix_a;                            // S x S, w = A w_aux
    std::vector<int> real_idx, imag_idx;
    std::vector<double> real_cw, real_cp, real_c0, imag_cw, imag_cp, imag_c0;
    std::vector<double> dt_scaling, h_lin, h_sq;
    double h_const = 0.0;
};

// Periodic 3D CNN of the learned saddle-point guess; weights on the device. One input
// (w- -> w+) or several (in_scales.size() channels, e.g. the state guess's w-, its
// change and the previous w+ -> the change of w+).
class LearnedGuess {
public:
    LearnedGuess(std::vector<int> nx, std::vector<double> in_scales, double out_scale);
    ~LearnedGuess();
    LearnedGuess(const LearnedGuess&) = delete;
    LearnedGuess& operator=(const LearnedGuess&) = delete;
    int n_params() const { return n_params_; }
    int n_inputs() const { return n_in_; }
    int n_grid() const { return M_; }
    void set_params(const float* p);
    void get_params(float* p);
    void predict(const double* w_minus, double* w_plus);
    double gradients(const double* w_minus, const double* target, float* grads);
    double train_step(const double* w_minus, const double* target, double lr);
    fts_cnn* handle() const { return cnn_; }

private:
    int M_, n_in_, n_params_ = 0;
    fts_cnn* cnn_ = nullptr;
};

class DeviceSaddlePoint {
public:
    // am: the Anderson mixing of a single replica; for a replica solver (n_replicas > 1) am is
    // null and am_params = {max_hist, start_error, mix_min, mix_init} configure one per replica
    DeviceSaddlePoint(PropagatorComputation* solver, AndersonMixing* am, const FieldTheoryTables& t,
                      const std::vector<double>& am_params = {});
    ~DeviceSaddlePoint();
    DeviceSaddlePoint(const DeviceSaddlePoint&) = delete;
    DeviceSaddlePoint& operator=(const DeviceSaddlePoint&) = delete;
    int n_types() const { return S_; }
    int n_real() const { return R_; }
    int n_imag() const { return I_; }
    int n_polymers() const { return P_; }
    int n_grid() const { return M_; }
    int n_replicas() const { return NR_; }
    void set_fields(const double* w_aux);
    void get_fields(double* w_aux);
    void get_concentration(int type, double* phi);
    void set_noise_prev(const double* noise);
    void get_noise_prev(double* noise);
    void backup();
    void restore();
    void restore_replicas(const int* flags);
    // returns (iterations, error level, hamiltonian) per replica; log rows as fts_saddle_find
    void find(double tolerance, int max_iter, int verbose, double* log, int* n_iter, double* error, double* ham);
    void langevin_update(const double* noise, double dt);
    void field_moments(double* mean, double* mean_sq);
    void prepare_guess(LearnedGuess& g);
    // mode 0: absolute, 1: delta (one-input networks)
    void predict_guess(LearnedGuess& g, int mode);
    // residual guess: record the first residual of each search; with g (three inputs),
    // apply its correction at the first iteration (g = null: record only or off)
    void set_residual_guess(LearnedGuess* g, bool record);
    // the residual guess's current sample: [w-, first residual, starting w+, target], host
    std::vector<double> guess_sample();
    // iterations on `lo` (an FP32 solver of the same molecules) while the error is at or above switch_error
    void set_low_precision_solver(PropagatorComputation* lo, double switch_error);
    // replicas: saddle iterations and Langevin updates replayed from captured HIP graphs (default on)
    void set_graphs(bool on);
    double train_guess(LearnedGuess& g, double lr, int iters, int replay);

private:
    int S_, R_, I_, P_, M_, NR_ = 1;
    int n_blocks_ = 0;   // per replica, of the saddle's own solver
    fts_saddle* saddle_ = nullptr;
};

class AbstractFactory {
public:
    AbstractFactory(std::string name, int precision, bool reduce_memory_usage)
        : name_(std::move(name)), precision_(precision), reduce_memory_usage_(reduce_memory_usage) {}
    ComputationBox* create_computation_box(std::vector<int> nx, std::vector<double> lx, std::vector<std::string> bc,
                                           const double* mask);
    Molecules* create_molecules_information(const std::string& chain_model, double ds,
                                            const std::map<std::string, double>& bond_lengths);
    PropagatorAnalyzer* create_propagator_analyzer(Molecules* molecules, bool aggregate);
    PropagatorComputation* create_replica_solver(ComputationBox* cb, Molecules* molecules,
                                                 PropagatorAnalyzer* analyzer, int n_replicas);
    PropagatorComputation* create_pseudospectral_solver(ComputationBox* cb, Molecules* molecules,
                                                        PropagatorAnalyzer* analyzer);
    PropagatorComputation* create_realspace_solver(ComputationBox* cb, Molecules* molecules,
                                                   PropagatorAnalyzer* analyzer);
    AndersonMixing* create_anderson_mixing(int n_var, int max_hist, double start_error, double mix_min,
                                           double mix_init);
    std::string get_model_name() const { return model_name_; }
    void display_info() const;

private:
    std::string name_;
    int precision_;
    bool reduce_memory_usage_;
    std::string model_name_;
};

struct PlatformSelector {
    static std::vector<std::string> avail_platforms();
    static AbstractFactory* create_factory(const std::string& platform);
    static AbstractFactory* create_factory(const std::string& platform, bool reduce_memory_usage);
};

}  // namespace lfts
