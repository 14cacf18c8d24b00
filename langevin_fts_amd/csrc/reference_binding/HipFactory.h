// The hip platform as a factory of the reference's own C++ tree: the classes a
// maintainer adds under src/platforms/hip/ of yongdd/Langevin_FTS so that
// PlatformSelector can hand out "hip" (PlatformSelector.cpp:16-54). They derive from
// the reference's abstract classes (AbstractFactory.h:21-62, PropagatorComputation.h,
// ComputationBox.h, AndersonMixing.h, Array.h) and call libfts_hip.so through the C
// ABI of include/fts_hip.h; nothing here depends on this repository's own host layer.
#ifndef HIP_FACTORY_H_
#define HIP_FACTORY_H_

#include <map>
#include <string>
#include <vector>

#include "AbstractFactory.h"
#include "fts_hip.h"

// the process's HIP context (device 0), made on first use
fts_context* hip_platform_context();

class HipArray : public Array {
    std::vector<double> data_;

public:
    explicit HipArray(unsigned int size);
    HipArray(double* data, unsigned int size);
    void add(const Array& a, const Array& b) override;
    void subtract(const Array& a, const Array& b) override;
    void multiply(const Array& a, const Array& b) override;
    void divide(const Array& a, const Array& b) override;
    void linear_scaling(const Array& src, const double a, const double b) override;
    void operator=(const Array& src) override;
    void set_data(double* data, unsigned int size) override;
    std::vector<double> to_vector() const override;
    double* get_ptr() const override;
    unsigned int get_size() const override;
    double operator[](unsigned int i) const override;
};

class HipComputationBox : public ComputationBox {
    fts_grid* g_ = nullptr;

public:
    HipComputationBox(std::vector<int> nx, std::vector<double> lx, std::vector<std::string> bc,
                      const double* mask = nullptr);
    ~HipComputationBox() override;
    void set_lx(std::vector<double> new_lx) override;
    double integral_device(const double* g) override;
    double inner_product_device(const double* g, const double* h) override;
    double inner_product_inverse_weight_device(const double* g, const double* h, const double* w) override;
    double multi_inner_product_device(int n_comp, const double* g, const double* h) override;
    void zero_mean_device(double* g) override;
    fts_grid* handle() const { return g_; }
};

// Pseudo-spectral (or, with realspace, Crank-Nicolson) propagators on the device, FP64
class HipPropagatorComputation : public PropagatorComputation {
    HipComputationBox* hcb_;
    fts_solver* s_ = nullptr;
    std::vector<std::string> monomers_;          // monomer types in name order
    std::vector<std::string> qinit_names_;
    std::map<std::string, int> prop_index_;      // propagator key -> index (ComparePropagatorKey order)
    std::vector<std::tuple<int, std::string, std::string>> block_keys_;
    std::map<std::string, std::vector<double>> last_w_;
    void partitions();
    void sum_blocks(int polymer, const std::string& monomer, bool with_solvents, double scale, double* phi);

public:
    HipPropagatorComputation(HipComputationBox* cb, Molecules* molecules, PropagatorAnalyzer* analyzer,
                             bool realspace);
    ~HipPropagatorComputation() override;
    void update_laplacian_operator() override;
    void compute_propagators(std::map<std::string, const double*> w_block,
                             std::map<std::string, const double*> q_init = {}) override;
    void compute_concentrations() override;
    void compute_statistics(std::map<std::string, const double*> w_block,
                            std::map<std::string, const double*> q_init = {}) override;
    void compute_stress() override;
    double get_total_partition(int polymer) override;
    void get_chain_propagator(double* q_out, int polymer, int v, int u, int n) override;
    void get_total_concentration(std::string monomer_type, double* phi) override;
    void get_total_concentration(int polymer, std::string monomer_type, double* phi) override;
    void get_block_concentration(int polymer, double* phi) override;
    double get_solvent_partition(int s) override;
    void get_solvent_concentration(int s, double* phi) override;
    void get_total_concentration_gce(double fugacity, int polymer, std::string monomer_type, double* phi) override;
    bool check_total_partition() override;
};

class HipAndersonMixing : public AndersonMixing {
    fts_am* am_ = nullptr;

public:
    HipAndersonMixing(int n_var, int max_hist, double start_error, double mix_min, double mix_init);
    ~HipAndersonMixing() override;
    void reset_count() override;
    void calculate_new_fields(double* w_new, double* w_current, double* w_deriv, double old_error_level,
                              double error_level) override;
};

class HipFactory : public AbstractFactory {
public:
    HipFactory(std::string chain_model, bool reduce_memory_usage);
    Array* create_array(unsigned int size) override;
    Array* create_array(double* data, unsigned int size) override;
    ComputationBox* create_computation_box(std::vector<int> nx, std::vector<double> lx, std::vector<std::string> bc,
                                           const double* mask = nullptr) override;
    Molecules* create_molecules_information(std::string chain_model, double ds,
                                            std::map<std::string, double> bond_lengths) override;
    PropagatorComputation* create_pseudospectral_solver(ComputationBox* cb, Molecules* molecules,
                                                        PropagatorAnalyzer* propagator_analyzer) override;
    PropagatorComputation* create_realspace_solver(ComputationBox* cb, Molecules* molecules,
                                                   PropagatorAnalyzer* propagator_analyzer) override;
    AndersonMixing* create_anderson_mixing(int n_var, int max_hist, double start_error, double mix_min,
                                           double mix_init) override;
    void display_info() override;
};

#endif
