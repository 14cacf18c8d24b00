// Python module `langevinfts` of the hip platform: the class and method surface
// of the reference binding (langevinfts.cpp:24-463) over the C++ host objects.
// Inputs are float64 arrays of exactly n_grid values (C order); outputs are new
// numpy arrays; C++ errors surface as RuntimeError.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "fts_platform.h"

namespace py = pybind11;
using namespace lfts;

namespace {

using darr = py::array_t<double, py::array::c_style | py::array::forcecast>;

const double* checked(const darr& a, ssize_t n, const std::string& what) {
    if (a.size() != n)
        LFTS_THROW("Size of input " + what + " (" + std::to_string(a.size()) + ") and 'n_grid' (" +
                   std::to_string(n) + ") must match");
    return a.data();
}

std::vector<BlockInput> block_list(const std::vector<std::vector<py::object>>& blocks) {
    std::vector<BlockInput> out;
    for (const auto& b : blocks) {
        if (b.size() != 4) LFTS_THROW("each block is [monomer_type, contour_length, v, u]");
        out.push_back({py::cast<std::string>(b[0]), py::cast<double>(b[1]), py::cast<int>(b[2]), py::cast<int>(b[3])});
    }
    return out;
}

// compute_propagators / compute_statistics argument handling (langevinfts.cpp:169-212)
struct Fields {
    std::vector<darr> keep;
    std::map<std::string, const double*> w, q;
};

Fields fields(PropagatorComputation& obj, const std::map<std::string, py::object>& w_input, const py::object& q_init) {
    Fields f;
    const ssize_t M = obj.get_n_grid();
    for (const auto& it : w_input) {
        f.keep.push_back(darr::ensure(it.second));
        if (!f.keep.back()) LFTS_THROW("w[" + it.first + "] is not convertible to a float64 array");
        f.w[it.first] = checked(f.keep.back(), M, "w[" + it.first + "]");
    }
    if (!q_init.is_none()) {
        for (const auto& it : q_init.cast<std::map<std::string, py::object>>()) {
            f.keep.push_back(darr::ensure(it.second));
            if (!f.keep.back()) LFTS_THROW("q[" + it.first + "] is not convertible to a float64 array");
            f.q[it.first] = checked(f.keep.back(), M, "q[" + it.first + "]");
        }
    }
    return f;
}

}  // namespace

PYBIND11_MODULE(langevinfts, m) {
    m.doc() = "MI355X (gfx950) propagator and concentration path with the langevinfts Python surface";

    py::class_<Block>(m, "Block")
        .def_readonly("monomer_type", &Block::monomer_type)
        .def_readonly("n_segment", &Block::n_segment)
        .def_readonly("contour_length", &Block::contour_length)
        .def_readonly("v", &Block::v)
        .def_readonly("u", &Block::u);

    py::class_<ComputationEdge>(m, "ComputationEdge")
        .def_readonly("max_n_segment", &ComputationEdge::max_n_segment)
        .def_readonly("monomer_type", &ComputationEdge::monomer_type)
        .def_readonly("deps", &ComputationEdge::deps)
        .def_readonly("height", &ComputationEdge::height)
        .def_readonly("junction_ends", &ComputationEdge::junction_ends);

    py::class_<ComputationBlock>(m, "ComputationBlock")
        .def_readonly("monomer_type", &ComputationBlock::monomer_type)
        .def_readonly("n_segment_left", &ComputationBlock::n_segment_left)
        .def_readonly("n_segment_right", &ComputationBlock::n_segment_right)
        .def_readonly("n_repeated", &ComputationBlock::n_repeated)
        .def_readonly("v_u", &ComputationBlock::v_u);

    py::class_<ComputationBox>(m, "ComputationBox")
        .def("get_dim", &ComputationBox::get_dim)
        .def("get_nx", py::overload_cast<>(&ComputationBox::get_nx, py::const_))
        .def("get_nx", py::overload_cast<int>(&ComputationBox::get_nx, py::const_))
        .def("get_lx", py::overload_cast<>(&ComputationBox::get_lx, py::const_))
        .def("get_lx", py::overload_cast<int>(&ComputationBox::get_lx, py::const_))
        .def("get_dx", py::overload_cast<>(&ComputationBox::get_dx, py::const_))
        .def("get_dx", py::overload_cast<int>(&ComputationBox::get_dx, py::const_))
        .def("get_dv", &ComputationBox::get_dv)
        .def("get_n_grid", &ComputationBox::get_n_grid)
        .def("get_volume", &ComputationBox::get_volume)
        .def("set_lx", &ComputationBox::set_lx)
        .def("integral",
             [](ComputationBox& obj, const darr& g) { return obj.integral(checked(g, obj.get_n_grid(), "")); })
        .def("inner_product", [](ComputationBox& obj, const darr& g, const darr& h) {
            return obj.inner_product(checked(g, obj.get_n_grid(), ""), checked(h, obj.get_n_grid(), ""));
        });

    py::class_<Polymer>(m, "Polymer")
        .def("get_alpha", &Polymer::get_alpha)
        .def("get_volume_fraction", &Polymer::get_volume_fraction)
        .def("get_n_blocks", &Polymer::get_n_blocks)
        .def("get_blocks", &Polymer::get_blocks, py::return_value_policy::copy)
        .def("get_block", &Polymer::get_block, py::return_value_policy::copy)
        .def("get_n_segment_total", &Polymer::get_n_segment_total)
        .def("get_n_segment", &Polymer::get_n_segment)
        .def("get_block_index_from_edge", &Polymer::get_block_index_from_edge)
        .def("get_adjacent_nodes", &Polymer::get_adjacent_nodes, py::return_value_policy::copy)
        .def("get_block_indexes", &Polymer::get_block_indexes, py::return_value_policy::copy)
        .def("get_propagator_key", &Polymer::get_propagator_key);

    py::class_<Molecules>(m, "Molecules")
        .def(py::init<std::string, double, std::map<std::string, double>>())
        .def("get_model_name", &Molecules::get_model_name)
        .def("get_ds", &Molecules::get_ds)
        .def("get_bond_lengths", &Molecules::get_bond_lengths)
        .def("get_n_polymer_types", &Molecules::get_n_polymer_types)
        .def("add_polymer",
             [](Molecules& obj, double vf, const std::vector<std::vector<py::object>>& blocks) {
                 obj.add_polymer(vf, block_list(blocks), {});
             })
        .def("add_polymer",
             [](Molecules& obj, double vf, const std::vector<std::vector<py::object>>& blocks,
                const std::map<int, std::string>& q_init) { obj.add_polymer(vf, block_list(blocks), q_init); })
        .def("get_polymer", &Molecules::get_polymer, py::return_value_policy::reference_internal)
        .def("get_n_solvent_types", &Molecules::get_n_solvent_types)
        .def("add_solvent", &Molecules::add_solvent);

    py::class_<PropagatorAnalyzer>(m, "PropagatorAnalyzer")
        // the reference binds this name with the parentheses (langevinfts.cpp:158)
        .def("get_computation_propagators()", &PropagatorAnalyzer::get_computation_propagators,
             py::return_value_policy::copy)
        .def("get_computation_propagators", &PropagatorAnalyzer::get_computation_propagators,
             py::return_value_policy::copy)
        .def("get_computation_propagator", &PropagatorAnalyzer::get_computation_propagator,
             py::return_value_policy::copy)
        .def("get_computation_blocks", &PropagatorAnalyzer::get_computation_blocks, py::return_value_policy::copy)
        .def("get_computation_block", &PropagatorAnalyzer::get_computation_block, py::return_value_policy::copy)
        .def("display_propagators", &PropagatorAnalyzer::display_propagators)
        .def("display_blocks", &PropagatorAnalyzer::display_blocks)
        .def("get_deps_from_key", [](PropagatorAnalyzer&, const std::string& k) { return propagator_code::deps_from_key(k); })
        .def("get_monomer_type_from_key",
             [](PropagatorAnalyzer&, const std::string& k) { return propagator_code::monomer_type_from_key(k); })
        .def("is_aggregated", &PropagatorAnalyzer::is_aggregated)
        .def("get_schedule", [](PropagatorAnalyzer& a, int n_streams) {
            Scheduler s(a.get_computation_propagators(), n_streams);
            return s.get_schedule();
        });

    py::class_<PropagatorComputation>(m, "PropagatorComputation")
        .def("update_laplacian_operator", &PropagatorComputation::update_laplacian_operator)
        .def(
            "compute_propagators",
            [](PropagatorComputation& obj, const std::map<std::string, py::object>& w_input, py::object q_init) {
                Fields f = fields(obj, w_input, q_init);
                obj.compute_propagators(f.w, f.q);
            },
            py::arg("w_input"), py::arg("q_init") = py::none())
        .def("compute_concentrations", &PropagatorComputation::compute_concentrations)
        // (device bytes, pinned host bytes) held by the solver (an addition: shows the reduce-memory placement)
        .def("get_memory_usage", &PropagatorComputation::memory_usage)
        .def(
            "compute_statistics",
            [](PropagatorComputation& obj, const std::map<std::string, py::object>& w_input, py::object q_init) {
                Fields f = fields(obj, w_input, q_init);
                obj.compute_statistics(f.w, f.q);
            },
            py::arg("w_input"), py::arg("q_init") = py::none())
        // device-resident variant: fields given as device addresses (e.g. torch tensor data_ptr())
        .def("compute_propagators_device",
             [](PropagatorComputation& obj, const std::map<std::string, uintptr_t>& w_ptrs,
                const std::map<std::string, uintptr_t>& q_ptrs) {
                 std::map<std::string, const double*> w, q;
                 for (const auto& it : w_ptrs) w[it.first] = reinterpret_cast<const double*>(it.second);
                 for (const auto& it : q_ptrs) q[it.first] = reinterpret_cast<const double*>(it.second);
                 obj.compute_propagators(w, q, true);
             },
             py::arg("w_ptrs"), py::arg("q_ptrs") = std::map<std::string, uintptr_t>{})
        .def("get_total_concentration_device",
             [](PropagatorComputation& obj, const std::string& monomer, uintptr_t out) {
                 obj.get_total_concentration(monomer, reinterpret_cast<double*>(out), true);
             })
        .def("get_total_concentration",
             [](PropagatorComputation& obj, const std::string& monomer) {
                 py::array_t<double> phi(obj.get_n_grid());
                 obj.get_total_concentration(monomer, phi.mutable_data());
                 return phi;
             })
        .def("get_total_concentration",
             [](PropagatorComputation& obj, int polymer, const std::string& monomer) {
                 py::array_t<double> phi(obj.get_n_grid());
                 obj.get_total_concentration(polymer, monomer, phi.mutable_data());
                 return phi;
             })
        .def("get_total_concentration_gce",
             [](PropagatorComputation& obj, double fugacity, int polymer, const std::string& monomer) {
                 py::array_t<double> phi(obj.get_n_grid());
                 obj.get_total_concentration_gce(fugacity, polymer, monomer, phi.mutable_data());
                 return phi;
             })
        .def("get_block_concentration",
             [](PropagatorComputation& obj, int polymer) {
                 const ssize_t nb = obj.get_n_blocks(polymer), M = obj.get_n_grid();
                 py::array_t<double> phi({nb, M});
                 obj.get_block_concentration(polymer, phi.mutable_data());
                 return phi;
             })
        .def("get_total_partition", &PropagatorComputation::get_total_partition)
        .def("get_solvent_partition", &PropagatorComputation::get_solvent_partition)
        .def("get_solvent_concentration",
             [](PropagatorComputation& obj, int s) {
                 py::array_t<double> phi(obj.get_n_grid());
                 obj.get_solvent_concentration(s, phi.mutable_data());
                 return phi;
             })
        .def("get_chain_propagator",
             [](PropagatorComputation& obj, int polymer, int v, int u, int n) {
                 py::array_t<double> q(obj.get_n_grid());
                 obj.get_chain_propagator(q.mutable_data(), polymer, v, u, n);
                 return q;
             })
        .def("compute_stress", &PropagatorComputation::compute_stress)
        .def("get_stress", &PropagatorComputation::get_stress)
        .def("get_stress_gce", &PropagatorComputation::get_stress_gce)
        .def("check_total_partition", &PropagatorComputation::check_total_partition)
        .def("get_schedule", &PropagatorComputation::schedule, py::return_value_policy::copy)
        .def("get_monomer_types", &PropagatorComputation::monomer_types, py::return_value_policy::copy);

    py::class_<AndersonMixing>(m, "AndersonMixing")
        .def("reset_count", &AndersonMixing::reset_count)
        .def("get_n_var", &AndersonMixing::get_n_var)
        .def("calculate_new_fields",
             [](AndersonMixing& obj, const darr& w_current, const darr& w_deriv, double old_err, double err) {
                 const ssize_t n = obj.get_n_var();
                 if (w_current.size() != n)
                     LFTS_THROW("Size of input w_current (" + std::to_string(w_current.size()) + ") and 'n_var' (" +
                                std::to_string(n) + ") must match");
                 if (w_deriv.size() != n)
                     LFTS_THROW("Size of input w_deriv (" + std::to_string(w_deriv.size()) + ") and 'n_var' (" +
                                std::to_string(n) + ") must match");
                 py::array_t<double> w_new(n);
                 obj.calculate_new_fields(w_new.mutable_data(), w_current.data(), w_deriv.data(), old_err, err);
                 return w_new;
             })
        .def("calculate_new_fields_device",
             [](AndersonMixing& obj, uintptr_t w_new, uintptr_t w_current, uintptr_t w_deriv, double old_err,
                double err) {
                 obj.calculate_new_fields(reinterpret_cast<double*>(w_new), reinterpret_cast<const double*>(w_current),
                                          reinterpret_cast<const double*>(w_deriv), old_err, err, true);
             });

    // Device-resident saddle point of the L-FTS driver (examples/fts/lfts.py:631-868);
    // an addition to the reference's surface, used by langevin_fts_amd.lfts.
    py::class_<DeviceSaddlePoint>(m, "DeviceSaddlePoint")
        .def(py::init([](PropagatorComputation* solver, AndersonMixing* am, std::vector<std::string> monomer_types,
                         std::map<std::string, std::map<std::string, double>> random_fraction,
                         std::vector<double> matrix_a, std::vector<int> real_idx, std::vector<int> imag_idx,
                         std::vector<double> real_cw, std::vector<double> real_cp, std::vector<double> real_c0,
                         std::vector<double> imag_cw, std::vector<double> imag_cp, std::vector<double> imag_c0,
                         std::vector<double> dt_scaling, std::vector<double> h_lin, std::vector<double> h_sq,
                         double h_const, std::vector<double> am_params) {
                 FieldTheoryTables t;
                 t.monomer_types = std::move(monomer_types);
                 t.random_fraction = std::move(random_fraction);
                 t.matrix_a = std::move(matrix_a);
                 t.real_idx = std::move(real_idx);
                 t.imag_idx = std::move(imag_idx);
                 t.real_cw = std::move(real_cw);
                 t.real_cp = std::move(real_cp);
                 t.real_c0 = std::move(real_c0);
                 t.imag_cw = std::move(imag_cw);
                 t.imag_cp = std::move(imag_cp);
                 t.imag_c0 = std::move(imag_c0);
                 t.dt_scaling = std::move(dt_scaling);
                 t.h_lin = std::move(h_lin);
                 t.h_sq = std::move(h_sq);
                 t.h_const = h_const;
                 return new DeviceSaddlePoint(solver, am, t, am_params);
             }),
             py::arg("solver"), py::arg("am"), py::arg("monomer_types"), py::arg("random_fraction"),
             py::arg("matrix_a"), py::arg("real_idx"), py::arg("imag_idx"), py::arg("real_cw"), py::arg("real_cp"),
             py::arg("real_c0"), py::arg("imag_cw"), py::arg("imag_cp"), py::arg("imag_c0"), py::arg("dt_scaling"),
             py::arg("h_lin"), py::arg("h_sq"), py::arg("h_const"), py::arg("am_params") = std::vector<double>{},
             py::keep_alive<1, 2>(), py::keep_alive<1, 3>())
        .def("get_n_replicas", &DeviceSaddlePoint::n_replicas)
        .def("get_n_types", &DeviceSaddlePoint::n_types)
        .def("get_n_real", &DeviceSaddlePoint::n_real)
        .def("get_n_imag", &DeviceSaddlePoint::n_imag)
        .def("set_fields",
             [](DeviceSaddlePoint& o, const darr& w) {
                 const ssize_t n = (ssize_t)o.n_replicas() * o.n_types() * o.n_grid();
                 if (w.size() != n)
                     LFTS_THROW("Size of w_aux (" + std::to_string(w.size()) + ") must be " + std::to_string(n));
                 o.set_fields(w.data());
             })
        .def("get_fields",
             [](DeviceSaddlePoint& o) {
                 std::vector<ssize_t> shape{(ssize_t)o.n_types(), (ssize_t)o.n_grid()};
                 if (o.n_replicas() > 1) shape.insert(shape.begin(), (ssize_t)o.n_replicas());
                 py::array_t<double> w(shape);
                 o.get_fields(w.mutable_data());
                 return w;
             })
        .def("get_concentration",
             [](DeviceSaddlePoint& o, int type) {
                 // type r n_types + i: monomer type i of replica r
                 if (type < 0 || type >= o.n_replicas() * o.n_types()) LFTS_THROW("type index out of range");
                 py::array_t<double> phi(o.n_grid());
                 o.get_concentration(type, phi.mutable_data());
                 return phi;
             })
        .def("set_noise_prev",
             [](DeviceSaddlePoint& o, const darr& noise) {
                 const ssize_t n = (ssize_t)o.n_replicas() * o.n_real() * o.n_grid();
                 if (noise.size() != n)
                     LFTS_THROW("Size of normal_noise_prev (" + std::to_string(noise.size()) + ") must be " +
                                std::to_string(n));
                 if (n > 0) o.set_noise_prev(noise.data());
             })
        .def("get_noise_prev",
             [](DeviceSaddlePoint& o) {
                 std::vector<ssize_t> shape{(ssize_t)o.n_real(), (ssize_t)o.n_grid()};
                 if (o.n_replicas() > 1) shape.insert(shape.begin(), (ssize_t)o.n_replicas());
                 py::array_t<double> w(shape);
                 if (o.n_real() > 0) o.get_noise_prev(w.mutable_data());
                 return w;
             })
        .def("backup", &DeviceSaddlePoint::backup)
        .def("restore", &DeviceSaddlePoint::restore)
        .def("restore_replicas",
             [](DeviceSaddlePoint& o, std::vector<int> flags) {
                 if ((int)flags.size() != o.n_replicas()) LFTS_THROW("one flag per replica is required");
                 o.restore_replicas(flags.data());
             })
        .def(
            "find_saddle_point",
            [](DeviceSaddlePoint& o, double tolerance, int max_iter, int verbose) {
                const int row = 3 + o.n_polymers() + o.n_imag();
                const int rows = verbose == 2 ? std::max(max_iter, 1) : 1;
                py::array_t<double> log({(ssize_t)rows, (ssize_t)row});
                std::fill(log.mutable_data(), log.mutable_data() + (size_t)rows * row, 0.0);
                const int R = o.n_replicas();
                std::vector<int> n_iter(R, 0);
                std::vector<double> err(R, 0.0), ham(R, 0.0);
                double* log_rows = R == 1 ? log.mutable_data() : nullptr;
                {
                    // the iteration waits on the device: let other Python threads (the next
                    // step's noise) run meanwhile
                    py::gil_scoped_release nogil;
                    o.find(tolerance, max_iter, R == 1 ? verbose : 0, log_rows, n_iter.data(), err.data(), ham.data());
                }
                if (R == 1) return py::tuple(py::make_tuple(n_iter[0], err[0], ham[0], log));
                return py::tuple(py::make_tuple(n_iter, err, ham, py::none()));
            },
            py::arg("tolerance"), py::arg("max_iter"), py::arg("verbose") = 1)
        .def("langevin_update",
             [](DeviceSaddlePoint& o, const darr& noise, double dt) {
                 const ssize_t n = (ssize_t)o.n_replicas() * o.n_real() * o.n_grid();
                 if (noise.size() != n)
                     LFTS_THROW("Size of the noise (" + std::to_string(noise.size()) + ") must be " +
                                std::to_string(n));
                 if (n > 0) o.langevin_update(noise.data(), dt);
             })
        .def("field_moments",
             [](DeviceSaddlePoint& o) {
                 const int n = o.n_replicas() * o.n_types();
                 std::vector<double> mean(n), mean_sq(n);
                 o.field_moments(mean.data(), mean_sq.data());
                 return py::make_tuple(mean, mean_sq);
             })
        .def("set_low_precision_solver", &DeviceSaddlePoint::set_low_precision_solver, py::arg("solver"),
             py::arg("switch_error"), py::keep_alive<1, 2>())
        .def("set_graphs", &DeviceSaddlePoint::set_graphs, py::arg("on"))
        .def("prepare_guess", &DeviceSaddlePoint::prepare_guess)
        .def("predict_guess", &DeviceSaddlePoint::predict_guess, py::arg("guess"), py::arg("mode") = 0)
        .def(
            "set_residual_guess",
            [](DeviceSaddlePoint& o, LearnedGuess* g, bool record) { o.set_residual_guess(g, record); },
            py::arg("guess"), py::arg("record") = true)
        .def("guess_sample",
             [](DeviceSaddlePoint& o) {
                 const std::vector<double> v = o.guess_sample();
                 py::array_t<double> out({(py::ssize_t)4, (py::ssize_t)o.n_grid()});
                 std::copy(v.begin(), v.end(), out.mutable_data());
                 return out;
             })
        .def("train_guess", &DeviceSaddlePoint::train_guess, py::arg("guess"), py::arg("lr"), py::arg("iters") = 1,
             py::arg("replay") = 1);

    // Learned initial guess of the saddle point: periodic 3D CNN w- -> w+ (an addition to
    // the reference's surface).
    using farr = py::array_t<float, py::array::c_style | py::array::forcecast>;
    py::class_<LearnedGuess>(m, "LearnedGuess")
        .def(py::init([](std::vector<int> nx, py::object in_scale, double out_scale) {
                 std::vector<double> scales;
                 if (py::isinstance<py::float_>(in_scale) || py::isinstance<py::int_>(in_scale))
                     scales.push_back(in_scale.cast<double>());
                 else
                     scales = in_scale.cast<std::vector<double>>();
                 return new LearnedGuess(nx, scales, out_scale);
             }),
             py::arg("nx"), py::arg("in_scale") = 1.0, py::arg("out_scale") = 1.0)
        .def_static("n_params_one_input", []() { return fts_cnn_n_params(); })
        .def("n_params", &LearnedGuess::n_params)
        .def("n_inputs", &LearnedGuess::n_inputs)
        .def("set_params",
             [](LearnedGuess& o, const farr& p) {
                 if (p.size() != o.n_params()) LFTS_THROW("expected " + std::to_string(o.n_params()) + " parameters");
                 o.set_params(p.data());
             })
        .def("get_params",
             [](LearnedGuess& o) {
                 farr p(o.n_params());
                 o.get_params(p.mutable_data());
                 return p;
             })
        .def("predict",
             [](LearnedGuess& o, const darr& w) {
                 if (w.size() != (py::ssize_t)o.n_inputs() * o.n_grid())
                     LFTS_THROW("the inputs must have n_inputs * n_grid values");
                 py::array_t<double> out(o.n_grid());
                 o.predict(w.data(), out.mutable_data());
                 return out;
             })
        .def("gradients",
             [](LearnedGuess& o, const darr& w, const darr& t) {
                 if (w.size() != (py::ssize_t)o.n_inputs() * o.n_grid() || t.size() != o.n_grid())
                     LFTS_THROW("the inputs must have n_inputs * n_grid values and the target n_grid");
                 farr g(o.n_params());
                 const double loss = o.gradients(w.data(), t.data(), g.mutable_data());
                 return py::make_tuple(loss, g);
             })
        .def("train_step", [](LearnedGuess& o, const darr& w, const darr& t, double lr) {
            if (w.size() != (py::ssize_t)o.n_inputs() * o.n_grid() || t.size() != o.n_grid())
                LFTS_THROW("the inputs must have n_inputs * n_grid values and the target n_grid");
            return o.train_step(w.data(), t.data(), lr);
        });

    py::class_<AbstractFactory>(m, "AbstractFactory")
        .def(
            "create_computation_box",
            [](AbstractFactory& obj, std::vector<int> nx, std::vector<double> lx, py::object bc, py::object mask) {
                std::vector<std::string> bcv;
                if (!bc.is_none()) bcv = bc.cast<std::vector<std::string>>();
                int M = 1;
                for (int n : nx) M *= n;
                darr mk;
                const double* mp = nullptr;
                if (!mask.is_none()) {
                    mk = darr::ensure(mask);
                    mp = checked(mk, M, "mask");
                }
                return obj.create_computation_box(nx, lx, bcv, mp);
            },
            py::arg("nx"), py::arg("lx"), py::arg("bc") = py::none(), py::arg("mask") = py::none())
        .def("create_molecules_information", &AbstractFactory::create_molecules_information)
        .def("create_propagator_analyzer", &AbstractFactory::create_propagator_analyzer, py::keep_alive<0, 2>())
        // the solver keeps the box, molecules and analyzer it was built on alive
        .def("create_pseudospectral_solver", &AbstractFactory::create_pseudospectral_solver, py::keep_alive<0, 2>(),
             py::keep_alive<0, 3>(), py::keep_alive<0, 4>())
        .def("create_realspace_solver", &AbstractFactory::create_realspace_solver, py::keep_alive<0, 2>(),
             py::keep_alive<0, 3>(), py::keep_alive<0, 4>())
        // n_replicas independent copies of the chains in one solver (replica L-FTS; an addition)
        .def("create_replica_solver", &AbstractFactory::create_replica_solver, py::keep_alive<0, 2>(),
             py::keep_alive<0, 3>(), py::keep_alive<0, 4>())
        .def("create_anderson_mixing", &AbstractFactory::create_anderson_mixing)
        .def("display_info", &AbstractFactory::display_info)
        .def("get_model_name", &AbstractFactory::get_model_name);

    py::class_<PlatformSelector>(m, "PlatformSelector")
        .def(py::init<>())
        .def_static("avail_platforms", &PlatformSelector::avail_platforms)
        .def_static("create_factory", py::overload_cast<const std::string&>(&PlatformSelector::create_factory))
        .def_static("create_factory", py::overload_cast<const std::string&, bool>(&PlatformSelector::create_factory));

    // Kernel timing of the slab and pencil passes on the device context (an
    // addition to the reference's surface, for roofline reports).
    m.def("set_kernel_timing",
          [](int sample_every) { check_fts(fts_context_set_kernel_timing(hip_context(), sample_every)); });
    m.def("get_kernel_timing", [](const std::string& family) {
        int f;
        if (family == "slab") f = FTS_KERNEL_SLAB;
        else if (family == "pencil") f = FTS_KERNEL_PENCIL;
        else if (family == "real") f = FTS_KERNEL_REAL;
        else LFTS_THROW("kernel family must be \"slab\", \"pencil\" or \"real\"");
        fts_kernel_stats st{};
        check_fts(fts_context_get_kernel_timing(hip_context(), f, &st));
        py::dict d;
        d["milliseconds"] = st.milliseconds;
        d["bytes"] = st.bytes;
        d["launches"] = st.launches;
        return d;
    });
    m.def("scheduler", [](PropagatorAnalyzer& a, int n_streams) {
        Scheduler s(a.get_computation_propagators(), n_streams);
        return s.get_schedule();
    });
}
