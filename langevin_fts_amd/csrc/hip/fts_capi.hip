// extern "C" entry points of include/fts_hip.h: argument checks, exception to
// status-code conversion, thread-local error text.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "fts_internal.hpp"

namespace {

thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return FTS_OK;
    } catch (const std::exception& e) {
        g_last_error = e.what();
    } catch (...) {
        g_last_error = "unknown error";
    }
    return FTS_ERROR;
}

void need(const void* p, const char* what) {
    if (!p) throw fts::Error(std::string(what) + " is null");
}

}  // namespace

extern "C" {

const char* fts_last_error(void) { return g_last_error.c_str(); }

int fts_abi_version(void) { return 1; }

int fts_context_create(int device, fts_context** out) {
    return guarded([&] {
        need(out, "out");
        int n = 0;
        FTS_HIP_CHECK(hipGetDeviceCount(&n));
        FTS_REQUIRE(device >= 0 && device < n, "device index out of range");
        FTS_HIP_CHECK(hipSetDevice(device));
        fts::noalloc_scratch_init(device);
        auto* c = new fts_context;
        c->device = device;
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            throw fts::Error("hipStreamCreate failed");
        }
        *out = c;
    });
}

int fts_context_destroy(fts_context* ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int fts_context_set_kernel_timing(fts_context* ctx, int sample_every) {
    return guarded([&] {
        need(ctx, "ctx");
        FTS_REQUIRE(sample_every >= 0, "sample_every must be non-negative");
        FTS_HIP_CHECK(hipSetDevice(ctx->device));
        ctx->timer.enable(sample_every);
    });
}

int fts_context_get_kernel_timing(fts_context* ctx, int family, fts_kernel_stats* out) {
    return guarded([&] {
        need(ctx, "ctx");
        need(out, "out");
        FTS_REQUIRE(family >= 0 && family < FTS_KERNEL_FAMILIES, "kernel family out of range");
        FTS_HIP_CHECK(hipSetDevice(ctx->device));
        *out = ctx->timer.get(family);
    });
}

int fts_context_synchronize(fts_context* ctx) {
    return guarded([&] {
        need(ctx, "ctx");
        FTS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    });
}

int fts_context_stream(fts_context* ctx, void** stream) {
    return guarded([&] {
        need(ctx, "ctx");
        need(stream, "stream");
        *stream = (void*)ctx->stream;
    });
}

int fts_context_info(fts_context* ctx, char* buf, int buflen) {
    return guarded([&] {
        need(ctx, "ctx");
        need(buf, "buf");
        FTS_REQUIRE(buflen > 0, "buflen must be positive");
        hipDeviceProp_t p;
        FTS_HIP_CHECK(hipGetDeviceProperties(&p, ctx->device));
        snprintf(buf, buflen, "%s, %s, %d CUs, %.1f GiB", p.name, p.gcnArchName, p.multiProcessorCount,
                 p.totalGlobalMem / 1073741824.0);
    });
}

int fts_grid_create(fts_context* ctx, int dim, const int* nx, const double* lx, const double* mask, fts_grid** out) {
    return guarded([&] {
        need(ctx, "ctx");
        need(nx, "nx");
        need(lx, "lx");
        need(out, "out");
        FTS_HIP_CHECK(hipSetDevice(ctx->device));
        auto* g = new fts_grid;
        try {
            g->g.reset(new fts::Grid(ctx, dim, nx, lx, mask));
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    });
}

int fts_grid_destroy(fts_grid* g) {
    return guarded([&] { delete g; });
}

int fts_grid_set_lx(fts_grid* g, const double* lx) {
    return guarded([&] {
        need(g, "grid");
        need(lx, "lx");
        g->g->set_lx(lx);
    });
}

int fts_grid_volume(fts_grid* g, double* out) {
    return guarded([&] {
        need(g, "grid");
        need(out, "out");
        *out = g->g->volume;
    });
}

int fts_grid_integral(fts_grid* g, const double* a, int on_device, double* out) {
    return guarded([&] {
        need(g, "grid");
        need(a, "a");
        need(out, "out");
        *out = g->g->integral(a, on_device != 0);
    });
}

int fts_grid_inner_product(fts_grid* g, const double* a, const double* b, int on_device, double* out) {
    return guarded([&] {
        need(g, "grid");
        need(a, "a");
        need(b, "b");
        need(out, "out");
        *out = g->g->inner_product(a, b, on_device != 0);
    });
}

int fts_grid_get_fft_schedule(fts_grid* g, int precision, int* slab_threads, int* pencil_columns, char* note,
                              int notelen) {
    return guarded([&] {
        need(g, "grid");
        FTS_REQUIRE(precision == FTS_FP64 || precision == FTS_FP32, "precision must be FTS_FP64 or FTS_FP32");
        const fts::Grid& gr = *g->g;
        const int p = precision == FTS_FP32 ? 1 : 0;
        if (slab_threads) *slab_threads = gr.wave_local ? gr.slab_bd[p] : 0;
        if (pencil_columns) *pencil_columns = gr.wave_local ? gr.pencil_nl[p] : 0;
        if (note && notelen > 0) {
            const std::string& s = gr.wave_local ? gr.fft_schedule[p] : std::string("runtime kernels (no schedule)");
            const size_t n = std::min(s.size(), (size_t)notelen - 1);
            std::memcpy(note, s.data(), n);
            note[n] = 0;
        }
    });
}

int fts_grid_set_fft_schedule(fts_grid* g, int precision, int slab_threads, int pencil_columns) {
    return guarded([&] {
        need(g, "grid");
        g->g->set_fft_schedule(precision == FTS_FP32 ? 1 : (precision == FTS_FP64 ? 0 : -1), slab_threads,
                               pencil_columns);
    });
}

int fts_grid_fft_forward(fts_grid* g, int precision, const void* d_in, void* d_out, int batch) {
    return guarded([&] {
        need(g, "grid");
        need(d_in, "in");
        need(d_out, "out");
        FTS_REQUIRE(batch >= 1, "batch must be positive");
        if (precision == FTS_FP64)
            g->g->fft_forward<double>(static_cast<const double*>(d_in), static_cast<double2*>(d_out), batch);
        else if (precision == FTS_FP32)
            g->g->fft_forward<float>(static_cast<const float*>(d_in), static_cast<float2*>(d_out), batch);
        else
            throw fts::Error("precision must be FTS_FP64 or FTS_FP32");
        FTS_HIP_CHECK(hipStreamSynchronize(g->g->stream()));
    });
}

int fts_grid_fft_backward(fts_grid* g, int precision, const void* d_in, void* d_out, int batch) {
    return guarded([&] {
        need(g, "grid");
        need(d_in, "in");
        need(d_out, "out");
        FTS_REQUIRE(batch >= 1, "batch must be positive");
        if (precision == FTS_FP64)
            g->g->fft_backward<double>(static_cast<const double2*>(d_in), static_cast<double*>(d_out), batch);
        else if (precision == FTS_FP32)
            g->g->fft_backward<float>(static_cast<const float2*>(d_in), static_cast<float*>(d_out), batch);
        else
            throw fts::Error("precision must be FTS_FP64 or FTS_FP32");
    });
}

int fts_grid_fft_backward_noalloc(fts_grid* g, int precision, const void* d_in, void* d_out, int batch) {
    return guarded([&] {
        need(g, "grid");
        need(d_in, "in");
        need(d_out, "out");
        FTS_REQUIRE(batch >= 1, "batch must be positive");
        if (precision == FTS_FP64)
            g->g->fft_backward_noalloc<double>(static_cast<const double2*>(d_in), static_cast<double*>(d_out), batch);
        else if (precision == FTS_FP32)
            g->g->fft_backward_noalloc<float>(static_cast<const float2*>(d_in), static_cast<float*>(d_out), batch);
        else
            throw fts::Error("precision must be FTS_FP64 or FTS_FP32");
    });
}

int fts_solver_create(fts_grid* g, const fts_plan* plan, int precision, fts_solver** out) {
    return guarded([&] {
        need(g, "grid");
        need(plan, "plan");
        need(out, "out");
        auto* s = new fts_solver;
        try {
            s->s = fts::make_solver(g->g.get(), plan, precision);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int fts_solver_create_reduce_memory(fts_grid* g, const fts_plan* plan, int precision, const int* bc,
                                    fts_solver** out) {
    return guarded([&] {
        need(g, "grid");
        need(plan, "plan");
        need(out, "out");
        auto* s = new fts_solver;
        try {
            s->s = fts::make_solver(g->g.get(), plan, bc ? FTS_FP64 : precision, bc, true);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int fts_solver_get_memory_usage(fts_solver* s, size_t* device_bytes, size_t* host_bytes) {
    return guarded([&] {
        need(s, "solver");
        need(device_bytes, "device_bytes");
        need(host_bytes, "host_bytes");
        s->s->memory_usage(device_bytes, host_bytes);
    });
}

int fts_solver_create_realspace(fts_grid* g, const fts_plan* plan, const int* bc, fts_solver** out) {
    return guarded([&] {
        need(g, "grid");
        need(plan, "plan");
        need(bc, "bc");
        need(out, "out");
        auto* s = new fts_solver;
        try {
            s->s = fts::make_solver(g->g.get(), plan, FTS_FP64, bc);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

namespace {

// R copies of a plan with their own monomer fields: replica r's propagator p is
// r n_props + p, block b is r n_blocks + b, polymer q is r n_polymers + q, solvent
// s is r n_solvents + s, monomer m is r n_monomers + m, q_init k is r n_qinit + k;
// every schedule slice holds the jobs of all replicas.
struct ReplicatedPlan {
    fts_plan plan{};
    std::vector<double> seg, vf, alpha, svf;
    std::vector<int> pm, pmax, pk, pq, pdo, pnd, dp, dn, dr, pjo, pnj, je, bp, bl, br, bnl, bnr, brep, bm, bnvu, smon,
        soff, jp, jf, jt;

    ReplicatedPlan(const fts_plan& p, int R) {
        for (int r = 0; r < R; r++) {
            seg.insert(seg.end(), p.segment_length, p.segment_length + p.n_monomers);
            for (int i = 0; i < p.n_props; i++) {
                pm.push_back(p.prop_monomer[i] + r * p.n_monomers);
                pmax.push_back(p.prop_max_n[i]);
                pk.push_back(p.prop_kind[i]);
                pq.push_back(p.prop_qinit && p.prop_qinit[i] >= 0 ? p.prop_qinit[i] + r * p.n_qinit : -1);
                pdo.push_back((int)dp.size());
                pnd.push_back(p.prop_n_deps[i]);
                for (int d = 0; d < p.prop_n_deps[i]; d++) {
                    const int k = p.prop_dep_off[i] + d;
                    dp.push_back(p.dep_prop[k] + r * p.n_props);
                    dn.push_back(p.dep_n[k]);
                    dr.push_back(p.dep_rep[k]);
                }
                pjo.push_back((int)je.size());
                pnj.push_back(p.prop_n_je[i]);
                for (int d = 0; d < p.prop_n_je[i]; d++) je.push_back(p.je[p.prop_je_off[i] + d]);
            }
            for (int b = 0; b < p.n_blocks; b++) {
                bp.push_back(p.blk_polymer[b] + r * p.n_polymers);
                bl.push_back(p.blk_left[b] + r * p.n_props);
                br.push_back(p.blk_right[b] + r * p.n_props);
                bnl.push_back(p.blk_n_left[b]);
                bnr.push_back(p.blk_n_right[b]);
                brep.push_back(p.blk_rep[b]);
                bm.push_back(p.blk_monomer[b] + r * p.n_monomers);
                bnvu.push_back(p.blk_n_vu[b]);
            }
            vf.insert(vf.end(), p.pol_vf, p.pol_vf + p.n_polymers);
            alpha.insert(alpha.end(), p.pol_alpha, p.pol_alpha + p.n_polymers);
            for (int k = 0; k < p.n_solvents; k++) {
                svf.push_back(p.sol_vf[k]);
                smon.push_back(p.sol_monomer[k] + r * p.n_monomers);
            }
        }
        for (int sl = 0; sl < p.n_slices; sl++) {
            soff.push_back((int)jp.size());
            for (int r = 0; r < R; r++)
                for (int j = p.slice_off[sl]; j < p.slice_off[sl + 1]; j++) {
                    jp.push_back(p.job_prop[j] + r * p.n_props);
                    jf.push_back(p.job_from[j]);
                    jt.push_back(p.job_to[j]);
                }
        }
        soff.push_back((int)jp.size());
        plan.model = p.model;
        plan.ds = p.ds;
        plan.n_monomers = p.n_monomers * R;
        plan.segment_length = seg.data();
        plan.n_props = p.n_props * R;
        plan.prop_monomer = pm.data();
        plan.prop_max_n = pmax.data();
        plan.prop_kind = pk.data();
        plan.prop_qinit = pq.data();
        plan.prop_dep_off = pdo.data();
        plan.prop_n_deps = pnd.data();
        plan.dep_prop = dp.empty() ? nullptr : dp.data();
        plan.dep_n = dn.empty() ? nullptr : dn.data();
        plan.dep_rep = dr.empty() ? nullptr : dr.data();
        plan.prop_je_off = pjo.data();
        plan.prop_n_je = pnj.data();
        plan.je = je.empty() ? nullptr : je.data();
        plan.n_blocks = p.n_blocks * R;
        plan.blk_polymer = bp.data();
        plan.blk_left = bl.data();
        plan.blk_right = br.data();
        plan.blk_n_left = bnl.data();
        plan.blk_n_right = bnr.data();
        plan.blk_rep = brep.data();
        plan.blk_monomer = bm.data();
        plan.blk_n_vu = bnvu.data();
        plan.n_polymers = p.n_polymers * R;
        plan.pol_vf = vf.data();
        plan.pol_alpha = alpha.data();
        plan.n_solvents = p.n_solvents * R;
        plan.sol_vf = svf.empty() ? nullptr : svf.data();
        plan.sol_monomer = smon.empty() ? nullptr : smon.data();
        plan.n_qinit = p.n_qinit * R;
        plan.n_slices = p.n_slices;
        plan.slice_off = soff.data();
        plan.job_prop = jp.data();
        plan.job_from = jf.data();
        plan.job_to = jt.data();
    }
};

}  // namespace

int fts_solver_create_replicas(fts_grid* g, const fts_plan* plan, int precision, int n_replicas, fts_solver** out) {
    return guarded([&] {
        need(g, "grid");
        need(plan, "plan");
        need(out, "out");
        FTS_REQUIRE(n_replicas >= 1, "n_replicas must be positive");
        ReplicatedPlan rp(*plan, n_replicas);
        auto* s = new fts_solver;
        try {
            s->s = fts::make_solver(g->g.get(), &rp.plan, precision);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int fts_solver_destroy(fts_solver* s) {
    return guarded([&] { delete s; });
}

int fts_solver_update_laplacian(fts_solver* s) {
    return guarded([&] {
        need(s, "solver");
        s->s->update_laplacian();
    });
}

int fts_solver_compute_propagators(fts_solver* s, const double* const* w, const double* const* q_init, int on_device) {
    return guarded([&] {
        need(s, "solver");
        need(w, "w");
        s->s->compute_propagators(w, q_init, on_device != 0);
    });
}

int fts_solver_compute_concentrations(fts_solver* s) {
    return guarded([&] {
        need(s, "solver");
        s->s->compute_concentrations();
    });
}

int fts_solver_set_graphs(fts_solver* s, int enable) {
    return guarded([&] {
        need(s, "solver");
        s->s->set_graphs(enable != 0);
    });
}

int fts_solver_compute_stress(fts_solver* s) {
    return guarded([&] {
        need(s, "solver");
        s->s->compute_stress();
    });
}

int fts_solver_get_partition(fts_solver* s, int polymer, double* out) {
    return guarded([&] {
        need(s, "solver");
        need(out, "out");
        *out = s->s->partition(polymer);
    });
}

int fts_solver_get_solvent_partition(fts_solver* s, int solvent, double* out) {
    return guarded([&] {
        need(s, "solver");
        need(out, "out");
        *out = s->s->solvent_partition(solvent);
    });
}

int fts_solver_get_dq_dl(fts_solver* s, int polymer, double* out) {
    return guarded([&] {
        need(s, "solver");
        need(out, "out");
        s->s->dq_dl(polymer, out);
    });
}

int fts_solver_sum_concentrations(fts_solver* s, int n_terms, const int* src_kind, const int* src_index,
                                  const double* scale, double* out, int on_device) {
    return guarded([&] {
        need(s, "solver");
        need(out, "out");
        FTS_REQUIRE(n_terms >= 0, "n_terms must be non-negative");
        if (n_terms) {
            need(src_kind, "src_kind");
            need(src_index, "src_index");
        }
        s->s->sum_concentrations(n_terms, src_kind, src_index, scale, out, on_device != 0);
    });
}

int fts_solver_get_propagator(fts_solver* s, int prop, int n, int half, double* out, int on_device) {
    return guarded([&] {
        need(s, "solver");
        need(out, "out");
        s->s->propagator(prop, n, half, out, on_device != 0);
    });
}

int fts_am_create_reduce_memory(fts_context* ctx, int n_var, int max_hist, double start_error, double mix_min,
                                double mix_init, fts_am** out) {
    return guarded([&] {
        need(ctx, "ctx");
        need(out, "out");
        auto* a = new fts_am;
        try {
            a->a.reset(new fts::AndersonMixingDevice(ctx, n_var, max_hist, start_error, mix_min, mix_init, true));
        } catch (...) {
            delete a;
            throw;
        }
        *out = a;
    });
}

int fts_am_create(fts_context* ctx, int n_var, int max_hist, double start_error, double mix_min, double mix_init,
                  fts_am** out) {
    return guarded([&] {
        need(ctx, "ctx");
        need(out, "out");
        auto* a = new fts_am;
        try {
            a->a.reset(new fts::AndersonMixingDevice(ctx, n_var, max_hist, start_error, mix_min, mix_init));
        } catch (...) {
            delete a;
            throw;
        }
        *out = a;
    });
}

int fts_am_destroy(fts_am* am) {
    return guarded([&] { delete am; });
}

int fts_am_reset(fts_am* am) {
    return guarded([&] {
        need(am, "am");
        am->a->reset();
    });
}

int fts_am_calculate_new_fields(fts_am* am, double* w_new, const double* w_current, const double* w_deriv,
                                double old_error_level, double error_level, int on_device) {
    return guarded([&] {
        need(am, "am");
        need(w_new, "w_new");
        need(w_current, "w_current");
        need(w_deriv, "w_deriv");
        am->a->calculate_new_fields(w_new, w_current, w_deriv, old_error_level, error_level, on_device != 0);
    });
}

int fts_saddle_create(fts_context* ctx, fts_solver* solver, fts_am* am, int n_grid, const fts_field_theory* ft,
                      fts_saddle** out) {
    return guarded([&] {
        need(ctx, "ctx");
        need(solver, "solver");
        need(am, "am");
        need(ft, "field theory");
        need(out, "out");
        const int S = ft->n_types, R = ft->n_real, I = ft->n_imag;
        FTS_REQUIRE(S >= 1 && S <= FTS_MAX_TYPES, "n_types must be in [1, 8]");
        need(ft->imag_idx, "imag_idx");
        need(ft->w_map, "w_map");
        need(ft->phi_term_off, "phi_term_off");
        if (ft->phi_term_off[S] > 0) {
            need(ft->phi_term_kind, "phi_term_kind");
            need(ft->phi_term_index, "phi_term_index");
            need(ft->phi_term_scale, "phi_term_scale");
        }
        if (R > 0) {
            need(ft->real_idx, "real_idx");
            need(ft->real_cw, "real_cw");
            need(ft->real_cp, "real_cp");
            need(ft->real_c0, "real_c0");
        }
        FTS_REQUIRE(I >= 1, "n_imag must be positive");
        need(ft->imag_cw, "imag_cw");
        need(ft->imag_cp, "imag_cp");
        need(ft->imag_c0, "imag_c0");
        need(ft->dt_scaling, "dt_scaling");
        need(ft->h_lin, "h_lin");
        need(ft->h_sq, "h_sq");
        need(ft->polymer_weight, "polymer_weight");
        auto* s = new fts_saddle;
        try {
            s->s = std::make_unique<fts::SaddlePoint>(ctx, solver->s.get(),
                                                      std::vector<fts::AndersonMixingDevice*>{am->a.get()}, n_grid,
                                                      *ft);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int fts_saddle_create_replicas(fts_context* ctx, fts_solver* solver, int n_replicas, int n_blocks, int n_solvents,
                               int max_hist, double start_error, double mix_min, double mix_init, int n_grid,
                               const fts_field_theory* ft, fts_saddle** out) {
    return guarded([&] {
        need(ctx, "ctx");
        need(solver, "solver");
        need(ft, "field theory");
        need(out, "out");
        FTS_REQUIRE(n_replicas >= 1, "n_replicas must be positive");
        FTS_REQUIRE(ft->n_types >= 1 && ft->n_types <= FTS_MAX_TYPES, "n_types must be in [1, 8]");
        need(ft->imag_idx, "imag_idx");
        need(ft->w_map, "w_map");
        need(ft->phi_term_off, "phi_term_off");
        auto* s = new fts_saddle;
        try {
            std::vector<fts::AndersonMixingDevice*> ams;
            for (int r = 0; r < n_replicas; r++) {
                s->owned_ams.push_back(std::make_unique<fts::AndersonMixingDevice>(
                    ctx, ft->n_imag * n_grid, max_hist, start_error, mix_min, mix_init));
                ams.push_back(s->owned_ams.back().get());
            }
            s->s = std::make_unique<fts::SaddlePoint>(ctx, solver->s.get(), ams, n_grid, *ft, n_blocks, n_solvents);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int fts_saddle_destroy(fts_saddle* s) {
    return guarded([&] { delete s; });
}

int fts_saddle_set_fields(fts_saddle* s, const double* w_aux, int on_device) {
    return guarded([&] {
        need(s, "saddle");
        need(w_aux, "w_aux");
        s->s->set_fields(w_aux, on_device != 0);
    });
}

int fts_saddle_get_fields(fts_saddle* s, double* w_aux, int on_device) {
    return guarded([&] {
        need(s, "saddle");
        need(w_aux, "w_aux");
        s->s->get_fields(w_aux, on_device != 0);
    });
}

int fts_saddle_get_concentration(fts_saddle* s, int type, double* out, int on_device) {
    return guarded([&] {
        need(s, "saddle");
        need(out, "out");
        s->s->get_concentration(type, out, on_device != 0);
    });
}

int fts_saddle_set_noise(fts_saddle* s, const double* noise_prev, int on_device) {
    return guarded([&] {
        need(s, "saddle");
        need(noise_prev, "noise_prev");
        s->s->set_noise(noise_prev, on_device != 0);
    });
}

int fts_saddle_get_noise(fts_saddle* s, double* noise_prev, int on_device) {
    return guarded([&] {
        need(s, "saddle");
        need(noise_prev, "noise_prev");
        s->s->get_noise(noise_prev, on_device != 0);
    });
}

int fts_saddle_backup(fts_saddle* s) {
    return guarded([&] {
        need(s, "saddle");
        s->s->backup();
    });
}

int fts_saddle_restore(fts_saddle* s) {
    return guarded([&] {
        need(s, "saddle");
        s->s->restore();
    });
}

int fts_saddle_restore_replicas(fts_saddle* s, const int* replicas) {
    return guarded([&] {
        need(s, "saddle");
        need(replicas, "replicas");
        s->s->restore(replicas);
    });
}

int fts_saddle_find(fts_saddle* s, double tolerance, int max_iter, int verbose, double* log, int* n_iter,
                    double* error_level, double* hamiltonian) {
    return guarded([&] {
        need(s, "saddle");
        s->s->find(tolerance, max_iter, verbose, log, n_iter, error_level, hamiltonian);
    });
}

int fts_saddle_langevin_update(fts_saddle* s, const double* noise, int on_device, double dt) {
    return guarded([&] {
        need(s, "saddle");
        need(noise, "noise");
        s->s->langevin_update(noise, on_device != 0, dt);
    });
}

int fts_saddle_field_moments(fts_saddle* s, double* mean, double* mean_sq) {
    return guarded([&] {
        need(s, "saddle");
        need(mean, "mean");
        need(mean_sq, "mean_sq");
        s->s->field_moments(mean, mean_sq);
    });
}

int fts_cnn_n_params(void) { return fts::cnn_n_params(1); }

int fts_cnn_create_inputs(fts_context* ctx, const int* nx, int n_in, const double* in_scales, double out_scale,
                          fts_cnn** out) {
    return guarded([&] {
        need(ctx, "ctx");
        need(nx, "nx");
        need(in_scales, "in_scales");
        need(out, "out");
        FTS_HIP_CHECK(hipSetDevice(ctx->device));
        auto* c = new fts_cnn;
        try {
            c->c = std::make_unique<fts::Cnn>(ctx, nx[0], nx[1], nx[2], n_in, in_scales, out_scale);
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
    });
}

int fts_cnn_create(fts_context* ctx, const int* nx, double in_scale, double out_scale, fts_cnn** out) {
    return fts_cnn_create_inputs(ctx, nx, 1, &in_scale, out_scale, out);
}

int fts_cnn_param_count(fts_cnn* c, int* n) {
    return guarded([&] {
        need(c, "cnn");
        need(n, "n");
        *n = c->c->n_params();
    });
}

int fts_cnn_destroy(fts_cnn* c) {
    return guarded([&] { delete c; });
}

int fts_cnn_set_params(fts_cnn* c, const float* params) {
    return guarded([&] {
        need(c, "cnn");
        need(params, "params");
        c->c->set_params(params);
    });
}

int fts_cnn_get_params(fts_cnn* c, float* params) {
    return guarded([&] {
        need(c, "cnn");
        need(params, "params");
        c->c->get_params(params);
    });
}

int fts_cnn_predict(fts_cnn* c, const double* w_minus, double* w_plus, int on_device) {
    return guarded([&] {
        need(c, "cnn");
        need(w_minus, "w_minus");
        need(w_plus, "w_plus");
        c->c->predict(w_minus, w_plus, on_device != 0);
    });
}

int fts_cnn_gradients(fts_cnn* c, const double* w_minus, const double* w_plus_target, int on_device, double* loss,
                      float* grads) {
    return guarded([&] {
        need(c, "cnn");
        need(w_minus, "w_minus");
        need(w_plus_target, "w_plus_target");
        const double l = c->c->gradients(w_minus, w_plus_target, on_device != 0);
        if (loss) *loss = l;
        if (grads) c->c->get_grads(grads);
    });
}

int fts_cnn_train_step(fts_cnn* c, const double* w_minus, const double* w_plus_target, int on_device, double lr,
                       double* loss) {
    return guarded([&] {
        need(c, "cnn");
        need(w_minus, "w_minus");
        need(w_plus_target, "w_plus_target");
        FTS_REQUIRE(lr > 0, "learning rate must be positive");
        const double l = c->c->train_step(w_minus, w_plus_target, on_device != 0, lr);
        if (loss) *loss = l;
    });
}

int fts_saddle_set_graphs(fts_saddle* s, int on) {
    return guarded([&] {
        need(s, "saddle");
        s->s->set_graphs(on != 0);
    });
}

int fts_saddle_set_low_precision(fts_saddle* s, fts_solver* lo, double switch_error) {
    return guarded([&] {
        need(s, "saddle");
        FTS_REQUIRE(switch_error >= 0.0, "switch_error must be non-negative");
        s->s->set_low_precision(lo ? lo->s.get() : nullptr, switch_error);
    });
}

int fts_saddle_prepare_guess(fts_saddle* s, fts_cnn* c) {
    return guarded([&] {
        need(s, "saddle");
        need(c, "cnn");
        s->s->prepare_guess(*c->c);
    });
}

int fts_saddle_predict_guess(fts_saddle* s, fts_cnn* c, int mode) {
    return guarded([&] {
        need(s, "saddle");
        need(c, "cnn");
        FTS_REQUIRE(mode == 0 || mode == 1, "guess mode must be 0 (absolute) or 1 (delta)");
        s->s->predict_guess(*c->c, mode);
    });
}

int fts_saddle_set_residual_guess(fts_saddle* s, fts_cnn* c, int record) {
    return guarded([&] {
        need(s, "saddle");
        s->s->set_residual_guess(c ? c->c.get() : nullptr, record != 0);
    });
}

int fts_saddle_get_guess_sample(fts_saddle* s, double* out) {
    return guarded([&] {
        need(s, "saddle");
        need(out, "out");
        s->s->guess_sample(out);
    });
}

int fts_saddle_train_guess(fts_saddle* s, fts_cnn* c, double lr, int iters, int replay, double* loss) {
    return guarded([&] {
        need(s, "saddle");
        need(c, "cnn");
        FTS_REQUIRE(lr > 0, "learning rate must be positive");
        const double l = s->s->train_guess(*c->c, lr, iters, replay);
        if (loss) *loss = l;
    });
}

}  // extern "C"
