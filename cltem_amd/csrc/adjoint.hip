// Adjoint slice loop with sqrt(N) checkpointing: gradients of an exit-intensity loss with
// respect to the atom positions of one configuration.  clTEM has no counterpart; the
// forward it differentiates is doMultiSliceStep (microscope/simulationgeneral.cpp:682-769)
// over the transmission functions of initialiseSimulation (simulationgeneral.cpp:345-646).
//
// Waves are held as S = F_x psi (the engine's mixed domain).  One forward slice is
//   psi' = U(t psi),  U = F2^-1 P F2 (propagator with the band-limit mask),
// so for a real loss L with g = dL/dpsi* (times 2) the adjoint slice is
//   u = U^H g' = F2^-1 conj(P) F2 g',  g = conj(t) u,  g_t = conj(psi) u,
// which is the forward column pass with conj(P') followed by a row pass that multiplies
// by conj(t) (and emits g_t).  Adjoint states are kept as res * F_x g so that the column
// pass sees the same scaling as in the forward direction.  The band limit t = BL(exp(i phi))
// is self-adjoint, so g_tau = BL(g_t) reuses the forward band-limit passes.
#include "adjoint_kernels.hpp"
#include "engine_internal.hpp"

namespace mse_rt {
namespace {

int ceil_sqrt(int n) {
    int k = 1;
    while (k * k < n) ++k;
    return k;
}

void require_cfg_waves(mse_engine* e, int cfg) {
    REQUIRE(e->mix_sets < 2, MSE_ERR_UNSUPPORTED, "position gradients are not defined under mixed static potentials");
    for (int w = 0; w < e->n_waves; ++w)
        REQUIRE(e->h_wave_cfg[w] == cfg, MSE_ERR_ARG,
                "every wave must refer to configuration " + std::to_string(cfg));
}

void copy_planes(mse_engine* e, float2* dst, const float2* src, size_t count) {
    HIPX(hipMemcpyAsync(dst, src, count * sizeof(float2), hipMemcpyDeviceToDevice, e->stream));
}

// One forward slice from `in` into `out` (in != out): the row pass reads one buffer and
// writes the other, so recomputed states need no copy of their predecessor.
void slice_step_into(mse_engine* e, const float2* in, float2* out, int nw, int slice) {
    mse::RowArgs ra = base_row_args(e);
    ra.slice = slice;
    row_pass(e, mse::ROW_IFFT_T_FFT, in, out, nw, e->N, ra);
    col_pass(e, mse::COL_PROP, out, out, nw, e->N, base_col_args(e));
}

}  // namespace
}  // namespace mse_rt

extern "C" {

int mse_forward_checkpointed(mse_engine* e, int cfg, double* intensity) {
    return run_guarded("mse_forward_checkpointed", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        check_waves(e);
        check_trans(e);
        require_cfg_waves(e, cfg);
        ConfigData& c = config(e, cfg);
        const int n = cfg_geom(e, c).n_slices;
        const int nw = e->n_waves;
        const size_t plane = (size_t)nw * e->N;
        // every state when all of them fit the budget (no recompute), else sqrt(n) checkpoints
        const bool keep_all = (double)n * (double)plane * sizeof(float2) <= (double)e->adj_budget;
        const int K = keep_all ? 1 : ceil_sqrt(std::max(n, 1));
        const int n_ck = (n + K - 1) / K;
        e->adj_ready = false;
        e->adj_ckpt.ensure(std::max(n_ck, 1) * plane);
        e->adj_exit.ensure(plane);
        // The state before slice s is checkpoint s / K when s % K == 0: the slice before writes
        // it there directly and slice s reads it there, so checkpoints cost no copies (other
        // slices step the engine's waves in place).
        if (n > 0) copy_planes(e, e->adj_ckpt.p, e->waves.p, plane);
        for (int s = 0; s < n; ++s) {
            const float2* in = s % K == 0 ? e->adj_ckpt.p + (size_t)(s / K) * plane : e->waves.p;
            float2* out = (s + 1 < n && (s + 1) % K == 0) ? e->adj_ckpt.p + (size_t)((s + 1) / K) * plane : e->waves.p;
            if (in == out)
                slice_step(e, out, nw, s);
            else
                slice_step_into(e, in, out, nw, s);
        }
        copy_planes(e, e->adj_exit.p, e->waves.p, plane);
        if (intensity) {
            e->diff_img.ensure(plane);
            mse::RowArgs ra = base_row_args(e);
            ra.scale = (float)(1.0 / e->res);
            ra.dout = e->diff_img.p;
            row_pass(e, mse::ROW_IFFT_INTENSITY, e->waves.p, e->scratch.p, nw, e->N, ra);
            HIPX(hipMemcpyAsync(intensity, e->diff_img.p, plane * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        }
        HIPX(hipStreamSynchronize(e->stream));
        e->adj_cfg = cfg;
        e->adj_k = K;
        e->adj_nw = nw;
        e->adj_nslices = n;
        e->adj_version = c.version;
        e->adj_ready = true;
    });
}

int mse_set_checkpoint_budget(mse_engine* e, unsigned long long bytes) {
    return run_guarded("mse_set_checkpoint_budget", e, [&]() {
        REQUIRE(e, MSE_ERR_ARG, "engine is NULL");
        e->adj_budget = (size_t)bytes;
    });
}

int mse_adjoint_positions(mse_engine* e, const double* dl_di, double* grad_xyz) {
    return run_guarded("mse_adjoint_positions", e, [&]() {
        REQUIRE(e && dl_di && grad_xyz, MSE_ERR_ARG, "NULL argument");
        REQUIRE(e->adj_ready, MSE_ERR_STATE, "mse_forward_checkpointed has not been called");
        ConfigData& c = config(e, e->adj_cfg);
        REQUIRE(c.version == e->adj_version && c.sorted && c.built, MSE_ERR_STATE,
                "atoms, geometry or transmission functions changed since mse_forward_checkpointed");
        REQUIRE(e->n_waves == e->adj_nw, MSE_ERR_STATE, "wave count changed since mse_forward_checkpointed");
        const mse_geometry& g = cfg_geom(e, c);
        const int n = e->adj_nslices, K = e->adj_k, nw = e->adj_nw;
        const size_t N = e->N, plane = (size_t)nw * N;
        const int n_atoms = c.n_input;

        // dL/dI, phases phi_s and the gradient accumulator
        e->adj_w.ensure(plane);
        HIPX(hipMemcpyAsync(e->adj_w.p, dl_di, plane * sizeof(double), hipMemcpyHostToDevice, e->stream));
        e->adj_phase.ensure((size_t)std::max(n, 1) * N);
        require_elements(e, c);
        mse::PotArgs pa = pot_args(e, c);
        // the phases from the same (tabulated or direct) potentials as the forward build
        pa.vtab = potential_tables(e, c, g);
        const bool tab = pa.vtab != nullptr;
        // as mse_build_transmission: full 3D 32 slices per launch, projected up to 256
        const int chunk = g.full3d ? 32 : 256;
        for (int s0 = 0; s0 < n; s0 += chunk) {
            const int cs = std::min(chunk, n - s0);
            dim3 grid(e->res / 16, e->res / 16, cs);
            float* o = e->adj_phase.p + (size_t)s0 * N;
            launch(e, PK_POT, [&]() {
                if (g.full3d) {
                    if (e->form == 0)
                        tab ? mse::k_phase_full3d<0, true><<<grid, 256, 0, e->stream>>>(o, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, pa, s0)
                            : mse::k_phase_full3d<0><<<grid, 256, 0, e->stream>>>(o, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, pa, s0);
                    else if (e->form == 1)
                        tab ? mse::k_phase_full3d<1, true><<<grid, 256, 0, e->stream>>>(o, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, pa, s0)
                            : mse::k_phase_full3d<1><<<grid, 256, 0, e->stream>>>(o, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, pa, s0);
                    else
                        tab ? mse::k_phase_full3d<2, true><<<grid, 256, 0, e->stream>>>(o, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, pa, s0)
                            : mse::k_phase_full3d<2><<<grid, 256, 0, e->stream>>>(o, c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, e->params3d.p, pa, s0);
                } else if (e->form == 0)
                    tab ? mse::k_phase_projected<0, true><<<grid, 256, 0, e->stream>>>(o, c.x.p, c.y.p, c.Z.p, c.block_start.p, e->params_proj.p, pa, s0)
                        : mse::k_phase_projected<0><<<grid, 256, 0, e->stream>>>(o, c.x.p, c.y.p, c.Z.p, c.block_start.p, e->params_proj.p, pa, s0);
                else if (e->form == 1)
                    tab ? mse::k_phase_projected<1, true><<<grid, 256, 0, e->stream>>>(o, c.x.p, c.y.p, c.Z.p, c.block_start.p, e->params_proj.p, pa, s0)
                        : mse::k_phase_projected<1><<<grid, 256, 0, e->stream>>>(o, c.x.p, c.y.p, c.Z.p, c.block_start.p, e->params_proj.p, pa, s0);
                else
                    tab ? mse::k_phase_projected<2, true><<<grid, 256, 0, e->stream>>>(o, c.x.p, c.y.p, c.Z.p, c.block_start.p, e->params_proj.p, pa, s0)
                        : mse::k_phase_projected<2><<<grid, 256, 0, e->stream>>>(o, c.x.p, c.y.p, c.Z.p, c.block_start.p, e->params_proj.p, pa, s0);
            });
        }
        const size_t ng = 3 * (size_t)std::max(n_atoms, 1);
        e->adj_grad.ensure(ng);
        launch(e, PK_OTHER, [&]() {
            mse::k_fill_zero<<<(unsigned)((ng + 255) / 256), 256, 0, e->stream>>>(e->adj_grad.p, ng);
        });
        // first sorted atom of every slice
        const int nblk = g.blocks_x * g.blocks_y;
        std::vector<int> bs((size_t)n * nblk + 1);
        HIPX(hipMemcpyAsync(bs.data(), c.block_start.p, bs.size() * sizeof(int), hipMemcpyDeviceToHost, e->stream));

        e->adj_g.ensure(plane);
        e->adj_gt.ensure(plane);  // per-wave g_t
        e->adj_seg.ensure((size_t)std::max(K - 1, 1) * plane);  // states 1..K-1 of a segment
        HIPX(hipStreamSynchronize(e->stream));

        // seed: res F_x (2 w psi_N)
        mse::RowArgs ra = base_row_args(e);
        ra.weights = e->adj_w.p;
        row_pass(e, mse::ROW_ADJ_SEED, e->adj_exit.p, e->adj_g.p, nw, N, ra);

        mse::ColArgs ca = base_col_args(e);
        // Without per-wave beam tables the backward column pass is the forward propagator pass
        // (with the variant tuned for nw waves) on the conjugated px, py tables; with them it
        // conjugates the per-wave factors itself (COL_PROP_CONJ).
        const bool conj_tables = ca.wave_prop == nullptr;
        if (conj_tables) {
            e->adj_pconj.ensure(2 * (size_t)e->res);
            launch(e, PK_OTHER, [&]() {
                mse::k_conj_tables<<<(unsigned)((2 * e->res + 255) / 256), 256, 0, e->stream>>>(e->px.p, e->py.p, e->res,
                                                                                              e->adj_pconj.p);
            });
            ca.px = e->adj_pconj.p;
            ca.py = e->adj_pconj.p + e->res;
        }
        e->adj_gtau.ensure((size_t)std::max(n, 1) * N);  // g_t of every slice, summed over the waves
        const int n_seg = (n + K - 1) / K;
        for (int j = n_seg - 1; j >= 0; --j) {
            const int start = j * K, len = std::min(K, n - start);
            // recompute the forward states of the segment from its checkpoint: state i of
            // the segment is the checkpoint for i = 0, else slot i - 1 of adj_seg
            const float2* ck = e->adj_ckpt.p + (size_t)j * plane;
            auto slot = [&](int i) { return e->adj_seg.p + (size_t)(i - 1) * plane; };
            auto state = [&](int i) -> const float2* { return i == 0 ? ck : slot(i); };
            for (int i = 1; i < len; ++i) slice_step_into(e, state(i - 1), slot(i), nw, start + i - 1);
            for (int s = start + len - 1; s >= start; --s) {
                col_pass(e, conj_tables ? mse::COL_PROP : mse::COL_PROP_CONJ, e->adj_g.p, e->adj_g.p, nw, N, ca);
                float2* gs = e->adj_gtau.p + (size_t)s * N;
                mse::RowArgs rb = base_row_args(e);
                rb.slice = s;
                rb.aux_in = state(s - start);
                rb.aux_out = nw == 1 ? gs : e->adj_gt.p;  // one wave: its g_t is the sum
                row_pass(e, mse::ROW_ADJ, e->adj_g.p, e->adj_g.p, nw, N, rb);
                if (nw > 1)
                    launch(e, PK_OTHER, [&]() {
                        mse::k_sum_planes<<<(unsigned)((N + 255) / 256), 256, 0, e->stream>>>(e->adj_gt.p, nw, N, N, gs);
                    });
            }
        }
        // the position terms of every slice in one pass: g_tau = BL(g_t) of all slices, then
        // one gradient launch over all atoms (projected: each atom reads its own slice's
        // planes) or over all slices (full 3D: each slice sums the atoms of s +- load_z)
        const int a_first = bs[0], a_end = bs[(size_t)n * nblk];
        if (n > 0 && a_end > a_first) {
            band_limit_planes(e, e->adj_gtau.p, n);
            const float2* gt = e->adj_gtau.p;
            const float* ph = e->adj_phase.p;
            if (g.full3d) {
                dim3 tiles(e->res / 16, e->res / 16, n);
                launch(e, PK_OTHER, [&]() {
                    if (e->form == 0)
                        mse::k_atom_grad_full3d<0><<<tiles, 256, 0, e->stream>>>(c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, c.orig.p, e->params3d.p, pa, 0, gt, ph, e->adj_grad.p);
                    else if (e->form == 1)
                        mse::k_atom_grad_full3d<1><<<tiles, 256, 0, e->stream>>>(c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, c.orig.p, e->params3d.p, pa, 0, gt, ph, e->adj_grad.p);
                    else
                        mse::k_atom_grad_full3d<2><<<tiles, 256, 0, e->stream>>>(c.x.p, c.y.p, c.z.p, c.Z.p, c.block_start.p, c.orig.p, e->params3d.p, pa, 0, gt, ph, e->adj_grad.p);
                });
            } else {
                dim3 grid(a_end - a_first);
                const int* bsp = c.block_start.p;
                launch(e, PK_OTHER, [&]() {
                    if (e->form == 0)
                        tab ? mse::k_atom_grad_projected<0, true><<<grid, 256, 0, e->stream>>>(c.x.p, c.y.p, c.Z.p, c.orig.p, a_first, bsp, nblk, n, e->params.p, pa, gt, ph, e->adj_grad.p)
                            : mse::k_atom_grad_projected<0><<<grid, 256, 0, e->stream>>>(c.x.p, c.y.p, c.Z.p, c.orig.p, a_first, bsp, nblk, n, e->params.p, pa, gt, ph, e->adj_grad.p);
                    else if (e->form == 1)
                        tab ? mse::k_atom_grad_projected<1, true><<<grid, 256, 0, e->stream>>>(c.x.p, c.y.p, c.Z.p, c.orig.p, a_first, bsp, nblk, n, e->params.p, pa, gt, ph, e->adj_grad.p)
                            : mse::k_atom_grad_projected<1><<<grid, 256, 0, e->stream>>>(c.x.p, c.y.p, c.Z.p, c.orig.p, a_first, bsp, nblk, n, e->params.p, pa, gt, ph, e->adj_grad.p);
                    else
                        tab ? mse::k_atom_grad_projected<2, true><<<grid, 256, 0, e->stream>>>(c.x.p, c.y.p, c.Z.p, c.orig.p, a_first, bsp, nblk, n, e->params.p, pa, gt, ph, e->adj_grad.p)
                            : mse::k_atom_grad_projected<2><<<grid, 256, 0, e->stream>>>(c.x.p, c.y.p, c.Z.p, c.orig.p, a_first, bsp, nblk, n, e->params.p, pa, gt, ph, e->adj_grad.p);
                });
            }
        }
        if (n_atoms > 0)
            HIPX(hipMemcpyAsync(grad_xyz, e->adj_grad.p, 3 * (size_t)n_atoms * sizeof(double), hipMemcpyDeviceToHost,
                                e->stream));
        HIPX(hipStreamSynchronize(e->stream));
    });
}

}  // extern "C"
