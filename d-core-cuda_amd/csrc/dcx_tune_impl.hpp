// Autotuner for the batched rows launch: included at the end of dcx_api.hip.
//
// The batched launch has four shape knobs -- workgroups in the launch,
// the most workgroups one k-pass may get (team_max), the alive-list length
// per team member below which a team narrows (shrink_at), and the largest
// core that runs LDS-resident beside the teams (lds cap).  Every
// shape gives the same bits (the passes are exact whatever their team), so
// the tuner only measures: coordinate descent over the three knobs, each
// candidate timed with HIP events around the rows kernel on the graph
// itself (best of `reps`), the winner left in the context's configuration.
// There is no counterpart in the reference, whose launch shape is fixed
// (BLK_NUMS/BLK_DIM, src/common.h:38-56).
#pragma once

namespace {

struct TuneShape {
    int32_t slots, team_max, shrink_at;
    uint32_t lds;
};

double time_rows_shape(dcx_ctx* c, dcx_graph* g, const TuneShape& t, int reps) {
    c->cfg.rows_slots = t.slots;
    c->cfg.team_max = t.team_max;
    c->cfg.shrink_at = t.shrink_at;
    c->lds_cap = t.lds;
    double best = 1e300;
    for (int r = 0; r < reps; ++r) {
        compute_rows(c, g, DCX_VARIANT_BATCHED);
        best = std::min(best, (double)c->timing.rows_kernel_ms);
    }
    return best;
}

}  // namespace

extern "C" {

dcx_status dcx_autotune(dcx_ctx* ctx, dcx_graph* g, int32_t reps, dcx_tune_result* out) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (reps <= 0) reps = 2;
    return guarded(ctx, [&]() -> dcx_status {
        if (!g->kmax_valid) compute_kmax(ctx, g);
        const int32_t cus = ctx->num_cus;
        const int32_t slot_opts[] = {cus, 2 * cus};
        const int32_t team_opts[] = {4, 8, 16, 32, 64, 128, 256};
        const int32_t shrink_opts[] = {1024, 2048, 4096, 8192, 16384};
        const uint32_t lds_opts[] = {ctx->lds_cap_max, 28672, 24576, 20480, 16384, 8192, 0};
        TuneShape best{2 * cus, (int32_t)kTeamMax, (int32_t)kShrinkAt, ctx->lds_cap};
        int tried = 0;
        const double def_ms = time_rows_shape(ctx, g, best, reps);
        double best_ms = def_ms;
        ++tried;
        auto consider = [&](TuneShape t) {
            const double ms = time_rows_shape(ctx, g, t, reps);
            ++tried;
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        };
        for (int32_t tm : team_opts)
            if (tm != best.team_max) consider(TuneShape{best.slots, tm, best.shrink_at, best.lds});
        const int32_t tm_best = best.team_max;
        for (int32_t sh : shrink_opts)
            if (sh != best.shrink_at) consider(TuneShape{best.slots, tm_best, sh, best.lds});
        const int32_t sh_best = best.shrink_at;
        for (int32_t sl : slot_opts)
            if (sl != best.slots) consider(TuneShape{sl, tm_best, sh_best, best.lds});
        const int32_t sl_best = best.slots;
        for (uint32_t lc : lds_opts)
            if (lc <= ctx->lds_cap_max && lc != best.lds)
                consider(TuneShape{sl_best, tm_best, sh_best, lc});
        // leave the winner configured and its rows resident
        time_rows_shape(ctx, g, best, 1);
        if (out) {
            std::memset(out, 0, sizeof *out);
            out->rows_slots = best.slots;
            out->team_max = best.team_max;
            out->shrink_at = best.shrink_at;
            out->candidates = tried;
            out->best_ms = best_ms;
            out->default_ms = def_ms;
            out->lds_members = best.lds ? (int32_t)best.lds : -1;
        }
        return DCX_OK;
    });
}

}  // extern "C"
