// Cost model of the rows scheduler: C-ABI (included at the end of
// dcx_api.hip).  No reference counterpart: the reference's launch shape and
// pass order are fixed (dcore, main.cu:197-262).  The model is trained on
// the pass log the decomposition itself records (dcx_graph_pass_log) and
// chooses, per team row, between peeling it and deriving it from the row
// above, and the team widths; both kinds of row are exact, so every plan
// gives the same bits.
#pragma once

extern "C" {

uint32_t dcx_mlp_param_count(void) { return (uint32_t)kMlpParams; }

dcx_status dcx_mlp_create(dcx_ctx* ctx, uint64_t seed, dcx_mlp** out) {
    if (!ctx || !out) return DCX_ERR_INVALID_ARG;
    *out = nullptr;
    std::unique_ptr<dcx_mlp> m(new (std::nothrow) dcx_mlp);
    if (!m) return DCX_ERR_OUT_OF_MEMORY;
    m->ctx = ctx;
    dcx_status s = guarded(ctx, [&]() -> dcx_status {
        dev_alloc(ctx, &m->p, kMlpParams);
        dev_alloc(ctx, &m->m, kMlpParams);
        dev_alloc(ctx, &m->v, kMlpParams);
        dev_alloc(ctx, &m->grad, kMlpParams);
        dev_alloc(ctx, &m->loss, 1);
        // He-uniform weights from SplitMix64, zero biases
        std::vector<float> p(kMlpParams, 0.f);
        SplitMix r(mix64(seed ^ 0x5eedull));
        auto fill = [&](int off, int n, int fan_in) {
            const double a = std::sqrt(6.0 / fan_in);
            for (int i = 0; i < n; ++i) p[off + i] = (float)((2.0 * r.uniform() - 1.0) * a);
        };
        fill(0, kMlpHidden * kMlpIn, kMlpIn);
        fill(kMlpHidden * kMlpIn + kMlpHidden, kMlpHidden * kMlpHidden, kMlpHidden);
        fill(kMlpHidden * kMlpIn + 2 * kMlpHidden + kMlpHidden * kMlpHidden, kMlpHidden,
             kMlpHidden);
        HIPX(ctx, hipMemcpyAsync(m->p, p.data(), 4ull * kMlpParams, hipMemcpyHostToDevice,
                                 ctx->stream));
        HIPX(ctx, hipMemsetAsync(m->m, 0, 4ull * kMlpParams, ctx->stream));
        HIPX(ctx, hipMemsetAsync(m->v, 0, 4ull * kMlpParams, ctx->stream));
        sync(ctx);
        return DCX_OK;
    });
    if (s != DCX_OK) {
        dev_free(m->p), dev_free(m->m), dev_free(m->v), dev_free(m->grad), dev_free(m->loss);
        return s;
    }
    *out = m.release();
    return DCX_OK;
}

dcx_status dcx_mlp_free(dcx_mlp* m) {
    if (!m) return DCX_ERR_INVALID_ARG;
    dcx_ctx* c = m->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->sched_mlp == m) c->sched_mlp = nullptr;
    dev_free(m->p), dev_free(m->m), dev_free(m->v), dev_free(m->grad), dev_free(m->loss);
    dev_free(m->buf);
    delete m;
    return DCX_OK;
}

dcx_status dcx_mlp_get_params(dcx_mlp* m, float* out) {
    if (!m || !out) return DCX_ERR_INVALID_ARG;
    return guarded(m->ctx, [&]() -> dcx_status {
        sync(m->ctx);
        HIPX(m->ctx, hipMemcpy(out, m->p, 4ull * kMlpParams, hipMemcpyDeviceToHost));
        return DCX_OK;
    });
}

dcx_status dcx_mlp_set_params(dcx_mlp* m, const float* in) {
    if (!m || !in) return DCX_ERR_INVALID_ARG;
    return guarded(m->ctx, [&]() -> dcx_status {
        sync(m->ctx);
        HIPX(m->ctx, hipMemcpy(m->p, in, 4ull * kMlpParams, hipMemcpyHostToDevice));
        HIPX(m->ctx, hipMemset(m->m, 0, 4ull * kMlpParams));
        HIPX(m->ctx, hipMemset(m->v, 0, 4ull * kMlpParams));
        m->step = 0;
        return DCX_OK;
    });
}

dcx_status dcx_mlp_predict(dcx_mlp* m, const float* X, uint32_t n, float* y) {
    if (!m || (n && (!X || !y))) return DCX_ERR_INVALID_ARG;
    return guarded(m->ctx, [&]() -> dcx_status {
        std::vector<float> x(X, X + (size_t)n * kMlpIn);
        const std::vector<float> r = mlp_predict(m->ctx, m, x, n);
        std::copy(r.begin(), r.end(), y);
        return DCX_OK;
    });
}

// Summed gradient of 0.5 * sum (f(X_s) - T_s)^2 and the summed loss.
dcx_status dcx_mlp_gradient(dcx_mlp* m, const float* X, const float* T, uint32_t n, float* grad,
                            float* loss) {
    if (!m || !grad || (n && (!X || !T))) return DCX_ERR_INVALID_ARG;
    return guarded(m->ctx, [&]() -> dcx_status {
        dcx_ctx* c = m->ctx;
        mlp_stage(c, m, (size_t)n * (kMlpIn + 1));
        if (n) {
            HIPX(c, hipMemcpyAsync(m->buf, X, 4ull * n * kMlpIn, hipMemcpyHostToDevice, c->stream));
            HIPX(c, hipMemcpyAsync(m->buf + (size_t)n * kMlpIn, T, 4ull * n, hipMemcpyHostToDevice,
                                   c->stream));
        }
        HIPX(c, launch_mlp_backward(m->p, m->buf, m->buf + (size_t)n * kMlpIn, n, m->grad, m->loss,
                                    c->stream));
        HIPX(c, hipMemcpyAsync(grad, m->grad, 4ull * kMlpParams, hipMemcpyDeviceToHost, c->stream));
        float l = 0.f;
        HIPX(c, hipMemcpyAsync(&l, m->loss, 4, hipMemcpyDeviceToHost, c->stream));
        sync(c);
        if (loss) *loss = l;
        return DCX_OK;
    });
}

// Full-batch Adam: `epochs` steps on the mean squared error; *final_loss =
// mean 0.5 (f - T)^2 before the last step.
dcx_status dcx_mlp_train(dcx_mlp* m, const float* X, const float* T, uint32_t n, uint32_t epochs,
                         float lr, float* final_loss) {
    if (!m || !n || !X || !T || lr <= 0.f) return DCX_ERR_INVALID_ARG;
    return guarded(m->ctx, [&]() -> dcx_status {
        dcx_ctx* c = m->ctx;
        mlp_stage(c, m, (size_t)n * (kMlpIn + 1));
        HIPX(c, hipMemcpyAsync(m->buf, X, 4ull * n * kMlpIn, hipMemcpyHostToDevice, c->stream));
        HIPX(c, hipMemcpyAsync(m->buf + (size_t)n * kMlpIn, T, 4ull * n, hipMemcpyHostToDevice,
                               c->stream));
        for (uint32_t e = 0; e < epochs; ++e) {
            HIPX(c, launch_mlp_backward(m->p, m->buf, m->buf + (size_t)n * kMlpIn, n, m->grad,
                                        m->loss, c->stream));
            HIPX(c, launch_mlp_adam(m->p, m->grad, m->m, m->v, n, lr, ++m->step, c->stream));
        }
        float l = 0.f;
        HIPX(c, hipMemcpyAsync(&l, m->loss, 4, hipMemcpyDeviceToHost, c->stream));
        sync(c);
        if (final_loss) *final_loss = l / (float)n;
        return DCX_OK;
    });
}

// The batched rows launch of this context takes its team widths from the
// model (NULL: the work-share rule).
dcx_status dcx_set_scheduler(dcx_ctx* ctx, dcx_mlp* m) {
    if (!ctx || (m && m->ctx != ctx)) return DCX_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->sched_mlp = m;
    return DCX_OK;
}

// Training samples from the resident decomposition's pass log: one per team
// row of the batched launch (peeled, or derived from the row above),
// features of pass_features with the team width and kind it ran as and the
// work statistics its run logged (with_log = 0: those as unknown), target
// log2(ms + 0.01).
dcx_status dcx_graph_pass_features(dcx_ctx* ctx, const dcx_graph* g, int32_t with_log,
                                   uint32_t cap, float* X, float* T, uint32_t* n) {
    dcx_status s = check_graph(ctx, g);
    if (s != DCX_OK) return s;
    if (!n || (cap && (!X || !T))) return DCX_ERR_INVALID_ARG;
    return guarded(ctx, [&]() -> dcx_status {
        *n = 0;
        if (!g->rows_valid || g->h_pstats.size() < ((size_t)g->kmax + 2) * kStatWords)
            return DCX_OK;
        uint32_t m = 0;
        for (int32_t k = 0; k <= g->kmax; ++k) {
            // team rows only (peeled or derived): the LDS-resident rows
            // (team 0) are not planned by the model
            const int kind = last_pass_kind(g, k);
            if (kind < 0) continue;
            const uint32_t* w = g->h_pstats.data() + (size_t)(k + 1) * kStatWords;
            const uint64_t ticks = (uint64_t)w[kStatTicksLo] | ((uint64_t)w[kStatTicksHi] << 32);
            const double ms = (double)ticks / ctx->wall_khz;
            if (m < cap) {
                pass_features(g, k, g->h_pass_team[k + 1], kind, with_log != 0,
                              X + (size_t)m * kMlpIn);
                T[m] = (float)std::log2(ms + 0.01);
            }
            ++m;
        }
        *n = m;
        return DCX_OK;
    });
}

}  // extern "C"
