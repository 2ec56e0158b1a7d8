// Context, error reporting, device memory helpers and RNS NTT plans.
#include <cstring>

#include "host_math.hpp"
#include "internal.hpp"

namespace ck {

static thread_local std::string g_last_error;

void set_error(int code, const std::string &msg) {
    (void)code;
    g_last_error = msg;
}

int fail(int code, const std::string &msg) {
    set_error(code, msg);
    return code;
}

int check_plan_limbs(const ckks_rns_plan *plan, size_t n_limbs, const char *what) {
    if (!plan) return fail(CKKS_ERR_INVALID, std::string(what) + ": null plan");
    if (n_limbs == 0 || n_limbs > plan->moduli.size())
        return fail(CKKS_ERR_INVALID,
                    std::string(what) + ": limb count must be within the plan basis");
    return CKKS_OK;
}

}  // namespace ck

using namespace ck;

int ckks_context::workspace(int slot_in_lane, size_t bytes, void **out) {
    const int slot = slot_in_lane + kSlots * lane;
    if (bytes > ws_bytes[slot] && capturing)
        return fail(CKKS_ERR_INVALID,
                    "workspace must be sized before graph capture: run the schedule once eagerly first");
    if (bytes > ws_bytes[slot]) {
        if (ws[slot]) {
            hipStreamSynchronize(stream);
            hipFree(ws[slot]);
            ws[slot] = nullptr;
            ws_bytes[slot] = 0;
        }
        hipError_t e = hipMalloc(&ws[slot], bytes);
        if (e != hipSuccess) {
            ws[slot] = nullptr;
            return fail(CKKS_ERR_OOM, std::string("workspace allocation failed: ") + hipGetErrorString(e));
        }
        ws_bytes[slot] = bytes;
    }
    *out = ws[slot];
    return CKKS_OK;
}

extern "C" {

const char *ckks_last_error(void) { return g_last_error.c_str(); }

const char *ckks_version(void) { return "ckks_hip 0.1 (gfx950)"; }

int ckks_context_create(int device, void *stream, ckks_context **out) {
    if (!out) return fail(CKKS_ERR_INVALID, "null output pointer");
    int count = 0;
    CK_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(CKKS_ERR_INVALID, "device index out of range");
    CK_HIP(hipSetDevice(device));
    std::unique_ptr<ckks_context> ctx(new ckks_context());
    ctx->device = device;
    ctx->stream = (hipStream_t)stream;
    int rc = ntt_scratch_lookup(device, &ctx->ntt_scratch);
    if (rc) return rc;
    *out = ctx.release();
    return CKKS_OK;
}

int ckks_context_destroy(ckks_context *ctx) {
    if (!ctx) return CKKS_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    while (!ctx->servers.empty()) ckks_server_destroy(*ctx->servers.begin());
    while (!ctx->graphs.empty()) ckks_graph_destroy(*ctx->graphs.begin());
    while (!ctx->keys.empty()) ckks_ksk_destroy(*ctx->keys.begin());
    while (!ctx->plans.empty()) ckks_rns_plan_destroy(*ctx->plans.begin());
    while (!ctx->ntt64_plans.empty()) ckks_ntt64_plan_destroy(*ctx->ntt64_plans.begin());
    for (auto &kv : ctx->tables) {
        hipFree(kv.second.fwd);
        hipFree(kv.second.inv);
        if (kv.second.rowf) hipFree(kv.second.rowf);
        if (kv.second.rowf_inv) hipFree(kv.second.rowf_inv);
    }
    for (void *p : ctx->owned) hipFree(p);
    for (int i = 0; i < 2 * ckks_context::kSlots; ++i)
        if (ctx->ws[i]) hipFree(ctx->ws[i]);
    if (ctx->ev_fork) hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) hipEventDestroy(ctx->ev_join);
    if (ctx->aux) hipStreamDestroy(ctx->aux);
    delete ctx;
    return CKKS_OK;
}

int ckks_context_set_stream(ckks_context *ctx, void *stream) {
    if (!ctx) return fail(CKKS_ERR_INVALID, "null context");
    CK_HIP(hipStreamSynchronize(ctx->stream));
    ctx->stream = (hipStream_t)stream;
    return CKKS_OK;
}

int ckks_sync(ckks_context *ctx) {
    if (!ctx) return fail(CKKS_ERR_INVALID, "null context");
    CK_HIP(hipStreamSynchronize(ctx->stream));
    return CKKS_OK;
}

int ckks_set_ntt_chunk_bytes(ckks_context *ctx, size_t bytes) {
    if (!ctx) return fail(CKKS_ERR_INVALID, "null context");
    ctx->ntt_chunk_bytes = bytes;
    return CKKS_OK;
}

}  // extern "C"

struct ckks_graph {
    ckks_context *ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ckks_execution_profile counts = {};  // operations recorded in the capture
};

static void add_profile(ckks_execution_profile &to, const ckks_execution_profile &d) {
    const size_t n = sizeof(ckks_execution_profile) / sizeof(uint64_t);
    auto *a = reinterpret_cast<uint64_t *>(&to);
    auto *b = reinterpret_cast<const uint64_t *>(&d);
    for (size_t i = 0; i < n; ++i) a[i] += b[i];
}

extern "C" {

int ckks_graph_begin(ckks_context *ctx) {
    if (!ctx) return fail(CKKS_ERR_INVALID, "null context");
    if (!ctx->stream) return fail(CKKS_ERR_INVALID, "graph capture needs a context on a non-null stream");
    if (ctx->capturing) return fail(CKKS_ERR_INVALID, "a graph capture is already open");
    CK_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    ctx->captured = ckks_execution_profile{};
    return CKKS_OK;
}

int ckks_graph_end(ckks_context *ctx, ckks_graph **out) {
    if (!ctx || !out) return fail(CKKS_ERR_INVALID, "null argument");
    if (!ctx->capturing) return fail(CKKS_ERR_INVALID, "no graph capture is open");
    ctx->capturing = false;
    std::unique_ptr<ckks_graph> g(new ckks_graph());
    g->ctx = ctx;
    g->counts = ctx->captured;
    CK_HIP(hipStreamEndCapture(ctx->stream, &g->graph));
    CK_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    ctx->graphs.insert(g.get());
    *out = g.release();
    return CKKS_OK;
}

int ckks_graph_launch(ckks_graph *g) {
    if (!g) return fail(CKKS_ERR_INVALID, "null graph");
    CK_HIP(hipGraphLaunch(g->exec, g->ctx->stream));
    add_profile(g->ctx->counts(), g->counts);
    return CKKS_OK;
}

int ckks_profile_get(const ckks_context *ctx, ckks_execution_profile *out) {
    if (!ctx || !out) return fail(CKKS_ERR_INVALID, "null argument");
    *out = ctx->profile;
    return CKKS_OK;
}

int ckks_profile_reset(ckks_context *ctx) {
    if (!ctx) return fail(CKKS_ERR_INVALID, "null context");
    ctx->profile = ckks_execution_profile{};
    return CKKS_OK;
}

int ckks_graph_destroy(ckks_graph *g) {
    if (!g) return CKKS_OK;
    hipStreamSynchronize(g->ctx->stream);
    g->ctx->graphs.erase(g);
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
    return CKKS_OK;
}

int ckks_set_base_conversion(ckks_context *ctx, int mode) {
    if (!ctx) return fail(CKKS_ERR_INVALID, "null context");
    if (mode != CKKS_BASECONV_AUTO && mode != CKKS_BASECONV_VALU && mode != CKKS_BASECONV_MFMA)
        return fail(CKKS_ERR_INVALID, "unknown base-conversion mode");
    ctx->base_conversion = mode;
    return CKKS_OK;
}

int ckks_set_ntt_variant(ckks_context *ctx, uint32_t logn, int variant) {
    if (!ctx) return fail(CKKS_ERR_INVALID, "null context");
    if (logn >= 32 || variant < 0 || variant >= CKKS_NTT_VARIANT_COUNT)
        return fail(CKKS_ERR_INVALID, "unknown NTT variant");
    ctx->ntt_variant[logn] = (uint8_t)variant;
    return CKKS_OK;
}

int ckks_set_key_switch_fusion(ckks_context *ctx, int mode) {
    if (!ctx) return fail(CKKS_ERR_INVALID, "null context");
    if (mode < CKKS_KS_FUSION_AUTO || mode > CKKS_KS_FUSION_INVERSE)
        return fail(CKKS_ERR_INVALID, "unknown key-switch fusion mode");
    ctx->ks_fusion = mode;
    return CKKS_OK;
}

int ckks_set_key_switch_split(ckks_context *ctx, int mode) {
    if (!ctx) return fail(CKKS_ERR_INVALID, "null context");
    if (mode != CKKS_KS_SPLIT_AUTO && mode != CKKS_KS_SPLIT_OFF && mode != CKKS_KS_SPLIT_ON)
        return fail(CKKS_ERR_INVALID, "unknown key-switch split mode");
    ctx->ks_split = mode;
    return CKKS_OK;
}

int ckks_get_ntt_variant(const ckks_context *ctx, uint32_t logn, int *variant) {
    if (!ctx || !variant || logn >= 32) return fail(CKKS_ERR_INVALID, "invalid argument");
    *variant = ctx->ntt_variant[logn];
    return CKKS_OK;
}

int ckks_malloc(ckks_context *ctx, size_t bytes, void **dev) {
    if (!ctx || !dev) return fail(CKKS_ERR_INVALID, "null argument");
    hipError_t e = hipMalloc(dev, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(CKKS_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return CKKS_OK;
}

int ckks_free(ckks_context *ctx, void *dev) {
    if (!ctx) return fail(CKKS_ERR_INVALID, "null context");
    CK_HIP(hipStreamSynchronize(ctx->stream));
    CK_HIP(hipFree(dev));
    return CKKS_OK;
}

int ckks_upload(ckks_context *ctx, void *dev, const void *host, size_t bytes) {
    if (!ctx) return fail(CKKS_ERR_INVALID, "null context");
    CK_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    CK_HIP(hipStreamSynchronize(ctx->stream));
    return CKKS_OK;
}

int ckks_download(ckks_context *ctx, void *host, const void *dev, size_t bytes) {
    if (!ctx) return fail(CKKS_ERR_INVALID, "null context");
    CK_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    CK_HIP(hipStreamSynchronize(ctx->stream));
    return CKKS_OK;
}

int ckks_find_ntt_primes(unsigned bits, size_t degree, size_t count, const uint64_t *skip,
                         size_t n_skip, uint64_t *out, size_t *found) {
    if (!out || !found || (n_skip && !skip)) return fail(CKKS_ERR_INVALID, "null argument");
    *found = 0;
    if (bits < 2 || bits > 62 || degree < 2 || (degree & (degree - 1)))
        return fail(CKKS_ERR_INVALID, "prime search needs 2 <= bits <= 62 and a power-of-two degree");
    const uint64_t two_n = 2 * (uint64_t)degree;
    uint64_t cand = (1ull << bits) - 1;
    cand -= (cand - 1) % two_n;
    while (*found < count) {
        if (cand > 2 && (cand - 1) % two_n == 0 && host::is_prime(cand)) {
            bool skipped = false;
            for (size_t i = 0; i < n_skip; ++i) skipped |= skip[i] == cand;
            if (!skipped) out[(*found)++] = cand;
        }
        if (cand <= two_n + 1) break;
        cand -= two_n;
    }
    return CKKS_OK;
}

int ckks_rns_plan_create(ckks_context *ctx, const uint64_t *moduli, size_t n_limbs,
                         size_t degree, ckks_rns_plan **out) {
    if (!ctx || !moduli || !out) return fail(CKKS_ERR_INVALID, "null argument");
    if (n_limbs == 0) return fail(CKKS_ERR_INVALID, "RNS NTT plan requires at least one modulus");
    if (degree == 0 || (degree & (degree - 1)))
        return fail(CKKS_ERR_INVALID, "NTT degree must be a positive power of two");
    uint32_t logn = 0;
    while (((size_t)1 << logn) < degree) ++logn;
    if (logn < 10 || logn > 17)
        return fail(CKKS_ERR_UNSUPPORTED, "GPU NTT supports degrees 2^10 .. 2^17");
    for (size_t i = 0; i < n_limbs; ++i)
        for (size_t j = 0; j < i; ++j)
            if (moduli[i] == moduli[j]) return fail(CKKS_ERR_INVALID, "RNS moduli must be distinct");
    CK_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<ckks_rns_plan> plan(new ckks_rns_plan());
    plan->ctx = ctx;
    plan->logn = logn;
    plan->degree = degree;
    for (size_t i = 0; i < n_limbs; ++i) {
        uint64_t q = moduli[i];
        if (q >= (1ull << 62)) return fail(CKKS_ERR_UNSUPPORTED, "GPU path requires every modulus < 2^62");
        if ((q - 1) % (2 * (uint64_t)degree) != 0) return fail(CKKS_ERR_INVALID, "2N must divide q - 1");
        if (!host::is_prime(q)) return fail(CKKS_ERR_INVALID, "RNS modulus must be prime");
        uint64_t psi = host::negacyclic_root(q, degree);
        if (!psi) return fail(CKKS_ERR_INVALID, "no primitive 2N-th root found");
        auto key = std::make_pair(q, logn);
        auto it = ctx->tables.find(key);
        if (it == ctx->tables.end()) {
            TwiddleTables tt;
            tt.q = q;
            tt.logn = logn;
            std::vector<uint64_t> fwd, inv;
            if (q < host::kFpModulusBound) {
                host::twiddle_table_fp(q, psi, logn, fwd);
                host::twiddle_table_fp(q, host::invmod(psi, q), logn, inv);
            } else {
                host::twiddle_table(q, psi, logn, fwd);
                host::twiddle_table(q, host::invmod(psi, q), logn, inv);
            }
            size_t bytes = fwd.size() * sizeof(uint64_t);
            CK_HIP(hipMalloc(&tt.fwd, bytes));
            CK_HIP(hipMalloc(&tt.inv, bytes));
            CK_HIP(hipMemcpy(tt.fwd, fwd.data(), bytes, hipMemcpyHostToDevice));
            CK_HIP(hipMemcpy(tt.inv, inv.data(), bytes, hipMemcpyHostToDevice));
            if (q < host::kFpModulusBound && logn > 12) {
                std::vector<uint64_t> rf, rfi;
                host::row_factor_table_fp(q, psi, logn, rf);
                host::row_factor_table_fp(q, host::invmod(psi, q), logn, rfi);
                const size_t rb = rf.size() * sizeof(uint64_t);
                CK_HIP(hipMalloc(&tt.rowf, rb));
                CK_HIP(hipMalloc(&tt.rowf_inv, rb));
                CK_HIP(hipMemcpy(tt.rowf, rf.data(), rb, hipMemcpyHostToDevice));
                CK_HIP(hipMemcpy(tt.rowf_inv, rfi.data(), rb, hipMemcpyHostToDevice));
            }
            it = ctx->tables.emplace(key, tt).first;
        }
        DevPrime p{};
        p.q = q;
        p.fwd = it->second.fwd;
        p.inv = it->second.inv;
        p.rowf = it->second.rowf;
        p.rowf_inv = it->second.rowf_inv;
        p.ninv = host::invmod(degree % q, q);
        p.ninv_q = host::shoup_quotient(p.ninv, q);
        host::barrett(q, p.br_hi, p.br_lo);
        p.logn = logn;
        p.fp = q < host::kFpModulusBound ? 1u : 0u;
        p.qd = (double)q;
        p.qinv = 1.0 / (double)q;
        p.ninv_d = (double)p.ninv;
        p.ninv_p = (double)p.ninv / (double)q;
        plan->moduli.push_back(q);
        plan->psi.push_back(psi);
        plan->h_primes.push_back(p);
    }
    std::vector<RescaleConst> resc(n_limbs * n_limbs);
    for (size_t l = 1; l < n_limbs; ++l)
        for (size_t i = 0; i < l; ++i) {
            uint64_t qi = plan->moduli[i], ql = plan->moduli[l];
            RescaleConst &r = resc[l * n_limbs + i];
            r.ql_mod = ql % qi;
            r.inv = host::invmod(r.ql_mod, qi);
            r.inv_q = host::shoup_quotient(r.inv, qi);
        }
    CK_HIP(hipMalloc(&plan->d_rescale, resc.size() * sizeof(RescaleConst)));
    CK_HIP(hipMemcpy(plan->d_rescale, resc.data(), resc.size() * sizeof(RescaleConst),
                     hipMemcpyHostToDevice));
    CK_HIP(hipMalloc(&plan->d_primes, n_limbs * sizeof(DevPrime)));
    CK_HIP(hipMemcpy(plan->d_primes, plan->h_primes.data(), n_limbs * sizeof(DevPrime),
                     hipMemcpyHostToDevice));
    ctx->plans.insert(plan.get());
    *out = plan.release();
    return CKKS_OK;
}

int ckks_rns_plan_destroy(ckks_rns_plan *plan) {
    if (!plan) return CKKS_OK;
    hipSetDevice(plan->ctx->device);
    hipStreamSynchronize(plan->ctx->stream);
    plan->ctx->plans.erase(plan);
    hipFree(plan->d_primes);
    hipFree(plan->d_rescale);
    for (auto &kv : plan->crt) hipFree(kv.second.first);
    for (auto &kv : plan->moddown) {
        hipFree(kv.second.first);
        hipFree(kv.second.second);
    }
    delete plan;
    return CKKS_OK;
}

int ckks_rns_plan_psi(const ckks_rns_plan *plan, uint64_t *psi_out) {
    if (!plan || !psi_out) return fail(CKKS_ERR_INVALID, "null argument");
    std::memcpy(psi_out, plan->psi.data(), plan->psi.size() * sizeof(uint64_t));
    return CKKS_OK;
}

}  // extern "C"
