// Serving loop on one GPU: continuous batching of encrypted jobs (see
// serve_core.hpp for the scheduler).  This file is the device side:
//   - register blocks from a size-class pool (released blocks are reused, so
//     steady-state serving calls no hipMalloc),
//   - one batched launch per op group: the group's operand registers are
//     gathered into contiguous batch buffers, the existing batched kernels
//     run once over them (key switch, plaintext multiply, add, rescale,
//     multiply-relinearize-rescale), and results are scattered back,
//   - a completion event per request, polled by the scheduler,
//   - at most kMaxStepsInFlight scheduler steps queued ahead of the GPU,
//     so admission follows the device's progress.
// Every launch is on the context stream, so a released register block can
// be handed to the next request at once: its new writes are ordered after
// the old request's last read.
#include <chrono>
#include <deque>
#include <map>
#include <memory>

#include "internal.hpp"
#include "serve_core.hpp"

namespace ck {

constexpr int kCopyItems = 64;
constexpr int kMaxStepsInFlight = 3;

struct CopyList {
    const u64 *src[kCopyItems];
    u64 *dst[kCopyItems];
    u64 elems;
};

__global__ __launch_bounds__(256) void k_copy_list(CopyList l) {
    const u64 *src = l.src[blockIdx.y];
    u64 *dst = l.dst[blockIdx.y];
    const u64 base = (u64)blockIdx.x * 1024 + threadIdx.x;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const u64 idx = base + (u64)i * 256;
        if (idx < l.elems) dst[idx] = src[idx];
    }
}

int ksk_info(const ckks_ksk *k, size_t *q_limbs, size_t *degree);
size_t ksk_bytes(const ckks_ksk *k);

class HipExec {
  public:
    explicit HipExec(ckks_context *ctx) : ctx_(ctx) {}

    ~HipExec() {
        hipStreamSynchronize(ctx_->stream);
        for (void *p : owned_) hipFree(p);
        for (auto &b : batch_) hipFree(b.ptr);
        for (hipEvent_t e : event_pool_) hipEventDestroy(e);
        for (hipEvent_t e : step_events_) hipEventDestroy(e);
    }

    bool alloc(serve::Request &r) {
        auto it = free_.find(r.bytes);
        if (it != free_.end()) {
            r.regs = it->second;
            free_.erase(it);
            return true;
        }
        if (ctx_->capturing) return false;
        void *p = nullptr;
        if (hipMalloc(&p, r.bytes) != hipSuccess) return false;
        owned_.push_back(p);
        r.regs = p;
        return true;
    }

    void release(serve::Request &r) {
        free_.emplace(r.bytes, r.regs);
        r.regs = nullptr;
    }

    static u64 *reg(const serve::Request &r, size_t i) {
        return (u64 *)r.regs + i * 2 * r.job.n_limbs * r.job.degree;
    }

    int load_input(serve::Request &r) {
        // registers are contiguous and input-sized: one copy loads them all
        const size_t bytes = r.job.n_inputs * 2 * r.job.n_limbs * r.job.degree * sizeof(u64);
        CK_HIP(hipMemcpyAsync(reg(r, 0), r.job.input, bytes, hipMemcpyDeviceToDevice, ctx_->stream));
        return CKKS_OK;
    }

    int finish(serve::Request &r) {
        const size_t limbs = r.limbs[r.job.out_reg];
        const size_t bytes = 2 * limbs * r.job.degree * sizeof(u64);
        CK_HIP(hipMemcpyAsync(r.job.output, reg(r, r.job.out_reg), bytes, hipMemcpyDeviceToDevice, ctx_->stream));
        hipEvent_t e;
        if (!event_pool_.empty()) {
            e = event_pool_.back();
            event_pool_.pop_back();
        } else {
            CK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        CK_HIP(hipEventRecord(e, ctx_->stream));
        r.done_event = e;
        return CKKS_OK;
    }

    bool done(serve::Request &r) {
        hipEvent_t e = (hipEvent_t)r.done_event;
        if (hipEventQuery(e) != hipSuccess) return false;
        event_pool_.push_back(e);
        r.done_event = nullptr;
        return true;
    }

    double now_ms() const {
        using namespace std::chrono;
        return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
    }

    // keeps at most kMaxStepsInFlight steps queued ahead of the device
    int end_step() {
        hipEvent_t e;
        if (!event_pool_.empty()) {
            e = event_pool_.back();
            event_pool_.pop_back();
        } else {
            CK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        CK_HIP(hipEventRecord(e, ctx_->stream));
        step_events_.push_back(e);
        while ((int)step_events_.size() > kMaxStepsInFlight) {
            CK_HIP(hipEventSynchronize(step_events_.front()));
            event_pool_.push_back(step_events_.front());
            step_events_.pop_front();
        }
        return CKKS_OK;
    }

    int run(serve::Group &g) {
        const size_t n = g.members.size(), N = g.degree, L = g.limbs;
        const size_t per = 2 * L * N, per_out = 2 * (L - 1) * N;
        const ckks_rns_plan *plan = (const ckks_rns_plan *)g.plan;
        int rc;
        if (g.code == serve::OP_MAC_PLAIN) {
            // per member (each has its own plaintext table), straight on the
            // registers: every output element is written by the thread that
            // read its sources, so dst may be one of them
            for (auto &m : g.members) {
                const serve::Op &o = *m.second;
                std::vector<const u64 *> src(o.b);
                for (uint32_t t = 0; t < o.b; ++t) src[t] = reg(*m.first, o.a + t);
                int r = ckks_multiply_plain_accumulate_ntt(plan, src.data(), o.pts, o.b, reg(*m.first, o.dst), L);
                if (r) return r;
            }
            return CKKS_OK;
        }
        if (g.code == serve::OP_NTT || g.code == serve::OP_INTT) {
            // straight on the registers, up to 16 requests per launch
            for (size_t i0 = 0; i0 < n; i0 += 16) {
                std::vector<NttCall> calls;
                for (size_t i = i0; i < std::min(n, i0 + 16); ++i) {
                    const serve::Op &o = *g.members[i].second;
                    calls.push_back(NttCall{plan, reg(*g.members[i].first, o.a), reg(*g.members[i].first, o.dst),
                                            2, L, L * N, L * N});
                }
                int r = launch_ntt(ctx_, calls.data(), calls.size(), g.code == serve::OP_NTT);
                if (r) return r;
            }
            return CKKS_OK;
        }
        if (g.code == serve::OP_COPY) {
            std::vector<const u64 *> s(n);
            std::vector<u64 *> d(n);
            for (size_t i = 0; i < n; ++i) {
                s[i] = reg(*g.members[i].first, g.members[i].second->a);
                d[i] = reg(*g.members[i].first, g.members[i].second->dst);
            }
            return copy_list(s, d, per);
        }
        u64 *in = nullptr, *in2 = nullptr, *out = nullptr, *pt = nullptr;
        if ((rc = batch(0, n * per, &in)) || (rc = batch(2, n * per, &out))) return rc;
        if ((rc = gather(g, in, per, [](const serve::Op &o) { return o.a; }))) return rc;
        size_t out_per = per;
        switch (g.code) {
            case serve::OP_ROTATE:
                rc = ckks_apply_galois((const ckks_ksk *)g.key, g.exponent, in, out, n, per, per);
                break;
            case serve::OP_MUL_PLAIN: {
                if ((rc = batch(3, n * L * N, &pt))) return rc;
                std::vector<const u64 *> s(n);
                std::vector<u64 *> d(n);
                for (size_t i = 0; i < n; ++i) {
                    s[i] = g.members[i].second->plaintext;
                    d[i] = pt + i * L * N;
                }
                if ((rc = copy_list(s, d, L * N))) return rc;
                rc = ckks_multiply_plain(plan, in, pt, out, n, L, per, L * N, per);
                break;
            }
            case serve::OP_ADD:
            case serve::OP_SUB:
                if ((rc = batch(1, n * per, &in2))) return rc;
                if ((rc = gather(g, in2, per, [](const serve::Op &o) { return o.b; }))) return rc;
                rc = (g.code == serve::OP_ADD ? ckks_poly_add : ckks_poly_sub)(plan, in, in2, out, 2 * n, L, L * N);
                break;
            case serve::OP_RESCALE:
                rc = ckks_rescale(plan, in, out, n, L, per, per_out);
                out_per = per_out;
                break;
            case serve::OP_MOD_SWITCH:
                rc = ckks_mod_switch(plan, in, out, n, L, per, per_out);
                out_per = per_out;
                break;
            case serve::OP_MUL_RELIN_RESCALE:
                if ((rc = batch(1, n * per, &in2))) return rc;
                if ((rc = gather(g, in2, per, [](const serve::Op &o) { return o.b; }))) return rc;
                rc = ckks_multiply_relinearize_rescale((const ckks_ksk *)g.key, in, in2, out, n, per, per_out);
                out_per = per_out;
                break;
            default:
                return fail(CKKS_ERR_INVALID, "unknown op code");
        }
        if (rc) return rc;
        std::vector<const u64 *> s(n);
        std::vector<u64 *> d(n);
        for (size_t i = 0; i < n; ++i) {
            s[i] = out + i * out_per;
            d[i] = reg(*g.members[i].first, g.members[i].second->dst);
        }
        return copy_list(s, d, out_per);
    }

  private:
    struct Buf {
        void *ptr = nullptr;
        size_t bytes = 0;
    };

    int batch(int slot, size_t elems, u64 **out) {
        Buf &b = batch_[slot];
        const size_t bytes = elems * sizeof(u64);
        if (bytes > b.bytes) {
            if (ctx_->capturing) return fail(CKKS_ERR_INVALID, "serving buffers cannot grow during graph capture");
            CK_HIP(hipStreamSynchronize(ctx_->stream));
            if (b.ptr) CK_HIP(hipFree(b.ptr));
            b.ptr = nullptr;
            b.bytes = 0;
            CK_HIP(hipMalloc(&b.ptr, bytes));
            b.bytes = bytes;
        }
        *out = (u64 *)b.ptr;
        return CKKS_OK;
    }

    template <class Pick>
    int gather(serve::Group &g, u64 *dst, size_t per, Pick pick) {
        const size_t n = g.members.size();
        std::vector<const u64 *> s(n);
        std::vector<u64 *> d(n);
        for (size_t i = 0; i < n; ++i) {
            s[i] = reg(*g.members[i].first, pick(*g.members[i].second));
            d[i] = dst + i * per;
        }
        return copy_list(s, d, per);
    }

    int copy_list(const std::vector<const u64 *> &s, const std::vector<u64 *> &d, size_t elems) {
        for (size_t i0 = 0; i0 < s.size(); i0 += kCopyItems) {
            CopyList l{};
            const size_t m = std::min<size_t>(kCopyItems, s.size() - i0);
            for (size_t i = 0; i < m; ++i) {
                l.src[i] = s[i0 + i];
                l.dst[i] = d[i0 + i];
            }
            l.elems = elems;
            const dim3 grid((u32)((elems + 1023) / 1024), (u32)m);
            hipLaunchKernelGGL(k_copy_list, grid, dim3(256), 0, ctx_->stream, l);
        }
        CK_HIP(hipGetLastError());
        return CKKS_OK;
    }

    ckks_context *ctx_;
    std::multimap<size_t, void *> free_;
    std::vector<void *> owned_;
    Buf batch_[4];
    std::vector<hipEvent_t> event_pool_;
    std::deque<hipEvent_t> step_events_;
};

}  // namespace ck

struct ckks_server {
    ckks_context *ctx;
    std::unique_ptr<ck::HipExec> ex;
    std::unique_ptr<ck::serve::Scheduler<ck::HipExec>> sched;
    std::map<const void *, ck::serve::KeyInfo> keys;
};

using namespace ck;

extern "C" {

int ckks_server_create(ckks_context *ctx, size_t hbm_budget_bytes, size_t max_active, ckks_server **out) {
    if (!ctx || !out) return fail(CKKS_ERR_INVALID, "null argument");
    if (hbm_budget_bytes == 0) return fail(CKKS_ERR_INVALID, "serving needs a nonzero HBM budget");
    std::unique_ptr<ckks_server> s(new ckks_server());
    s->ctx = ctx;
    s->ex.reset(new HipExec(ctx));
    s->sched.reset(new serve::Scheduler<HipExec>(*s->ex, hbm_budget_bytes, max_active));
    ctx->servers.insert(s.get());
    *out = s.release();
    return CKKS_OK;
}

int ckks_server_submit(ckks_server *s, const ckks_job *job, uint64_t *ticket) {
    if (!s || !job || !ticket || (job->n_ops && !job->ops)) return fail(CKKS_ERR_INVALID, "null argument");
    if (!job->plan) return fail(CKKS_ERR_INVALID, "job needs a plan");
    if (job->plan->ctx != s->ctx) return fail(CKKS_ERR_INVALID, "job plan belongs to another context");
    if (job->n_limbs > job->plan->moduli.size()) return fail(CKKS_ERR_INVALID, "job limbs exceed its plan");
    serve::Job j;
    j.plan = job->plan;
    j.degree = job->plan->degree;
    j.n_limbs = job->n_limbs;
    j.n_regs = job->n_regs;
    j.out_reg = job->out_reg;
    j.n_inputs = job->n_inputs ? job->n_inputs : 1;
    j.input = job->input;
    j.output = job->output;
    j.ops.resize(job->n_ops);
    for (size_t i = 0; i < job->n_ops; ++i) {
        const ckks_op &o = job->ops[i];
        j.ops[i] = serve::Op{o.code, o.dst, o.a, o.b, o.key, o.exponent, o.plaintext};
        if (o.code == serve::OP_MAC_PLAIN && o.plaintext && o.b && o.b <= serve::kMaxMacTerms) {
            // the host table of the terms' plaintext pointers, copied
            const auto *tab = reinterpret_cast<const uint64_t *const *>(o.plaintext);
            j.tables.emplace_back(tab, tab + o.b);
            j.ops[i].plaintext = nullptr;
        }
    }
    // tables are complete (no more reallocation of the outer vector): point
    // the ops at them
    for (size_t i = 0, t = 0; i < job->n_ops; ++i)
        if (job->ops[i].code == serve::OP_MAC_PLAIN && job->ops[i].plaintext && job->ops[i].b &&
            job->ops[i].b <= serve::kMaxMacTerms)
            j.ops[i].pts = j.tables[t++].data();
    auto lookup = [&](const void *k) -> const serve::KeyInfo * {
        auto it = s->keys.find(k);
        if (it != s->keys.end()) return &it->second;
        serve::KeyInfo ki{};
        if (ksk_info((const ckks_ksk *)k, &ki.q_limbs, &ki.degree)) return nullptr;
        ki.bytes = ksk_bytes((const ckks_ksk *)k);
        return &(s->keys[k] = ki);
    };
    const std::string why = serve::validate(j, lookup);
    if (!why.empty()) {
        ++s->sched->stats().rejected;
        return fail(CKKS_ERR_INVALID, "serving job rejected: " + why);
    }
    j.resident = serve::resident_set(j, lookup);
    if (!s->sched->fits_ever(j)) {
        ++s->sched->stats().rejected;
        return fail(CKKS_ERR_OOM, "serving job footprint exceeds the HBM budget");
    }
    *ticket = s->sched->submit(std::move(j));
    return CKKS_OK;
}

int ckks_server_step(ckks_server *s, size_t *launched) {
    if (!s) return fail(CKKS_ERR_INVALID, "null server");
    CK_HIP(hipSetDevice(s->ctx->device));
    int rc = s->sched->step(launched);
    if (rc) return rc;
    return s->ex->end_step();
}

int ckks_server_poll(ckks_server *s, uint64_t ticket, int *state, double *latency_ms) {
    if (!s || !state) return fail(CKKS_ERR_INVALID, "null argument");
    s->sched->retire();
    const serve::Request *r = s->sched->find(ticket);
    if (!r) return fail(CKKS_ERR_INVALID, "unknown ticket");
    switch (r->state) {
        case serve::RS_WAITING: *state = 0; break;
        case serve::RS_DONE: *state = 2; break;
        case serve::RS_CANCELLED: *state = 3; break;
        case serve::RS_FAILED: *state = 4; break;
        default: *state = 1;
    }
    if (latency_ms) *latency_ms = r->state == serve::RS_DONE ? r->t_done - r->t_submit : -1.0;
    return CKKS_OK;
}

int ckks_server_cancel(ckks_server *s, uint64_t ticket) {
    if (!s) return fail(CKKS_ERR_INVALID, "null server");
    if (!s->sched->find(ticket)) return fail(CKKS_ERR_INVALID, "unknown ticket");
    if (!s->sched->cancel(ticket)) return fail(CKKS_ERR_INVALID, "request already finished");
    return CKKS_OK;
}

int ckks_server_drain(ckks_server *s) {
    if (!s) return fail(CKKS_ERR_INVALID, "null server");
    while (s->sched->has_work()) {
        int rc = ckks_server_step(s, nullptr);
        if (rc) return rc;
    }
    CK_HIP(hipStreamSynchronize(s->ctx->stream));
    s->sched->retire();
    return CKKS_OK;
}

int ckks_server_stats(ckks_server *s, ckks_server_statistics *out) {
    if (!s || !out) return fail(CKKS_ERR_INVALID, "null argument");
    s->sched->retire();
    auto &st = s->sched->stats();
    out->submitted = st.submitted;
    out->completed = st.completed;
    out->rejected = st.rejected;
    out->waiting = s->sched->waiting();
    out->active = s->sched->active();
    out->peak_active = st.peak_active;
    out->launches = st.launches;
    out->steps = st.steps;
    out->peak_bytes = st.peak_bytes;
    out->p50_ms = serve::percentile(st.latencies_ms, 50);
    out->p99_ms = serve::percentile(st.latencies_ms, 99);
    double sum = 0, mx = 0;
    for (double v : st.latencies_ms) {
        sum += v;
        mx = std::max(mx, v);
    }
    out->mean_ms = st.latencies_ms.empty() ? 0 : sum / (double)st.latencies_ms.size();
    out->max_ms = mx;
    out->cancelled = st.cancelled;
    out->failed = st.failed;
    out->resident_bytes = s->sched->resident_bytes();
    return CKKS_OK;
}

int ckks_server_destroy(ckks_server *s) {
    if (!s) return CKKS_OK;
    hipSetDevice(s->ctx->device);
    hipStreamSynchronize(s->ctx->stream);
    s->ctx->servers.erase(s);
    delete s;
    return CKKS_OK;
}

}  // extern "C"
