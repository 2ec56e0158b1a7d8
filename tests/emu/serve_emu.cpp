// CPU build of the serving scheduler (ccmm-rs_amd/csrc/serve_core.hpp) with
// a recording executor: launches are counted instead of run, completion
// markers are reached `lag` scheduler steps after a request finishes, and
// an op code can be made to fail.  Test infrastructure only.
#include <cstring>
#include <string>
#include <vector>

#include "../../ccmm-rs_amd/csrc/serve_core.hpp"

using namespace ck::serve;

struct FakeExec {
    int lag = 1, fail_code = 0;
    int allocs = 0, fail_allocs_from = -1, fail_allocs_to = -1;  // allocation indices that fail
    double clock = 0;
    long launches = 0, group_members = 0;
    std::vector<uint64_t> done_order;
    std::vector<std::pair<Request *, int>> pending;  // request, steps left
    char pool[1 << 16];
    bool alloc(Request &r) {
        const int i = allocs++;
        if (i >= fail_allocs_from && i < fail_allocs_to) return false;
        r.regs = pool;
        return true;
    }
    void release(Request &) {}
    int load_input(Request &) { return 0; }
    int run(Group &g) {
        if ((int)g.code == fail_code) return 7;
        ++launches;
        group_members += (long)g.members.size();
        return 0;
    }
    int finish(Request &r) {
        pending.emplace_back(&r, lag);
        return 0;
    }
    bool done(Request &r) {
        for (auto &p : pending)
            if (p.first == &r && p.second <= 0) {
                done_order.push_back(r.ticket);
                return true;
            }
        return false;
    }
    void tick() {
        clock += 1.0;
        for (auto &p : pending) --p.second;
    }
    double now_ms() const { return clock; }
};

static KeyInfo kKey{3, 8, 1000};
static const KeyInfo *lookup(const void *k) { return k == (const void *)&kKey ? &kKey : nullptr; }
static uint64_t dummy_in[1], dummy_out[1];

static Job make_job(int kind, size_t regs) {
    Job j{};
    j.plan = &kKey;
    j.degree = 8;
    j.n_limbs = 3;
    j.n_regs = regs;
    j.out_reg = regs - 1;
    j.input = dummy_in;
    j.output = dummy_out;
    // kind 0: rotate, mul_plain, add, rescale (4 ops); kind 1: add, copy (2 ops)
    if (kind == 0) {
        j.ops.push_back(Op{OP_ROTATE, 1, 0, 0, &kKey, 5, nullptr});
        j.ops.push_back(Op{OP_MUL_PLAIN, 1, 1, 0, nullptr, 0, dummy_in});
        j.ops.push_back(Op{OP_ADD, 1, 1, 0, nullptr, 0, nullptr});
        j.ops.push_back(Op{OP_RESCALE, (uint32_t)(regs - 1), 1, 0, nullptr, 0, nullptr});
    } else {
        j.ops.push_back(Op{OP_ADD, 1, 0, 0, nullptr, 0, nullptr});
        j.ops.push_back(Op{OP_COPY, (uint32_t)(regs - 1), 1, 0, nullptr, 0, nullptr});
    }
    return j;
}

extern "C" {

// Runs n_jobs jobs (kinds[i], submitted at step arrive[i]) under a budget
// of `budget_jobs` 2-register footprints and an active cap; fills
// out[0..7] = completed, peak_active, launches, group_members, steps,
// peak_bytes / footprint, last error code, and done_order into order[].
int serve_emu_run(int n_jobs, const int *kinds, const int *arrive, int budget_jobs, int max_active, int lag,
                  int fail_code, long *out, uint64_t *order) {
    FakeExec ex;
    ex.lag = lag;
    ex.fail_code = fail_code;
    const size_t foot = footprint(make_job(0, 2));
    Scheduler<FakeExec> s(ex, (size_t)budget_jobs * foot, (size_t)max_active);
    int submitted = 0, rc = 0, step = 0;
    while ((submitted < n_jobs || !s.idle()) && step < 10000) {
        for (int i = 0; i < n_jobs; ++i)
            if (arrive[i] == step) {
                Job j = make_job(kinds[i], 2);
                if (!validate(j, lookup).empty()) return -1;
                s.submit(std::move(j));
                ++submitted;
            }
        rc = s.step();
        if (rc) break;
        ex.tick();
        ++step;
    }
    auto &st = s.stats();
    out[0] = (long)st.completed;
    out[1] = (long)st.peak_active;
    out[2] = ex.launches;
    out[3] = ex.group_members;
    out[4] = (long)st.steps;
    out[5] = (long)(st.peak_bytes / foot);
    out[6] = rc;
    for (size_t i = 0; i < ex.done_order.size(); ++i) order[i] = ex.done_order[i];
    return 0;
}

static int run_until_idle(Scheduler<FakeExec> &s, FakeExec &ex, int max_steps = 1000) {
    for (int i = 0; i < max_steps && !s.idle(); ++i) {
        int rc = s.step();
        if (rc) return rc;
        ex.tick();
    }
    return s.idle() ? 0 : -2;
}

static Job resident_job(int kind) {
    Job j = make_job(kind, 2);
    j.resident = resident_set(j, lookup);
    return j;
}

// Scenarios of cancellation, allocation failure and resident accounting;
// out[] receives the scenario's observations (see each case).
int serve_emu_scenario(int which, long *out) {
    FakeExec ex;
    const size_t foot = footprint(make_job(0, 2));
    const size_t shared = 1000 + 3 * 8 * 8;  // the key and the plaintext of a kind-0 job
    switch (which) {
        case 0: {  // cancel a waiting and an active request
            Scheduler<FakeExec> s(ex, foot + shared, 0);  // room for one job, its key and plaintext
            uint64_t t[3];
            for (int i = 0; i < 3; ++i) t[i] = s.submit(resident_job(0));
            s.step();  // admits t0
            out[0] = s.find(t[0])->state;
            out[1] = s.cancel(t[2]);  // waiting -> cancelled at once
            out[2] = s.find(t[2])->state;
            out[3] = s.cancel(t[0]);  // active -> dropped at the next step
            s.step();                  // drops t0, admits t1
            out[4] = s.find(t[0])->state;
            out[5] = s.find(t[1])->state;
            out[6] = run_until_idle(s, ex);
            out[7] = s.find(t[1])->state;
            out[8] = (long)s.stats().completed;
            out[9] = (long)s.stats().cancelled;
            out[10] = (long)s.in_use_bytes();
            out[11] = s.cancel(t[1]);  // finished: refused
            return 0;
        }
        case 1: {  // allocation fails with nothing in flight: the request fails
            ex.fail_allocs_from = 0;
            ex.fail_allocs_to = 1;
            Scheduler<FakeExec> s(ex, 4 * foot + 1000, 0);
            uint64_t a = s.submit(resident_job(0)), b = s.submit(resident_job(1));
            out[0] = run_until_idle(s, ex);
            out[1] = s.find(a)->state;
            out[2] = s.find(b)->state;
            out[3] = (long)s.stats().failed;
            out[4] = (long)s.stats().completed;
            out[5] = (long)s.in_use_bytes();
            return 0;
        }
        case 2: {  // allocation fails while another request runs: it waits
            ex.fail_allocs_from = 1;
            ex.fail_allocs_to = 3;
            Scheduler<FakeExec> s(ex, 4 * foot + 1000, 0);
            uint64_t a = s.submit(resident_job(0)), b = s.submit(resident_job(0));
            s.step();
            out[0] = s.find(a)->state;
            out[1] = s.find(b)->state;  // still waiting
            out[2] = run_until_idle(s, ex);
            out[3] = s.find(b)->state;
            out[4] = (long)s.stats().failed;
            out[5] = (long)s.stats().completed;
            return 0;
        }
        case 3: {  // a shared key is charged once; released with its last reader
            Scheduler<FakeExec> s(ex, 2 * foot + shared, 0);  // two jobs fit only if key and plaintext count once
            s.submit(resident_job(0));
            s.submit(resident_job(0));
            s.step();
            out[0] = (long)s.active();
            out[1] = (long)s.resident_bytes();  // key 1000 + one shared plaintext (3 limbs x 8 x 8 B)
            out[2] = (long)s.in_use_bytes();
            out[3] = run_until_idle(s, ex);
            out[4] = (long)s.resident_bytes();
            out[5] = (long)s.in_use_bytes();
            out[6] = (long)s.fits_ever(resident_job(0));
            Scheduler<FakeExec> tight(ex, foot + shared - 1, 0);  // registers fit, key and plaintext do not
            out[7] = (long)tight.fits_ever(resident_job(0));
            out[8] = (long)foot;
            return 0;
        }
        default: return -1;
    }
}

// Validation message of a malformed program, case 0..6.
int serve_emu_validate(int which, char *msg, int cap) {
    Job j = make_job(0, 2);
    switch (which) {
        case 0: j.ops[0].key = nullptr; break;                   // rotation without key
        case 1: j.ops[0].exponent = 4; break;                    // even exponent
        case 2: j.ops.insert(j.ops.begin(), Op{OP_ADD, 1, 0, 1, nullptr, 0, nullptr}); break;  // reads empty reg
        case 3: j.out_reg = 5; break;                            // outside the file
        case 4: j.ops.push_back(Op{OP_RESCALE, 1, 1, 0, nullptr, 0, nullptr});
                j.ops.push_back(Op{OP_RESCALE, 1, 1, 0, nullptr, 0, nullptr}); break;  // past the last level
        case 5: j.ops.push_back(Op{OP_ROTATE, 1, 1, 0, &kKey, 5, nullptr}); break;  // key level mismatch
        case 6: break;                                            // valid
        case 7: j.ops.insert(j.ops.begin(), Op{OP_INTT, 1, 0, 0, nullptr, 0, nullptr}); break;  // INTT of a coefficient reg
        case 8: j.ops.push_back(Op{OP_NTT, 1, 1, 0, nullptr, 0, nullptr});
                j.ops.push_back(Op{OP_ROTATE, 1, 1, 0, &kKey, 5, nullptr}); break;  // rotation of an NTT reg
        case 9: j.ops.push_back(Op{OP_NTT, (uint32_t)(j.n_regs - 1), 1, 0, nullptr, 0, nullptr}); break;  // NTT output
        case 10: j.ops.push_back(Op{OP_MAC_PLAIN, 1, 0, 1, nullptr, 0, nullptr}); break;  // no table
        default: return -1;
    }
    std::string m = validate(j, lookup);
    std::strncpy(msg, m.c_str(), (size_t)cap - 1);
    msg[cap - 1] = 0;
    return (int)m.size();
}

}
