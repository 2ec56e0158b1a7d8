// Continuous-batching request scheduler for encrypted jobs, independent of
// the device: the HIP executor (serve.hip) plugs in the memory pool and the
// batched launches; the CPU tests plug in a recording executor.
//
// A job is a straight-line program over ciphertext registers (register 0
// holds the input).  Every scheduler step
//   1. retires requests whose last launch the executor reports complete,
//   2. admits waiting requests in arrival order while their register
//      footprint fits the byte budget and the active-request cap,
//   3. takes the next op of every active request, groups the ops that can
//      run as one batched launch (same op, key, exponent, plan and level)
//      and hands each group to the executor.
// Requests that arrive while others run join at the next step: their ops
// batch with the ops of requests already in flight.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <map>
#include <string>
#include <tuple>
#include <vector>

namespace ck {
namespace serve {

constexpr uint32_t kMaxMacTerms = 64;

enum OpCode : uint32_t {
    OP_ROTATE = 1,             // dst = key switch of sigma_exponent(a)
    OP_MUL_PLAIN = 2,          // dst = a * plaintext (NTT-domain, a's level)
    OP_ADD = 3,                // dst = a + b (same level)
    OP_RESCALE = 4,            // dst = rescale(a): one limb fewer
    OP_MUL_RELIN_RESCALE = 5,  // dst = rescale(relin(a x b)): one limb fewer
    OP_COPY = 6,               // dst = a
    OP_MOD_SWITCH = 7,         // dst = a without its last limb (scale kept)
    OP_SUB = 8,                // dst = a - b (same level)
    OP_NTT = 9,                // dst = NTT(a): the register moves to the NTT domain
    OP_INTT = 10,              // dst = NTT^-1(a): back to the coefficient domain
    OP_MAC_PLAIN = 11,         // dst = sum_{i < b} reg[a + i] * pts[i], NTT domain
};

// Coefficient-domain registers feed every op except OP_INTT and
// OP_MAC_PLAIN, which read NTT-domain ones; ADD / SUB / COPY keep the
// domain of their operands; a program's output is coefficient domain.
enum RegDomain : uint8_t { DOM_COEFF = 0, DOM_NTT = 1 };

struct Op {
    uint32_t code, dst, a, b;
    const void *key;  // ROTATE: Galois key, MUL_RELIN_RESCALE: relinearization key
    uint64_t exponent;
    const uint64_t *plaintext;
    const uint64_t *const *pts = nullptr;  // OP_MAC_PLAIN: b NTT-domain plaintexts (job-owned table)
};

// Key facts the validator needs (the executor resolves them from the key).
struct KeyInfo {
    size_t q_limbs;  // limbs of the ciphertexts the key applies to
    size_t degree;
    size_t bytes = 0;  // device bytes the key holds resident
};

// Device memory a job reads without owning it: a key or a plaintext.  It is
// charged against the budget once while any admitted job reads it
// (evaluation keys stay resident across requests; the accounting model of
// evaluation_key_storage, ckks/key_accounting.rs:276-344, per resident key).
struct Resource {
    const void *ptr;
    size_t bytes;
};

struct Job {
    const void *plan;  // basis of the job's ciphertexts (the chain top)
    size_t degree, n_limbs, n_regs, out_reg, n_inputs = 1;
    const uint64_t *input;
    uint64_t *output;
    std::vector<Op> ops;
    std::vector<Resource> resident;  // distinct keys and plaintexts the ops read
    std::vector<std::vector<const uint64_t *>> tables;  // storage behind Op::pts
};

struct Request {
    uint64_t ticket;
    Job job;
    size_t pc = 0;                  // next op
    std::vector<size_t> limbs;      // current limb count per register
    void *regs = nullptr;           // executor-owned register block
    size_t bytes = 0;               // register footprint
    double t_submit = 0, t_admit = 0, t_done = 0;
    // 0 waiting, 1 active, 2 finishing (launched), 3 done, 4 cancelled,
    // 5 failed (its registers could not be allocated with nothing in flight)
    int state = 0;
    bool cancel = false;            // active: drop at the next step
    void *done_event = nullptr;     // executor-owned completion marker
};

enum RequestState { RS_WAITING = 0, RS_ACTIVE = 1, RS_FINISHING = 2, RS_DONE = 3, RS_CANCELLED = 4, RS_FAILED = 5 };

inline size_t footprint(const Job &j) { return j.n_regs * 2 * j.n_limbs * j.degree * sizeof(uint64_t); }

// Checks a program before it is queued; returns "" or the reason it is
// rejected.  `key_info` maps a key pointer to its facts (nullptr: unknown).
template <class KeyLookup>
std::string validate(const Job &j, KeyLookup &&key_info) {
    if (!j.plan || !j.input || !j.output) return "job needs a plan, an input and an output";
    if (j.n_regs == 0 || j.out_reg >= j.n_regs) return "output register outside the register file";
    if (j.n_limbs == 0 || j.degree == 0) return "job ciphertexts need limbs";
    if (j.n_inputs == 0 || j.n_inputs > j.n_regs) return "inputs must fit the register file";
    std::vector<size_t> limbs(j.n_regs, 0);
    std::vector<uint8_t> dom(j.n_regs, DOM_COEFF);
    for (size_t i = 0; i < j.n_inputs; ++i) limbs[i] = j.n_limbs;
    for (size_t i = 0; i < j.ops.size(); ++i) {
        const Op &o = j.ops[i];
        const std::string at = " (op " + std::to_string(i) + ")";
        if (o.dst >= j.n_regs || o.a >= j.n_regs) return "register index outside the register file" + at;
        if (limbs[o.a] == 0) return "op reads a register that holds no ciphertext" + at;
        const bool ntt_in = o.code == OP_INTT || o.code == OP_MAC_PLAIN;
        const bool any_in = o.code == OP_ADD || o.code == OP_SUB || o.code == OP_COPY;
        if (!any_in && dom[o.a] != (ntt_in ? DOM_NTT : DOM_COEFF))
            return std::string(ntt_in ? "op needs an NTT-domain ciphertext" : "op needs a coefficient-domain ciphertext") + at;
        if ((o.code == OP_ADD || o.code == OP_SUB) && o.b < j.n_regs && dom[o.b] != dom[o.a])
            return "addition needs two ciphertexts in one domain" + at;
        uint8_t dom_out = o.code == OP_NTT || o.code == OP_MAC_PLAIN ? DOM_NTT
                          : any_in                                      ? dom[o.a]
                                                                        : DOM_COEFF;
        switch (o.code) {
            case OP_ROTATE: {
                if (!o.key || !(o.exponent & 1)) return "rotation needs a key and an odd exponent" + at;
                const KeyInfo *ki = key_info(o.key);
                if (!ki) return "unknown key" + at;
                if (ki->q_limbs != limbs[o.a] || ki->degree != j.degree)
                    return "key basis must match the ciphertext level" + at;
                limbs[o.dst] = limbs[o.a];
                break;
            }
            case OP_MUL_PLAIN:
                if (!o.plaintext) return "plaintext multiply needs a plaintext" + at;
                limbs[o.dst] = limbs[o.a];
                break;
            case OP_ADD:
            case OP_SUB:
                if (o.b >= j.n_regs || limbs[o.b] != limbs[o.a]) return "addition needs two ciphertexts at one level" + at;
                limbs[o.dst] = limbs[o.a];
                break;
            case OP_RESCALE:
            case OP_MOD_SWITCH:
                if (limbs[o.a] < 2) return "cannot leave the final level" + at;
                limbs[o.dst] = limbs[o.a] - 1;
                break;
            case OP_MUL_RELIN_RESCALE: {
                if (o.b >= j.n_regs || limbs[o.b] != limbs[o.a]) return "multiplication needs two ciphertexts at one level" + at;
                if (limbs[o.a] < 2) return "cannot rescale the final level" + at;
                const KeyInfo *ki = key_info(o.key);
                if (!o.key || !ki) return "multiplication needs a relinearization key" + at;
                if (ki->q_limbs != limbs[o.a] || ki->degree != j.degree)
                    return "key basis must match the ciphertext level" + at;
                limbs[o.dst] = limbs[o.a] - 1;
                break;
            }
            case OP_COPY:
            case OP_NTT:
            case OP_INTT: limbs[o.dst] = limbs[o.a]; break;
            case OP_MAC_PLAIN: {
                if (o.b == 0 || (size_t)o.a + o.b > j.n_regs || !o.pts)
                    return "plaintext multiply-accumulate needs 1.. registers and a plaintext table" + at;
                if (o.b > kMaxMacTerms) return "plaintext multiply-accumulate of at most 64 terms" + at;
                for (uint32_t t = 0; t < o.b; ++t) {
                    if (limbs[o.a + t] != limbs[o.a] || dom[o.a + t] != DOM_NTT)
                        return "plaintext multiply-accumulate needs NTT-domain registers at one level" + at;
                    if (!o.pts[t]) return "plaintext multiply-accumulate needs a plaintext per term" + at;
                }
                limbs[o.dst] = limbs[o.a];
                break;
            }
            default: return "unknown op code" + at;
        }
        dom[o.dst] = dom_out;
    }
    if (limbs[j.out_reg] == 0) return "output register holds no ciphertext";
    if (dom[j.out_reg] != DOM_COEFF) return "output register must be coefficient domain";
    return "";
}

// The keys and plaintexts a validated job reads, each once, with their
// device bytes (a plaintext holds its op's limb count of residues).
template <class KeyLookup>
std::vector<Resource> resident_set(const Job &j, KeyLookup &&key_info) {
    std::vector<Resource> out;
    auto add = [&](const void *p, size_t bytes) {
        for (const Resource &r : out)
            if (r.ptr == p) return;
        out.push_back(Resource{p, bytes});
    };
    std::vector<size_t> limbs(j.n_regs, 0);
    for (size_t i = 0; i < j.n_inputs; ++i) limbs[i] = j.n_limbs;
    for (const Op &o : j.ops) {
        const size_t la = limbs[o.a];
        if ((o.code == OP_ROTATE || o.code == OP_MUL_RELIN_RESCALE) && o.key) {
            const KeyInfo *ki = key_info(o.key);
            add(o.key, ki ? ki->bytes : 0);
        }
        if (o.code == OP_MUL_PLAIN && o.plaintext) add(o.plaintext, la * j.degree * sizeof(uint64_t));
        if (o.code == OP_MAC_PLAIN && o.pts)
            for (uint32_t t = 0; t < o.b; ++t) add(o.pts[t], la * j.degree * sizeof(uint64_t));
        limbs[o.dst] = (o.code == OP_RESCALE || o.code == OP_MOD_SWITCH || o.code == OP_MUL_RELIN_RESCALE) ? la - 1 : la;
    }
    return out;
}

// Ops of several requests that one batched launch serves.
struct Group {
    uint32_t code;
    const void *key;
    uint64_t exponent;
    const void *plan;
    size_t degree, limbs;
    std::vector<std::pair<Request *, const Op *>> members;
};

struct Stats {
    size_t submitted = 0, completed = 0, rejected = 0, cancelled = 0, failed = 0;
    size_t peak_active = 0, peak_bytes = 0, launches = 0, steps = 0;
    std::vector<double> latencies_ms;
};

// Executor concept:
//   bool alloc(Request&)                    register block of r.bytes
//   void release(Request&)
//   int  load_input(Request&)                register 0 <- job input
//   int  run(Group&)                         one batched launch per group
//   int  finish(Request&)                    output <- out_reg, completion marker
//   bool done(Request&)                      marker reached
//   double now_ms()
template <class Exec>
class Scheduler {
  public:
    Scheduler(Exec &ex, size_t budget_bytes, size_t max_active)
        : ex_(ex), budget_(budget_bytes), max_active_(max_active ? max_active : (size_t)-1) {}

    ~Scheduler() {
        for (auto &kv : all_) delete kv.second;
    }

    // Queues a validated job; the caller validates (it knows the keys).
    uint64_t submit(Job job) {
        auto *r = new Request();
        r->ticket = ++next_ticket_;
        r->job = std::move(job);
        r->bytes = footprint(r->job);
        r->t_submit = ex_.now_ms();
        all_[r->ticket] = r;
        waiting_.push_back(r);
        ++stats_.submitted;
        return r->ticket;
    }

    bool fits_ever(const Job &j) const {
        size_t bytes = footprint(j);
        for (const Resource &res : j.resident) bytes += res.bytes;
        return bytes <= budget_;
    }

    // Cancels a waiting or active request: a waiting one leaves the queue, an
    // active one issues no further op and gives its registers back at the
    // next step (its launched work is ordered before any reuse on the one
    // stream).  Returns false once the request has finished.
    bool cancel(uint64_t ticket) {
        auto it = all_.find(ticket);
        if (it == all_.end()) return false;
        Request *r = it->second;
        if (r->state == RS_WAITING) {
            waiting_.erase(std::find(waiting_.begin(), waiting_.end(), r));
            r->state = RS_CANCELLED;
            ++stats_.cancelled;
            return true;
        }
        if (r->state == RS_ACTIVE) {
            r->cancel = true;
            return true;
        }
        return false;
    }

    // One scheduler iteration; returns an executor error code or 0.
    int step(size_t *launched = nullptr) {
        ++stats_.steps;
        retire();
        for (Request *r : active_)
            if (r->state == RS_ACTIVE && r->cancel) {
                drop(*r);
                r->state = RS_CANCELLED;
                ++stats_.cancelled;
            }
        active_.erase(std::remove_if(active_.begin(), active_.end(), [](Request *r) { return r->state != RS_ACTIVE; }),
                      active_.end());
        // admission in arrival order; a request that does not fit blocks the
        // ones behind it (no overtaking, so large jobs cannot starve)
        while (!waiting_.empty() && active_.size() < max_active_) {
            Request *r = waiting_.front();
            const size_t cost = r->bytes + unaccounted(*r);
            if (used_ + cost > budget_) break;
            if (!ex_.alloc(*r)) {
                // memory frees up only when active requests finish: with none
                // in flight the request cannot ever be admitted
                if (!active_.empty()) break;
                waiting_.pop_front();
                r->state = RS_FAILED;
                ++stats_.failed;
                continue;
            }
            charge(*r);
            stats_.peak_bytes = std::max(stats_.peak_bytes, used_);
            r->limbs.assign(r->job.n_regs, 0);
            for (size_t i = 0; i < r->job.n_inputs; ++i) r->limbs[i] = r->job.n_limbs;
            r->state = RS_ACTIVE;
            r->t_admit = ex_.now_ms();
            waiting_.pop_front();
            active_.push_back(r);
            int rc = ex_.load_input(*r);
            if (rc) return rc;
        }
        stats_.peak_active = std::max(stats_.peak_active, active_.size());
        // the next op of every active request, grouped
        std::map<std::tuple<uint32_t, const void *, uint64_t, const void *, size_t, size_t>, Group> groups;
        std::vector<Request *> finishing;
        for (Request *r : active_) {
            if (r->state != 1) continue;
            if (r->pc >= r->job.ops.size()) {
                finishing.push_back(r);
                continue;
            }
            const Op &o = r->job.ops[r->pc];
            const size_t limbs = r->limbs[o.a];
            auto key = std::make_tuple(o.code, o.code == OP_ROTATE || o.code == OP_MUL_RELIN_RESCALE ? o.key : nullptr,
                                       o.code == OP_ROTATE ? o.exponent : 0, r->job.plan, r->job.degree, limbs);
            Group &g = groups[key];
            if (g.members.empty()) {
                g.code = o.code;
                g.key = std::get<1>(key);
                g.exponent = std::get<2>(key);
                g.plan = r->job.plan;
                g.degree = r->job.degree;
                g.limbs = limbs;
            }
            g.members.emplace_back(r, &o);
        }
        size_t n = 0;
        for (auto &kv : groups) {
            Group &g = kv.second;
            int rc = ex_.run(g);
            if (rc) return rc;
            ++n;
            for (auto &m : g.members) {
                Request *r = m.first;
                const Op &o = *m.second;
                const size_t la = r->limbs[o.a];
                r->limbs[o.dst] =
                    (o.code == OP_RESCALE || o.code == OP_MOD_SWITCH || o.code == OP_MUL_RELIN_RESCALE) ? la - 1 : la;
                ++r->pc;
                if (r->pc >= r->job.ops.size()) finishing.push_back(r);
            }
        }
        for (Request *r : finishing) {
            if (r->state != 1) continue;
            int rc = ex_.finish(*r);
            if (rc) return rc;
            r->state = 2;
            // registers go back now: later launches are ordered after this
            // request's last one on the same stream
            drop(*r);
        }
        active_.erase(std::remove_if(active_.begin(), active_.end(), [](Request *r) { return r->state != 1; }),
                      active_.end());
        for (Request *r : finishing)
            if (r->state == 2) finishing_.push_back(r);
        stats_.launches += n;
        if (launched) *launched = n;
        return 0;
    }

    // Marks finished requests whose completion marker was reached.
    void retire() {
        for (auto it = finishing_.begin(); it != finishing_.end();) {
            Request *r = *it;
            if (ex_.done(*r)) {
                r->state = 3;
                r->t_done = ex_.now_ms();
                stats_.latencies_ms.push_back(r->t_done - r->t_submit);
                ++stats_.completed;
                it = finishing_.erase(it);
            } else {
                ++it;
            }
        }
    }

    bool idle() const { return waiting_.empty() && active_.empty() && finishing_.empty(); }
    bool has_work() const { return !waiting_.empty() || !active_.empty(); }

    const Request *find(uint64_t ticket) const {
        auto it = all_.find(ticket);
        return it == all_.end() ? nullptr : it->second;
    }

    size_t waiting() const { return waiting_.size(); }
    size_t active() const { return active_.size(); }
    size_t in_use_bytes() const { return used_; }
    size_t resident_bytes() const {
        size_t b = 0;
        for (const auto &kv : resident_) b += kv.second.first;
        return b;
    }
    Stats &stats() { return stats_; }

  private:
    // bytes of r's keys and plaintexts that no admitted request holds yet
    size_t unaccounted(const Request &r) const {
        size_t b = 0;
        for (const Resource &res : r.job.resident)
            if (!resident_.count(res.ptr)) b += res.bytes;
        return b;
    }

    void charge(Request &r) {
        used_ += r.bytes;
        for (const Resource &res : r.job.resident) {
            auto &e = resident_[res.ptr];
            if (e.second++ == 0) {
                e.first = res.bytes;
                used_ += res.bytes;
            }
        }
    }

    // registers back to the executor, resident bytes released with their
    // last reader
    void drop(Request &r) {
        ex_.release(r);
        used_ -= r.bytes;
        for (const Resource &res : r.job.resident) {
            auto it = resident_.find(res.ptr);
            if (it != resident_.end() && --it->second.second == 0) {
                used_ -= it->second.first;
                resident_.erase(it);
            }
        }
    }

    Exec &ex_;
    size_t budget_, max_active_, used_ = 0;
    uint64_t next_ticket_ = 0;
    std::deque<Request *> waiting_;
    std::vector<Request *> active_;
    std::vector<Request *> finishing_;
    std::map<uint64_t, Request *> all_;
    std::map<const void *, std::pair<size_t, size_t>> resident_;  // ptr -> (bytes, readers)
    Stats stats_;
};

inline double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double idx = p / 100.0 * (double)(v.size() - 1);
    const size_t lo = (size_t)idx, hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (idx - (double)lo);
}

}  // namespace serve
}  // namespace ck
