// Internal state shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <set>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ckks_hip.h"
#include "modarith.hpp"
#include "ntt_tile.hpp"

namespace ck {

// Per-limb constants as the kernels see them.
struct DevPrime {
    u64 q;
    const Tw *fwd;  // psi^bitrev(k) with Shoup quotients, k in [0, N)
    const Tw *inv;  // psi^-bitrev(k)
    // FP64 primes of a two-pass degree: row factors (row_factor_table_fp)
    // of the forward and inverse tables, null otherwise
    const Tw *rowf, *rowf_inv;
    u64 ninv, ninv_q;
    u64 br_hi, br_lo;  // floor(2^128 / q)
    double qd, qinv;      // FP64 path: q and 1/q
    double ninv_d, ninv_p;  // FP64 path: N^-1 and N^-1 / q
    u32 logn;
    u32 fp;  // 1: q < 2^41, the NTT runs on the exact FP64 path
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Renumbering so that consecutive logical blocks land on one XCD keeps
// the blocks that share source data (the chunks of one ciphertext, whose
// Galois gathers touch the same lines) behind one L2.
constexpr u32 kXcds = 8;

__device__ __forceinline__ u32 xcd_contiguous_block(u32 b, u32 nb) {
    const u32 full = nb - nb % kXcds;
    return b < full ? (b % kXcds) * (full / kXcds) + b / kXcds : b;
}

// Constants of the exact rescale by a dropped limb l into limb i < l
// (ckks/rns_rescale.rs:107-167).
struct RescaleConst {
    u64 inv, inv_q;  // q_l^-1 mod q_i and its Shoup quotient
    u64 ql_mod;      // q_l mod q_i
};

struct Error {
    int code;
    std::string msg;
};

void set_error(int code, const std::string &msg);
int fail(int code, const std::string &msg);

#define CK_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::ck::fail(CKKS_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct TwiddleTables {
    u64 q;
    u32 logn;
    Tw *fwd = nullptr, *inv = nullptr;  // device
    Tw *rowf = nullptr, *rowf_inv = nullptr;
};

}  // namespace ck

struct ckks_rns_plan;
struct ckks_ntt64_plan;
struct ckks_ksk;
struct ckks_graph;
struct ckks_server;

struct ckks_context {
    int device = 0;
    hipStream_t stream = nullptr;  // nullptr = the null stream
    // plans and keys created from this context, released with it
    std::set<ckks_rns_plan *> plans;
    std::set<ckks_ntt64_plan *> ntt64_plans;
    std::set<ckks_ksk *> keys;
    std::set<ckks_graph *> graphs;
    std::set<ckks_server *> servers;  // destroyed before the keys and plans they use
    // (q, logn) -> device twiddle tables, shared by plans of this context
    std::map<std::pair<uint64_t, uint32_t>, ck::TwiddleTables> tables;
    // grow-only scratch slots: slot 0 is the NTT's inter-pass buffer, the
    // others belong to the composite operations (tensor, key switch).
    // two lanes of slots: lane 1 belongs to the auxiliary stream of a key
    // switch whose batch is split in two (see key_switch_core)
    static constexpr int kSlots = 8;
    void *ws[2 * kSlots] = {};
    size_t ws_bytes[2 * kSlots] = {};
    int lane = 0;
    hipStream_t aux = nullptr;  // default priority, non-blocking
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int ks_split = CKKS_KS_SPLIT_AUTO;
    // NTT inter-pass buffer per chunk in bytes (0 = whole batch in one
    // chunk).  Measured on MI355X over 6144 limb transforms at N = 2^16:
    // 512 MiB chunks run a lone forward transform 7 % faster than one chunk
    // and the inverse equally fast, but the 64 + 16 ciphertext rotation call
    // gained nothing; 32-128 MiB chunks are slower (each chunk's launches
    // drain the GPU).
    size_t ntt_chunk_bytes = 0;
    int base_conversion = CKKS_BASECONV_AUTO;
    uint8_t ntt_variant[32] = {};  // two-pass NTT kernel variant per log2 N
    int ks_fusion = CKKS_KS_FUSION_AUTO;  // digits' row pass fused with the key MAC
    bool capturing = false;  // a hipGraph capture of `stream` is open
    // operations run so far (ckks_profile_get); during a capture the counts
    // go to `captured` and are added on each launch of the graph
    ckks_execution_profile profile = {}, captured = {};
    ckks_execution_profile &counts() { return capturing ? captured : profile; }
    std::vector<void *> owned;  // misc device allocations freed at destroy
    uint64_t *ntt_scratch = nullptr;  // this device's no-allocation NTT scratch

    int workspace(int slot, size_t bytes, void **out);
};

// Keeps trial runs (autotuning) out of the execution profile.
struct ProfileGuard {
    ckks_context *ctx;
    ckks_execution_profile saved;
    explicit ProfileGuard(ckks_context *c) : ctx(c), saved(c->profile) {}
    ~ProfileGuard() { ctx->profile = saved; }
};

// Single-modulus plan for primes up to 64 bits (canonical butterflies).
struct ckks_ntt64_plan {
    ckks_context *ctx = nullptr;
    uint64_t q = 0, psi = 0;
    uint32_t logn = 0;
    size_t degree = 0;
    ck::DevPrime prime{};  // device twiddle tables, N^-1
    ck::Tw *d_fwd = nullptr, *d_inv = nullptr;
};

struct ckks_rns_plan {
    ckks_context *ctx = nullptr;
    std::vector<uint64_t> moduli;
    std::vector<uint64_t> psi;
    uint32_t logn = 0;
    size_t degree = 0;
    ck::DevPrime *d_primes = nullptr;  // device array, one per limb
    std::vector<ck::DevPrime> h_primes;
    ck::RescaleConst *d_rescale = nullptr;  // [dropped l][target i], L x L
    // exact-CRT tables of the decoder per limb count: device blob, words
    std::map<size_t, std::pair<uint64_t *, unsigned>> crt;
    // standalone ModDown tables per Q limb count: constants, device table
    std::map<size_t, std::pair<void *, void *>> moddown;
};

namespace ck {

// Key inner product read by a fused inverse NTT in place of its input:
// polynomial p = 2 ct + comp of the call is sum_j digits[ct][j] * key_comp[j]
// (NTT domain, limb by limb; comp 0 takes kb, comp 1 ka).
struct NttMac {
    const u64 *digits;   // [ct][dnum][limbs][N]
    const u64 *kb, *ka;  // [dnum][limbs][N]
    u32 dnum;            // 1..8
    // hoisted rotations: a [limbs][N] table per component added to every
    // ciphertext's sum (the canonical-lift correction), null otherwise
    const u64 *tb = nullptr, *ta = nullptr;
};

// Launch helpers (ntt.hip).
struct NttCall {
    const ckks_rns_plan *plan;
    const u64 *in;
    u64 *out;
    size_t n_polys, n_limbs;
    size_t in_stride, out_stride;  // elements between consecutive polynomials
    const NttMac *mac = nullptr;   // inverse two-pass degrees only: see NttMac
    // inverse only, 0 = off: the transform reads its input through the
    // NTT-domain Galois permutation by this odd exponent (in != out); with
    // mac, the key inner product is what the permutation reads
    u64 galois_k = 0;
};
// scratch / scratch_elems: an inter-pass buffer supplied by the caller
// (the batch is cut into chunks that fit it) instead of workspace slot 0.
int launch_ntt(ckks_context *ctx, const NttCall *calls, size_t count, bool forward, u64 *ext_scratch = nullptr,
               size_t scratch_elems = 0);
// Device address of the no-allocation NTT scratch array on the current
// device (looked up once per device, at context creation).
int ntt_scratch_lookup(int device, u64 **out);
int ntt_variant_for(const ckks_context *ctx, u32 logn);
// Forward NTT of key-switch digits [ct][dnum][limbs][N] fused with the key
// inner product into acc [ct][2][limbs][N] (two-pass degrees only).
int launch_ntt_mac(ckks_context *ctx, const ckks_rns_plan *plan, const u64 *digits, size_t n_cts,
                   size_t dnum, size_t limbs, const u64 *kb, const u64 *ka, u64 *acc,
                   bool xcd_order);

int check_plan_limbs(const ckks_rns_plan *plan, size_t n_limbs, const char *what);

}  // namespace ck
