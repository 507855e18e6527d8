// Serving benchmark driver: Poisson arrivals of mixed NTT / INTT /
// element-wise requests through the continuous-batching server.
//
// Requests: 40 % forward NTT, 40 % inverse NTT, 20 % EltwiseMulMod; degree
// N = 2^U{13..17}, limbs L = U{1..48}; moduli are distinct NTT primes of
// 50..60 bits per (N, limb); operands are device-resident (the ciphertext
// data of an HE pipeline stays in HBM), drawn from a pre-filled arena.
//
// Phase 1 measures capacity: a closed loop keeps `inflight` requests
// outstanding.  Phase 2 offers Poisson load at fractions of that capacity
// and reports arrival-to-completion p50 / p99 and limbs/s.  Prints one JSON
// object on stdout.
//
//   axhel_serve_bench [--seconds S] [--fractions 0.5,0.8] [--slots K] [--seed X]
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/axhel_hip.h"

namespace {

using Clock = std::chrono::steady_clock;
typedef unsigned __int128 u128;

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }
uint64_t powmod(uint64_t a, uint64_t e, uint64_t m) {
    uint64_t r = 1;
    a %= m;
    while (e) {
        if (e & 1) r = mulmod(r, a, m);
        a = mulmod(a, a, m);
        e >>= 1;
    }
    return r;
}
// Deterministic Miller-Rabin for 64-bit n (bases of tests/common/parameters.hpp:29-51).
bool is_prime(uint64_t n) {
    if (n < 2) return false;
    for (uint64_t p : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 17ull, 19ull, 23ull, 29ull, 31ull, 37ull})
        if (n % p == 0) return n == p;
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) {
        d >>= 1;
        ++s;
    }
    for (uint64_t a : {2ull, 325ull, 9375ull, 28178ull, 450775ull, 9780504ull, 1795265022ull}) {
        if (a % n == 0) continue;
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int r = 1; r < s; ++r) {
            x = mulmod(x, x, n);
            if (x == n - 1) {
                comp = false;
                break;
            }
        }
        if (comp) return false;
    }
    return true;
}
// Smallest-base primitive 2N-th root (parameters.hpp:110-122).
uint64_t root_2n(uint64_t n, uint64_t q) {
    const uint64_t e = (q - 1) / (2 * n);
    for (uint64_t a = 2; a < 100000; ++a) {
        const uint64_t psi = powmod(a, e, q);
        if (psi != 1 && powmod(psi, n, q) == q - 1) return psi;
    }
    return 0;
}

constexpr int kMinLog = 13, kMaxLog = 17, kMaxLimbs = 48, kArenaSlots = 64;

struct Work {
    int op;
    uint32_t logn, limbs;
};

struct Bench {
    axhel_hip_server *srv = nullptr;
    std::vector<std::vector<uint64_t>> primes;                 // [logn][limb]
    std::vector<std::vector<const axhel_hip_plan *>> plans;    // [logn][limb]
    uint64_t *arena_in = nullptr, *arena_out = nullptr, *arena_b = nullptr;
    size_t slot_words = 0;
    std::mt19937_64 rng;
    std::mutex mu;
    std::vector<bool> slot_busy;
    std::atomic<uint64_t> limbs_done{0};
};

Work draw(Bench &B) {
    std::lock_guard<std::mutex> lk(B.mu);
    Work w;
    const uint64_t u = B.rng() % 10;
    w.op = u < 4 ? AXHEL_HIP_OP_NTT_FORWARD : (u < 8 ? AXHEL_HIP_OP_NTT_INVERSE : AXHEL_HIP_OP_MUL);
    w.logn = kMinLog + (uint32_t)(B.rng() % (kMaxLog - kMinLog + 1));
    w.limbs = 1 + (uint32_t)(B.rng() % kMaxLimbs);
    return w;
}

int take_slot(Bench &B) {
    std::lock_guard<std::mutex> lk(B.mu);
    for (int i = 0; i < kArenaSlots; ++i)
        if (!B.slot_busy[i]) {
            B.slot_busy[i] = true;
            return i;
        }
    return -1;
}

void give_slot(Bench &B, int i) {
    std::lock_guard<std::mutex> lk(B.mu);
    B.slot_busy[i] = false;
}

axhel_hip_status submit(Bench &B, const Work &w, int slot, uint64_t *ticket) {
    axhel_hip_request r;
    std::memset(&r, 0, sizeof(r));
    r.op = w.op;
    r.limbs = w.limbs;
    r.degree = 1ull << w.logn;
    r.plans = B.plans[w.logn].data();
    r.moduli = B.primes[w.logn].data();
    r.operand1 = B.arena_in + (size_t)slot * B.slot_words;
    r.result = B.arena_out + (size_t)slot * B.slot_words;
    r.operand2 = w.op == AXHEL_HIP_OP_MUL ? B.arena_b + (size_t)slot * B.slot_words : nullptr;
    r.input_mod_factor = 1;
    r.output_mod_factor = 1;
    return axhel_hip_server_submit(B.srv, &r, ticket);
}

struct Pending {
    uint64_t ticket;
    int slot;
    uint32_t limbs;
};

// Collector: waits tickets in submission order and frees arena slots.
void collect(Bench &B, std::deque<Pending> &q, std::mutex &qm, std::atomic<bool> &done_submitting) {
    for (;;) {
        Pending p;
        {
            std::lock_guard<std::mutex> lk(qm);
            if (q.empty()) {
                if (done_submitting) return;
                p.ticket = 0;
            } else {
                p = q.front();
                q.pop_front();
            }
        }
        if (!p.ticket) {
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            continue;
        }
        if (axhel_hip_server_wait(B.srv, p.ticket, -1) == AXHEL_HIP_OK) B.limbs_done += p.limbs;
        give_slot(B, p.slot);
    }
}

double capacity(Bench &B, double seconds, int inflight) {
    std::deque<Pending> q;
    std::mutex qm;
    std::atomic<bool> done{false};
    std::thread col(collect, std::ref(B), std::ref(q), std::ref(qm), std::ref(done));
    uint64_t n = 0;
    const auto t0 = Clock::now();
    while (std::chrono::duration<double>(Clock::now() - t0).count() < seconds) {
        size_t out;
        {
            std::lock_guard<std::mutex> lk(qm);
            out = q.size();
        }
        if ((int)out >= inflight) {
            std::this_thread::sleep_for(std::chrono::microseconds(10));
            continue;
        }
        const int slot = take_slot(B);
        if (slot < 0) {
            std::this_thread::sleep_for(std::chrono::microseconds(10));
            continue;
        }
        const Work w = draw(B);
        uint64_t t = 0;
        if (submit(B, w, slot, &t) != AXHEL_HIP_OK) {
            give_slot(B, slot);
            continue;
        }
        std::lock_guard<std::mutex> lk(qm);
        q.push_back({t, slot, w.limbs});
        ++n;
    }
    done = true;
    col.join();
    return (double)n / std::chrono::duration<double>(Clock::now() - t0).count();
}

std::string poisson(Bench &B, double rate, double seconds) {
    std::deque<Pending> q;
    std::mutex qm;
    std::atomic<bool> done{false};
    axhel_hip_server_reset_stats(B.srv);
    B.limbs_done = 0;
    std::thread col(collect, std::ref(B), std::ref(q), std::ref(qm), std::ref(done));
    std::exponential_distribution<double> gap(rate);
    std::mt19937_64 arr_rng(12345);
    uint64_t n = 0, stalls = 0;
    const auto t0 = Clock::now();
    double next = 0;
    while (next < seconds) {
        const auto at = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(next));
        while (Clock::now() < at) {
        }
        int slot;
        while ((slot = take_slot(B)) < 0) {
            ++stalls;
            std::this_thread::sleep_for(std::chrono::microseconds(5));
        }
        const Work w = draw(B);
        uint64_t t = 0;
        if (submit(B, w, slot, &t) == AXHEL_HIP_OK) {
            std::lock_guard<std::mutex> lk(qm);
            q.push_back({t, slot, w.limbs});
            ++n;
        } else {
            give_slot(B, slot);
        }
        next += gap(arr_rng);
    }
    done = true;
    col.join();
    axhel_hip_serve_stats st;
    axhel_hip_server_stats(B.srv, &st);
    const double span = st.first_arrival_to_last_completion_s;
    char buf[512];
    std::snprintf(buf, sizeof(buf),
                  "{\"offered_req_per_s\": %.1f, \"seconds\": %.2f, \"requests\": %llu, \"completed\": %llu, "
                  "\"p50_us\": %.1f, \"p99_us\": %.1f, \"mean_us\": %.1f, \"max_us\": %.1f, "
                  "\"limbs_per_s\": %.1f, \"batches\": %llu, \"arena_stalls\": %llu}",
                  rate, seconds, (unsigned long long)n, (unsigned long long)st.completed, st.p50_us, st.p99_us,
                  st.mean_us, st.max_us, span > 0 ? (double)st.limbs_completed / span : 0.0,
                  (unsigned long long)st.batches, (unsigned long long)stalls);
    return buf;
}

}  // namespace

int main(int argc, char **argv) {
    double seconds = 5.0;
    std::vector<double> fractions = {0.5, 0.8};
    uint32_t slots = 4;
    uint64_t seed = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--seconds" && i + 1 < argc) seconds = std::atof(argv[++i]);
        else if (a == "--slots" && i + 1 < argc) slots = (uint32_t)std::atoi(argv[++i]);
        else if (a == "--seed" && i + 1 < argc) seed = std::strtoull(argv[++i], nullptr, 10);
        else if (a == "--fractions" && i + 1 < argc) {
            fractions.clear();
            std::string s = argv[++i];
            size_t p = 0;
            while (p < s.size()) {
                size_t c = s.find(',', p);
                if (c == std::string::npos) c = s.size();
                fractions.push_back(std::atof(s.substr(p, c - p).c_str()));
                p = c + 1;
            }
        }
    }
    Bench B;
    B.rng.seed(seed);
    B.slot_busy.assign(kArenaSlots, false);
    axhel_hip_serve_config cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.slots = slots;
    if (axhel_hip_server_create(&cfg, &B.srv) != AXHEL_HIP_OK) {
        std::fprintf(stderr, "server_create failed: %s\n", axhel_hip_last_error());
        return 1;
    }
    // moduli and resident plans: 48 distinct primes per degree, 50..60 bits
    B.primes.resize(kMaxLog + 1);
    B.plans.resize(kMaxLog + 1);
    for (int logn = kMinLog; logn <= kMaxLog; ++logn) {
        const uint64_t n = 1ull << logn, step = 2 * n;
        std::vector<uint64_t> cursor(61, 0);
        for (int l = 0; l < kMaxLimbs; ++l) {
            const int bits = 50 + l % 11;
            uint64_t q = cursor[bits];
            if (!q) {
                const uint64_t hi = (1ull << bits) - 1;
                q = hi - ((hi - 1) % step);
            } else {
                q -= step;
            }
            while (!is_prime(q)) q -= step;
            cursor[bits] = q;
            B.primes[logn].push_back(q);
            const axhel_hip_plan *p = nullptr;
            if (axhel_hip_server_plan(B.srv, n, q, root_2n(n, q), &p) != AXHEL_HIP_OK) {
                std::fprintf(stderr, "plan failed\n");
                return 1;
            }
            B.plans[logn].push_back(p);
        }
    }
    // device arena: residues below 2^49 (under every modulus), device-resident
    B.slot_words = (size_t)kMaxLimbs << kMaxLog;
    const size_t bytes = B.slot_words * kArenaSlots * sizeof(uint64_t);
    if (hipMalloc(&B.arena_in, bytes) != hipSuccess || hipMalloc(&B.arena_out, bytes) != hipSuccess ||
        hipMalloc(&B.arena_b, bytes) != hipSuccess) {
        std::fprintf(stderr, "arena allocation failed\n");
        return 1;
    }
    {
        std::vector<uint64_t> h(B.slot_words);
        std::mt19937_64 g(seed + 1);
        for (auto &x : h) x = g() & ((1ull << 49) - 1);
        for (int i = 0; i < kArenaSlots; ++i) {
            (void)hipMemcpy(B.arena_in + (size_t)i * B.slot_words, h.data(), B.slot_words * 8, hipMemcpyHostToDevice);
            (void)hipMemcpy(B.arena_b + (size_t)i * B.slot_words, h.data(), B.slot_words * 8, hipMemcpyHostToDevice);
        }
    }
    // warm-up and capacity
    capacity(B, 1.0, 32);
    B.limbs_done = 0;
    axhel_hip_server_reset_stats(B.srv);
    const double cap = capacity(B, seconds, 32);
    axhel_hip_serve_stats cst;
    axhel_hip_server_stats(B.srv, &cst);
    std::fprintf(stderr, "capacity: %.1f req/s over %.0f s\n", cap, seconds);
    const double cap_limbs = cst.first_arrival_to_last_completion_s > 0
                                 ? (double)cst.limbs_completed / cst.first_arrival_to_last_completion_s
                                 : 0.0;
    std::string loads;
    for (double f : fractions) {
        if (!loads.empty()) loads += ", ";
        std::string one = poisson(B, f * cap, seconds);
        char head[64];
        std::snprintf(head, sizeof(head), "{\"fraction\": %.2f, ", f);
        loads += head + one.substr(1);
        std::fprintf(stderr, "load %.2f: %s\n", f, one.c_str());
    }
    std::printf(
        "{\"mix\": \"40%% NTT fwd, 40%% INTT, 20%% MulMod; N=2^13..2^17, L=1..48, device-resident operands\", "
        "\"slots\": %u, \"capacity_req_per_s\": %.1f, \"capacity_limbs_per_s\": %.1f, \"capacity_p50_us\": %.1f, "
        "\"capacity_p99_us\": %.1f, \"loads\": [%s]}\n",
        slots, cap, cap_limbs, cst.p50_us, cst.p99_us, loads.c_str());
    axhel_hip_server_destroy(B.srv);
    (void)hipFree(B.arena_in);
    (void)hipFree(B.arena_out);
    (void)hipFree(B.arena_b);
    return 0;
}
