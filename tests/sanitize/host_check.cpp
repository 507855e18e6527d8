// Host-side stress of the serving scheduler (NULL backend: no GPU is touched)
// and of the C++ drop-in's error and no-op paths, built with AddressSanitizer
// or ThreadSanitizer by tests/test_host_sanitizers.py.  The reference builds
// its Debug configuration with AddressSanitizer
// (cmake/axhel/axhel-util.cmake:66-74).
//
// Exercised: many client threads submitting, waiting, cancelling and timing
// out against a small memory budget with injected staging and batch
// failures; two threads waiting on one ticket; shutdown with queued
// requests and a waiter inside axhel_hip_server_wait; unsupported factors
// and invalid arguments through unipi::axhel.  Exit status 0 when every
// request ended with an expected status.
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <thread>
#include <vector>

#include "axhel/axhel.hpp"
#include "axhel_hip.h"

namespace {

std::atomic<int> g_bad{0};

void expect(bool ok, const char *what) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what);
        g_bad.fetch_add(1);
    }
}

axhel_hip_server *make_server(uint64_t budget, uint32_t slots, uint32_t latency_us, uint64_t fail_alloc,
                              uint64_t fail_batch) {
    axhel_hip_serve_config cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.memory_budget_bytes = budget;
    cfg.slots = slots;
    cfg.backend = AXHEL_HIP_BACKEND_NULL;
    cfg.null_latency_us = latency_us;
    cfg.fail_every_alloc = fail_alloc;
    cfg.fail_every_batch = fail_batch;
    axhel_hip_server *s = nullptr;
    expect(axhel_hip_server_create(&cfg, &s) == AXHEL_HIP_OK && s, "server_create");
    return s;
}

struct Buffers {
    std::vector<uint64_t> a, b, r, moduli, scalars;
};

axhel_hip_request make_request(Buffers &B, std::mt19937_64 &rng, int64_t timeout_us) {
    const uint64_t degree = 1ull << (10 + rng() % 6);
    const uint32_t limbs = 1 + rng() % 8;
    B.a.assign(degree * limbs, 1);
    B.b.assign(degree * limbs, 2);
    B.r.assign(degree * limbs, 0);
    B.moduli.assign(limbs, 97);
    B.scalars.assign(limbs, 5);
    axhel_hip_request r;
    std::memset(&r, 0, sizeof(r));
    const int ops[] = {AXHEL_HIP_OP_ADD, AXHEL_HIP_OP_SUB, AXHEL_HIP_OP_MUL, AXHEL_HIP_OP_FMA, AXHEL_HIP_OP_REDUCE};
    r.op = ops[rng() % 5];
    r.limbs = limbs;
    r.degree = degree;
    r.moduli = B.moduli.data();
    r.scalars = B.scalars.data();
    r.result = B.r.data();
    r.operand1 = B.a.data();
    r.operand2 = (r.op == AXHEL_HIP_OP_REDUCE) ? nullptr : B.b.data();
    r.input_mod_factor = r.op == AXHEL_HIP_OP_REDUCE ? 4 : 1;
    r.output_mod_factor = 1;
    r.timeout_us = timeout_us;
    return r;
}

void client(axhel_hip_server *s, int id, int count) {
    std::mt19937_64 rng(1234 + id);
    for (int i = 0; i < count; ++i) {
        Buffers B;
        const int64_t timeout = (rng() % 4 == 0) ? 50 : 0;
        axhel_hip_request r = make_request(B, rng, timeout);
        uint64_t t = 0;
        const axhel_hip_status st = axhel_hip_server_submit(s, &r, &t);
        if (st == AXHEL_HIP_REJECTED) continue;
        expect(st == AXHEL_HIP_OK, "submit");
        if (st != AXHEL_HIP_OK) continue;
        if (rng() % 5 == 0) {
            const axhel_hip_status c = axhel_hip_server_cancel(s, t);
            expect(c == AXHEL_HIP_OK || c == AXHEL_HIP_REJECTED, "cancel status");
        }
        // a short wait either reports "still pending" (the ticket stays
        // valid) or collects the final status
        axhel_hip_status w = AXHEL_HIP_PENDING;
        if (rng() % 7 == 0) w = axhel_hip_server_wait(s, t, 10);
        expect(w != AXHEL_HIP_INVALID_ARGUMENT, "short wait");
        if (w == AXHEL_HIP_PENDING) w = axhel_hip_server_wait(s, t, -1);
        const bool known =
            w == AXHEL_HIP_OK || w == AXHEL_HIP_CANCELLED || w == AXHEL_HIP_TIMEOUT || w == AXHEL_HIP_NO_MEMORY;
        if (!known) std::fprintf(stderr, "unexpected final status %d\n", (int)w);
        expect(known, "final status");
        expect(axhel_hip_server_wait(s, t, 0) == AXHEL_HIP_INVALID_ARGUMENT, "ticket released");
    }
}

void stress_scheduler() {
    axhel_hip_server *s = make_server(8ull << 20, 3, 200, 7, 5);
    std::vector<std::thread> th;
    for (int i = 0; i < 8; ++i) th.emplace_back(client, s, i, 150);
    for (auto &t : th) t.join();
    axhel_hip_serve_stats st;
    expect(axhel_hip_server_stats(s, &st) == AXHEL_HIP_OK, "stats");
    expect(st.bytes_in_use == 0, "all staging released");
    expect(st.completed > 0 && st.failed > 0, "both outcomes seen");
    expect(axhel_hip_server_destroy(s) == AXHEL_HIP_OK, "destroy");
}

void two_waiters_and_shutdown() {
    axhel_hip_server *s = make_server(64ull << 20, 1, 20000, 0, 0);
    std::mt19937_64 rng(7);
    Buffers B1, B2, B3;
    axhel_hip_request r1 = make_request(B1, rng, 0), r2 = make_request(B2, rng, 0), r3 = make_request(B3, rng, 0);
    uint64_t t1 = 0, t2 = 0, t3 = 0;
    expect(axhel_hip_server_submit(s, &r1, &t1) == AXHEL_HIP_OK, "submit 1");
    std::atomic<int> ok{0}, invalid{0};
    auto waiter = [&] {
        const axhel_hip_status w = axhel_hip_server_wait(s, t1, -1);
        if (w == AXHEL_HIP_OK) ok.fetch_add(1);
        if (w == AXHEL_HIP_INVALID_ARGUMENT) invalid.fetch_add(1);
    };
    std::thread a(waiter), b(waiter);
    a.join();
    b.join();
    expect(ok == 1 && invalid == 1, "one waiter collects");
    // one running, one queued behind it; destroy while a thread waits on the queued one
    expect(axhel_hip_server_submit(s, &r2, &t2) == AXHEL_HIP_OK, "submit 2");
    expect(axhel_hip_server_submit(s, &r3, &t3) == AXHEL_HIP_OK, "submit 3");
    std::atomic<int> w3{-1};
    std::thread c([&] { w3 = axhel_hip_server_wait(s, t3, -1); });
    std::this_thread::sleep_for(std::chrono::milliseconds(2));
    expect(axhel_hip_server_destroy(s) == AXHEL_HIP_OK, "destroy with work");
    c.join();
    expect(w3 == AXHEL_HIP_CANCELLED || w3 == AXHEL_HIP_OK, "queued request at shutdown");
}

void cxx_paths() {
    using namespace unipi::axhel;
    std::vector<uint64_t> a(64, 3), b(64, 5), r(64, 11);
    EltwiseMulMod(r.data(), a.data(), b.data(), 64, 97, 3);  // unsupported factor: silent no-op
    EltwiseFMAMod(r.data(), a.data(), 7, nullptr, 64, 97, 5);
    EltwiseReduceMod(r.data(), a.data(), 64, 97, 3, 1);
    bool untouched = true;
    for (uint64_t x : r) untouched = untouched && x == 11;
    expect(untouched, "unsupported factors leave the output untouched");
    bool threw = false;
    try {
        NTT bad(1000, 97, 3);  // not a power of two
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    expect(threw, "invalid degree throws invalid_argument");
}

}  // namespace

int main(int argc, char **argv) {
    const bool tsan = argc > 1 && std::strcmp(argv[1], "--scheduler-only") == 0;
    stress_scheduler();
    two_waiters_and_shutdown();
    if (!tsan) cxx_paths();
    if (g_bad.load() == 0) std::printf("host_check ok\n");
    return g_bad.load() == 0 ? 0 : 1;
}
