// Drop-in consumer program: drives the unipi::axhel C++ API with ordinary host
// std::vector buffers, the way the reference's callers do
// (example/example.cpp, tests/public/*.cpp, the SEAL patch).
//
// stdin : N q psi, then N words a[], then N words b[]
// stdout: one line per result, words separated by spaces:
//   kat         1 if the N=8, q=17, psi=3 known-answer round trip holds
//   fwd_lazy    ComputeForward(a, 1, 4)
//   fwd         ComputeForward(a, 1, 1)
//   inv_lazy    ComputeInverse(ReduceMod(fwd_lazy, 4->2), 2, 2)
//   inv         ComputeInverse(fwd, 1, 1)
//   oop         1 if out-of-place forward == in-place forward
//   ext         1 if the external-table NTT gives the same forward output
//   add sub mul fma reduce   the element-wise results on (a, b)
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <vector>

#include "axhel/axhel.hpp"

using namespace unipi::axhel;

static void put(const char *name, const std::vector<uint64_t> &v) {
    std::printf("%s", name);
    for (uint64_t x : v) std::printf(" %llu", (unsigned long long)x);
    std::printf("\n");
}

int main() {
    // tests/public/test_ntt.cpp:32-46
    {
        NTT ntt(8, 17, 3);
        std::vector<uint64_t> x{1, 2, 3, 4, 5, 6, 7, 8};
        ntt.ComputeForward(x.data(), x.data(), 1, 1);
        const bool f = x == std::vector<uint64_t>{5, 0, 13, 8, 9, 11, 5, 8};
        ntt.ComputeInverse(x.data(), x.data(), 1, 1);
        const bool r = x == std::vector<uint64_t>{1, 2, 3, 4, 5, 6, 7, 8};
        std::printf("kat %d\n", (f && r) ? 1 : 0);
    }
    unsigned long long n, q, psi;
    if (!(std::cin >> n >> q >> psi)) return 2;
    std::vector<uint64_t> a(n), b(n);
    for (auto &x : a) {
        unsigned long long v;
        std::cin >> v;
        x = v;
    }
    for (auto &x : b) {
        unsigned long long v;
        std::cin >> v;
        x = v;
    }
    NTT ntt(n, q, psi);
    std::vector<uint64_t> fl = a, fn = a;
    ntt.ComputeForward(fl.data(), fl.data(), 1, 4);
    ntt.ComputeForward(fn.data(), fn.data(), 1, 1);
    put("fwd_lazy", fl);
    put("fwd", fn);
    std::vector<uint64_t> il = fl;
    EltwiseReduceMod(il.data(), il.data(), n, q, 4, 2);
    ntt.ComputeInverse(il.data(), il.data(), 2, 2);
    put("inv_lazy", il);
    std::vector<uint64_t> in = fn;
    ntt.ComputeInverse(in.data(), in.data(), 1, 1);
    put("inv", in);
    std::vector<uint64_t> oop(n);
    ntt.ComputeForward(oop.data(), a.data(), 1, 1);
    std::printf("oop %d\n", oop == fn ? 1 : 0);
    {
        NTT copy = ntt;  // shallow copy shares the plan (ntt.cpp:155-167)
        std::vector<uint64_t> c = a;
        copy.ComputeForward(c.data(), c.data(), 1, 1);
        (void)c;
    }
    {
        // External tables: regenerate through a tiny autonomous plan is not
        // possible from the public API, so rebuild them from the definition.
        typedef unsigned __int128 u128;
        auto mulmod = [&](uint64_t x, uint64_t y) { return (uint64_t)(((u128)x * y) % q); };
        auto powmod = [&](uint64_t base, uint64_t e) {
            uint64_t r = 1;
            while (e) {
                if (e & 1) r = mulmod(r, base);
                base = mulmod(base, base);
                e >>= 1;
            }
            return r;
        };
        unsigned logn = 0;
        while ((1ull << logn) < n) ++logn;
        auto brv = [&](uint64_t i) {
            uint64_t r = 0;
            for (unsigned k = 0; k < logn; ++k) r |= ((i >> k) & 1ull) << (logn - 1 - k);
            return r;
        };
        auto quot = [&](uint64_t w) { return (uint64_t)(((u128)w << 64) / q); };
        std::vector<NTTMultiplyOperand> fw(n), iw(n);
        const uint64_t ipsi = powmod(psi, q - 2);
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t w = powmod(psi, brv(i));
            fw[i] = {w, quot(w)};
        }
        iw[0] = {0, 0};
        for (uint64_t i = 1; i < n; ++i) {
            const uint64_t w = powmod(ipsi, brv(i - 1) + 1);
            iw[i] = {w, quot(w)};
        }
        const uint64_t ni = powmod(n % q, q - 2);
        NTT ext(n, q, fw.data(), iw.data(), NTTMultiplyOperand{ni, quot(ni)});
        std::vector<uint64_t> e = a;
        ext.ComputeForward(e.data(), e.data(), 1, 4);
        std::printf("ext %d\n", e == fl ? 1 : 0);
    }
    std::vector<uint64_t> r(n);
    EltwiseAddMod(r.data(), a.data(), b.data(), n, q);
    put("add", r);
    EltwiseSubMod(r.data(), a.data(), b.data(), n, q);
    put("sub", r);
    EltwiseMulMod(r.data(), a.data(), b.data(), n, q, 1);
    put("mul", r);
    EltwiseFMAMod(r.data(), a.data(), n ? b[0] : 0, b.data(), n, q, 1);
    put("fma", r);
    EltwiseReduceMod(r.data(), b.data(), n, q, q, 1);
    put("reduce", r);
    return 0;
}
