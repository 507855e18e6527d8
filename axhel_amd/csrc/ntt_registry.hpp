// Kernel configurations: the helpers that turn template parameters into
// PassKernel / NttConfig entries, and the per-degree lists, which live in
// several translation units (ntt_cfg_*.hip) so the many kernel
// instantiations compile in parallel.  A kernel instantiated in two units is
// one kernel: both units register the same host stub.
#pragma once

#include <vector>

#include "ntt_host.hpp"
#include "polymul_kernels.hpp"

namespace axhel_host {

using Registry = std::vector<std::vector<NttConfig>>;


template <int LG, int LE, int GPW, bool COL, bool MULTI, bool FWD>
PassKernel pk() {
    PassKernel k;
    k.fn = reinterpret_cast<const void *>(&axhel_dev::ntt_pass_kernel<LG, LE, GPW, FWD, COL, MULTI>);
    if constexpr (FWD)
        k.fn_relaxed = reinterpret_cast<const void *>(&axhel_dev::ntt_pass_kernel<LG, LE, GPW, true, COL, MULTI, true>);
    k.log_g = LG;
    k.log_e = LE;
    k.gpw = GPW;
    k.col = COL;
    k.multi = MULTI;
    k.threads = GPW << (LG - LE);
    constexpr size_t tw = axhel_dev::Lds<LG, LE, GPW, COL>::TW_WORDS;
    k.lds_bytes = ((size_t)axhel_dev::pass_tile_words<LG, LE, GPW, COL, false>() + tw) * sizeof(uint64_t);
    k.lds_bytes_relaxed = ((size_t)axhel_dev::pass_tile_words<LG, LE, GPW, COL, true>() + tw) * sizeof(uint64_t);
    return k;
}

template <int LG, int LE, int GPW>
PassKernel pk_fused() {
    PassKernel k;
    k.fn = reinterpret_cast<const void *>(&axhel_dev::polymul_row_kernel<LG, LE, GPW>);
    k.fn_relaxed = reinterpret_cast<const void *>(&axhel_dev::polymul_row_kernel<LG, LE, GPW, true>);
    k.log_g = LG;
    k.log_e = LE;
    k.gpw = GPW;
    k.col = false;
    k.multi = false;
    k.threads = GPW << (LG - LE);
    k.lds_bytes = (size_t)axhel_dev::Lds<LG, LE, GPW, false>::WORDS * sizeof(uint64_t);
    k.lds_bytes_relaxed = k.lds_bytes;
    return k;
}

template <int LG, int LE, int GPW>
PassKernel pk_fused_stack() {
    PassKernel k;
    k.fn = reinterpret_cast<const void *>(&axhel_dev::polymul_row_stack_kernel<LG, LE, GPW>);
    k.fn_relaxed = reinterpret_cast<const void *>(&axhel_dev::polymul_row_stack_kernel<LG, LE, GPW, true>);
    k.log_g = LG;
    k.log_e = LE;
    k.gpw = GPW;
    k.col = false;
    k.multi = false;
    k.stack = true;
    k.threads = GPW << (LG - LE);
    k.lds_bytes = ((size_t)axhel_dev::Lds<LG, LE, GPW, false>::WORDS + 4 * ((1u << LG) - 1)) * sizeof(uint64_t);
    k.lds_bytes_relaxed = k.lds_bytes;
    return k;
}

template <int LG, int LE, int GPW, bool FWD>
PassKernel pk_stack() {
    PassKernel k;
    k.fn = reinterpret_cast<const void *>(&axhel_dev::ntt_row_stack_kernel<LG, LE, GPW, FWD>);
    if constexpr (FWD)
        k.fn_relaxed = reinterpret_cast<const void *>(&axhel_dev::ntt_row_stack_kernel<LG, LE, GPW, true, true>);
    k.log_g = LG;
    k.log_e = LE;
    k.gpw = GPW;
    k.col = false;
    k.multi = false;
    k.stack = true;
    k.threads = GPW << (LG - LE);
    k.lds_bytes = (size_t)axhel_dev::stack_lds_words<LG, LE, GPW>() * sizeof(uint64_t);
    k.lds_bytes_relaxed = k.lds_bytes;
    return k;
}

template <int LG, int LE, int GPW>
NttConfig single() {
    NttConfig c;
    c.pub.log_elems = LE;
    c.pub.row_groups = GPW;
    c.pub.col_groups = 0;
    c.pub.log_col = 0;
    c.pub.log_elems_col = 0;
    c.npass = 1;
    c.fwd[0] = pk<LG, LE, GPW, false, (GPW > 1), true>();
    c.inv[0] = pk<LG, LE, GPW, false, (GPW > 1), false>();
    if constexpr (GPW == 1 && LG >= 12) c.fused = pk_fused<LG, LE, GPW>();
    return c;
}

// Two passes: COL over the top LC bits, ROW over the low LR bits.
template <int LC, int LEC, int GC, int LR, int LER, int GR>
NttConfig two() {
    NttConfig c;
    c.pub.log_elems = LER;
    c.pub.row_groups = GR;
    c.pub.col_groups = GC;
    c.pub.log_col = LC;
    c.pub.log_elems_col = LEC;
    c.npass = 2;
    c.fwd[0] = pk<LC, LEC, GC, true, false, true>();
    c.fwd[1] = pk<LR, LER, GR, false, false, true>();
    c.inv[0] = pk<LR, LER, GR, false, false, false>();
    c.inv[1] = pk<LC, LEC, GC, true, false, false>();
    c.fused = pk_fused<LR, LER, GR>();
    c.fwd_stack = pk_stack<LR, LER, GR, true>();
    c.inv_stack = pk_stack<LR, LER, GR, false>();
    c.fused_stack = pk_fused_stack<LR, LER, GR>();
    return c;
}

// Per-degree configuration lists, in index order (the tuning table's indices).
void add_configs_1_13(Registry &r);
void add_configs_14_15(Registry &r);
void add_configs_16a(Registry &r);
void add_configs_16b(Registry &r);
void add_configs_17(Registry &r);
void add_configs_18_19(Registry &r);
void add_configs_20(Registry &r);

}  // namespace axhel_host
