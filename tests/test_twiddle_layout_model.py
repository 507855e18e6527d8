"""CPU model of the stacked ROW passes' LDS twiddle layout (TwLayout,
TwLdsT and TwStage::fetch_row in axhel_amd/csrc/ntt_kernels.hpp): with the
kernels' own index functions (Geom::kt, Geom::kr), every twiddle read of
every stage, thread and register finds the entry the untransposed layout
holds at index lp = (kt + kr) >> (p + 1), for every pass width and radix
the kernels instantiate, forward and inverse.  The staged copy is a
permutation of the table's 2^LOG_G - 1 entries."""
import pytest


def spread(o, lo, b):
    return (o & ((1 << lo) - 1)) | ((o >> lo) << (lo + b))


def check(log_g, log_e, fwd):
    nc = (log_g + log_e - 1) // log_e
    last_b = log_g - (nc - 1) * log_e
    B = lambda c: log_e if c < nc - 1 else last_b
    LO = lambda c: (log_g - (c * log_e + B(c))) if fwd else c * log_e
    kr = lambda c, r: spread(r >> B(c), LO(c), B(c)) | ((r & ((1 << B(c)) - 1)) << LO(c))
    kt = lambda c, tid: spread(tid << (log_e - B(c)), LO(c), B(c))

    def chunk_of(p):
        return next(c for c in range(nc) if LO(c) <= p < LO(c) + B(c))

    top = lambda c: LO(c) + B(c) + max(0, log_e - B(c) - LO(c))
    nth = lambda p: log_g - top(chunk_of(p))
    G, E = 1 << log_g, 1 << log_e
    idx = lambda j, lp: ((1 << j) - 1 if fwd else G - (2 << j)) + lp
    table = {log_g - 1 - p: nth(p) for p in range(log_g)}
    slot = {}
    for i in range(G - 1):  # TwStage::fetch_row: slot i <- entry (j, lp_of(j, pos))
        j = (i + 1).bit_length() - 1 if fwd else (G - 1 - i).bit_length() - 1
        pos = i + 1 - (1 << j) if fwd else i - (G - (2 << j))
        nt = table[j]
        nr = j - nt
        slot[i] = (j, ((pos & ((1 << nt) - 1)) << nr) | (pos >> nt))
    assert sorted(slot.values()) == sorted((j, lp) for j in range(log_g) for lp in range(1 << j))
    for c in range(nc):
        for p in range(LO(c), LO(c) + B(c)):
            j = log_g - 1 - p
            nt = nth(p)
            nr = j - nt
            bit = 1 << (p - LO(c))
            for tid in range(G // E):
                for r in range(E):
                    if r & bit:
                        continue
                    pr = kr(c, r) >> (p + 1)
                    th = kt(c, tid) >> (p + 1 + nr)  # TwLdsT::get
                    assert pr < 1 << nr and th < 1 << nt
                    assert slot[idx(j, (pr << nt) | th)] == (j, (kt(c, tid) + kr(c, r)) >> (p + 1)), \
                        (log_g, log_e, fwd, c, p, tid, r)


@pytest.mark.parametrize("log_g", range(2, 11))
@pytest.mark.parametrize("log_e", [1, 2, 3, 4])
def test_transposed_twiddle_layout(log_g, log_e):
    if log_e > log_g:
        pytest.skip("radix wider than the pass")
    for fwd in (True, False):
        check(log_g, log_e, fwd)
