"""The sym_kernel.c-compatible batch ABI (mf_wasm_*), checked like the
reference's own test (mfah/wasm/test_wasm.mjs:61-65: payloads (n, depth,
stopDepth, #prefixes) = (5,14,20,16), (6,14,20,16), (7,14,28,64)) -- with the
prefixes and expected counts from the oracle instead of the reference Python."""
import ctypes as C

import numpy as np
import pytest

STRIDE = 112


def _lib(mf):
    return mf._abi.lib()


def _write_cases(mf, cycles, mults, n):
    L = _lib(mf)
    base = C.cast(L.mf_wasm_input_ptr(), C.c_void_p).value
    buf = (C.c_uint8 * (STRIDE * 256)).from_address(base)
    arr = np.frombuffer(buf, dtype=np.uint8)
    arr[:] = 0
    for k, (cyc, m) in enumerate(zip(cycles, mults)):
        rec = arr[k * STRIDE:(k + 1) * STRIDE]
        rec[: n * n] = cyc[: n * n]
        rec[100:108] = np.frombuffer(int(m).to_bytes(8, "little"), dtype=np.uint8)


def _read_out(mf):
    L = _lib(mf)
    out = L.mf_wasm_output_ptr()
    w = [int(out[i]) for i in range(8)]
    count = w[0] | (w[1] << 32) | (w[2] << 64) | (w[3] << 96)
    nodes = w[4] | (w[5] << 32)
    return count, nodes, w[6]


def test_wasm_constants_and_status_codes(mf):
    L = _lib(mf)
    assert L.mf_wasm_case_stride() == 112 and L.mf_wasm_max_cases() == 256
    assert L.mf_wasm_solve_cases(1, 3, 4, 1) == 1        # bad n (sym_kernel.c:252-255)
    assert _read_out(mf) == (0, 0, 1)
    assert L.mf_wasm_solve_cases(10, 3, 4, 1) == 1       # n above MAX_SIDE = 9
    assert L.mf_wasm_solve_cases(9, 3, 82, 1) == 2       # stop > 81
    assert L.mf_wasm_solve_cases(6, 3, 4, 257) == 1      # too many cases
    assert L.mf_wasm_solve_cases(6, 10, 4, 1) == 2       # start > stop (:256-259)
    assert L.mf_wasm_solve_cases(6, 3, 37, 1) == 2       # stop > n^2
    assert _read_out(mf)[2] == 2


@pytest.mark.gpu
@pytest.mark.parametrize("n,depth,stop,k", [(5, 14, 20, 16), (6, 14, 20, 16), (7, 14, 28, 64), (5, 14, 25, 100)])
def test_wasm_payloads_match_oracle(gpu, oracle, n, depth, stop, k):
    prefixes = oracle.sym_prefixes(n, depth)
    step = max(1, prefixes.shape[0] // k)
    chosen = prefixes[::step][:k]
    mults = [1 + (i % 7) * 3 for i in range(len(chosen))]   # exercise the multiplier field
    want_c, want_n = 0, 0
    for cyc, m in zip(chosen, mults):
        c, nodes = oracle.sym_count_to(n, cyc, depth, stop)
        want_c += c * m
        want_n += nodes
    _write_cases(gpu, chosen, mults, n)
    status = _lib(gpu).mf_wasm_solve_cases(n, depth, stop, len(chosen))
    assert status == 0
    assert _read_out(gpu) == (want_c, want_n, 0)


def test_wasm_level_zero_is_one_childless_node_per_case(mf):
    """start_i = 0 < stop_i: get_good(cycle, 0) is empty, so sym_kernel.c
    counts 0 with one node per case and status 0 (sym_kernel.c:183,212-230,
    263-276).  No kernel runs for it."""
    L = _lib(mf)
    _write_cases(mf, [np.zeros(64, dtype=np.uint8)] * 5, [3] * 5, 6)
    assert L.mf_wasm_solve_cases(6, 0, 20, 5) == 0
    assert _read_out(mf) == (0, 5, 0)


@pytest.mark.gpu
def test_wasm_records_stay_spiral_under_another_numbering(gpu, oracle, monkeypatch):
    """The wasm records are in spiral order whatever the process-wide
    numbering (and the environment) say: a 7x7 payload under MF_ORDER=reverse
    and a reverse library numbering still equals the spiral oracle."""
    prefixes = oracle.sym_prefixes(7, 14)
    chosen = prefixes[:: max(1, prefixes.shape[0] // 32)][:32]
    want_c, want_n = 0, 0
    for cyc in chosen:
        c, nodes = oracle.sym_count_to(7, cyc, 14, 28)
        want_c += c
        want_n += nodes
    monkeypatch.setenv("MF_ORDER", "reverse")
    with gpu.sym_numbering(order="reverse"):
        _write_cases(gpu, chosen, [1] * len(chosen), 7)
        assert _lib(gpu).mf_wasm_solve_cases(7, 14, 28, len(chosen)) == 0
    assert _read_out(gpu) == (want_c, want_n, 0)
