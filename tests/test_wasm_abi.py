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
