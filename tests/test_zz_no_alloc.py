"""hevcr_residual_blocks_noalloc (device scratch in a static array, no
allocation) against hevcr_residual_blocks, both called through the C-ABI with
the same host inputs and the same pre-filled outputs: equal outputs, bit for
bit (samples outside every block included), on two calls in a row with the
second smaller, plus the reference's residual KATs; sizes beyond its limits
are rejected."""
import ctypes as C

import numpy as np
import pytest

import hevcr
import oracle_lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    e = hevcr.Engine(0)
    yield e
    e.close()


def _blocks(rng, n, bd):
    log2 = rng.integers(2, 6, n).astype(np.uint8)
    cidx = rng.integers(0, 3, n).astype(np.uint8)
    qp = rng.integers(0, 52 + 6 * (bd - 8), n).astype(np.uint8)
    flags = rng.integers(0, 2, n).astype(np.uint8)
    flags |= ((log2 == 2) & (rng.random(n) < 0.3)).astype(np.uint8) << 1  # transform skip at 4x4 only
    sizes = (1 << (2 * log2.astype(np.int64)))
    gaps = rng.integers(0, 8, n)                       # samples no block covers
    off = (np.concatenate([[0], np.cumsum(sizes + gaps)[:-1]]) + gaps).astype(np.uint32)
    total = int(off[-1] + sizes[-1] + 5)
    coeffs = rng.integers(-300, 301, total).astype(np.int16)
    coeffs[rng.random(total) < 0.5] = 0
    coeffs[rng.random(total) < 0.01] = rng.choice([-32768, 32767], 1)[0]
    return log2, cidx, qp, flags, off, coeffs


def _call(engine, fn, bd, log2, cidx, qp, flags, off, coeffs, prefill):
    out = prefill.copy()
    rc = getattr(hevcr.engine_lib(), fn)(engine.h, bd, len(log2), log2.ctypes.data, cidx.ctypes.data,
                                         qp.ctypes.data, flags.ctypes.data, off.ctypes.data, coeffs.ctypes.data,
                                         coeffs.size, out.ctypes.data)
    return rc, out


def test_noalloc_equals_allocating_entry_point(engine):
    rng = np.random.default_rng(7)
    for bd, n in ((8, 1500), (10, 700)):  # the second call is smaller
        args = _blocks(rng, n, bd)
        prefill = rng.integers(-5, 5, args[-1].size).astype(np.int16)
        print(f"bit depth {bd}: {n} blocks, {args[-1].size} coefficients", flush=True)
        rc_a, want = _call(engine, "hevcr_residual_blocks", bd, *args, prefill)
        rc_b, got = _call(engine, "hevcr_residual_blocks_noalloc", bd, *args, prefill)
        assert rc_a == 0 and rc_b == 0
        np.testing.assert_array_equal(got, want)
        assert not np.array_equal(got, prefill)


def test_noalloc_reproduces_reference_kats(engine):
    kat = np.load(oracle_lib.os.path.join(oracle_lib.GOLDEN_DIR, "residual_kat.npz"))
    blocks = kat["blocks"]
    for bd in (8, 10):
        sel = blocks[blocks[:, 0] == bd]
        args = [np.ascontiguousarray(a, dtype=np.uint8) for a in (sel[:, 1], sel[:, 2], sel[:, 4],
                                                                  sel[:, 3] | (sel[:, 5] << 1))]
        off = np.ascontiguousarray(sel[:, 6], dtype=np.uint32)
        rc, out = _call(engine, "hevcr_residual_blocks_noalloc", bd, *args, off, kat["coeffs"],
                        np.zeros_like(kat["coeffs"]))
        assert rc == 0
        for _, log2, *_rest, o in sel:
            n2 = 1 << (2 * log2)
            np.testing.assert_array_equal(out[o:o + n2], kat["residual"][o:o + n2])


def test_noalloc_rejects_sizes_beyond_its_scratch(engine):
    rng = np.random.default_rng(9)
    log2 = np.full(1100, 5, np.uint8)  # 1100 x 32 rows > 32768
    z = np.zeros(1100, np.uint8)
    off = (np.arange(1100) * 1024).astype(np.uint32)
    coeffs = np.zeros(1100 * 1024, np.int16)  # also more than 2^20 coefficients
    rc, _ = _call(engine, "hevcr_residual_blocks_noalloc", 8, log2, z, z, z, off, coeffs, coeffs)
    assert rc == -3  # HEVCR_EINVAL
    log2 = np.full(1100, 5, np.uint8)
    off = np.zeros(1100, np.uint32)
    coeffs = np.zeros(1024, np.int16)  # coefficients fit, rows do not
    rc, _ = _call(engine, "hevcr_residual_blocks_noalloc", 8, log2, z, z, z, off, coeffs, coeffs)
    assert rc == -3
    # a call within the limits still works after the rejected ones
    args = _blocks(rng, 50, 8)
    prefill = np.zeros_like(args[-1])
    assert _call(engine, "hevcr_residual_blocks", 8, *args, prefill)[1].tolist() == \
        _call(engine, "hevcr_residual_blocks_noalloc", 8, *args, prefill)[1].tolist()
