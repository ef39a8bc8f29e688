"""HIP engine parity: every golden case reconstructed on the GPU through the
C-ABI matches the reference's golden SHA-256 per plane after each stage
(pre-deblock, post-deblock, post-SAO); on a mismatch the sample differences
against the CPU oracle are reported. Residual blocks match the reference's
dequant + inverse-transform KATs bit for bit."""
import numpy as np
import pytest

import hevcr
import oracle_lib

pytestmark = pytest.mark.gpu
CASES = oracle_lib.load_goldens()
STOPS = ((hevcr.RUN_STOP_RECON, "recon", 0), (hevcr.RUN_STOP_DEBLOCK, "deblock", 1), (hevcr.RUN_FULL, "sao", 2))


@pytest.fixture(scope="module")
def engine():
    e = hevcr.Engine(0)
    yield e
    e.close()


def _diff_report(frame, got, idx):
    ref = oracle_lib.recon(frame)[idx]
    msg = []
    for c in range(3):
        d = np.argwhere(got.planes[c] != ref.planes[c])
        if len(d):
            y, x = d[0]
            msg.append(f"plane {c}: {len(d)} samples differ, first at (x={x}, y={y}) "
                       f"gpu={got.planes[c][y, x]} oracle={ref.planes[c][y, x]}")
    return "; ".join(msg)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_engine_matches_reference_goldens(engine, case):
    f = oracle_lib.make_frame(case)
    for stop, stage, idx in STOPS:
        print(f"case {case['name']} stage {stage}", flush=True)
        pic = hevcr.decode_on_engine(engine, f, stop)
        if pic.sha256() != case["stages"][stage]:
            pytest.fail(f"{stage}: " + _diff_report(f, pic, idx))


def test_engine_residual_blocks_match_reference_kats(engine):
    kat = np.load(oracle_lib.os.path.join(oracle_lib.GOLDEN_DIR, "residual_kat.npz"))
    blocks = kat["blocks"]
    for bd in (8, 10):
        sel = blocks[blocks[:, 0] == bd]
        out = engine.residual_blocks(bd, sel[:, 1], sel[:, 2], sel[:, 4], sel[:, 3] | (sel[:, 5] << 1), sel[:, 6],
                                     kat["coeffs"])
        for _, log2, *_rest, off in sel:
            n2 = 1 << (2 * log2)
            np.testing.assert_array_equal(out[off:off + n2], kat["residual"][off:off + n2])
