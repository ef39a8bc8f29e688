"""The CPU oracle reproduces the reference decoder's own reconstruction
(golden SHA-256 per plane and stage, written by tests/golden/make_goldens.py
from the reference's code) on every golden case, and the seeded generator
still produces the exact syntax the goldens were made from."""
import numpy as np
import pytest

import oracle_lib

CASES = oracle_lib.load_goldens()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_reference_goldens(case):
    f = oracle_lib.make_frame(case)
    assert oracle_lib.syntax_digest(f) == case["syntax_sha256"], "generator drift"
    pics = oracle_lib.recon(f)
    for stage, pic in zip(("recon", "deblock", "sao"), pics):
        assert pic.sha256() == case["stages"][stage], f"{stage} differs"


def test_residual_kats():
    kat = np.load(oracle_lib.os.path.join(oracle_lib.GOLDEN_DIR, "residual_kat.npz"))
    for bd, log2, cidx, intra, qp, tskip, off in kat["blocks"]:
        n2 = 1 << (2 * log2)
        got = oracle_lib.residual_block(int(log2), int(cidx), int(intra), int(qp), int(bd), int(tskip),
                                        kat["coeffs"][off:off + n2])
        np.testing.assert_array_equal(got, kat["residual"][off:off + n2])
