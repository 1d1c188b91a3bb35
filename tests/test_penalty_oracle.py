"""CPU checks that pin the SPS penalty oracle (oracle/penalty_oracle.c, restating the
reference's huber.c / tikhonov.c / haar.c / penalty_weights.c and penalty.pyx) and the
penalised, clipped SPS restatement (oracle/solvers.py, sir.py:348-560).

* golden vectors from the reference's own penalty module compiled with its flags and
  with IEEE flags (tests/golden/penalty_ref.npz) -- bitwise, and the live build when
  present;
* closed forms: Tikhonov of a linear ramp, Haar of a constant image, the Huber terms of a
  constant image (numerator 0, denominator 2 * sum of the neighbour weights);
* SPS: with beta = 0 a penalty changes nothing; the clip keeps x >= 0 and equals
  np.clip of the unclipped step.
"""
import os

import numpy as np
import pytest

from oracle import penalty as P
from oracle import ref as R
from oracle import solvers as S

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "penalty_ref.npz")
Z = np.load(GOLDEN)
CASES = sorted({k.rsplit("_", 1)[0] for k in Z.files if not k.endswith("/image")})


@pytest.mark.parametrize("case", CASES)
def test_penalty_oracle_matches_reference_golden(case):
    img_key, rest = case.split("/")
    kind, delta, flags = rest.split("_")
    mode = P.FAST if flags == "fast" else P.IEEE
    rn, rd = P.penalty(kind, Z[img_key + "/image"], float(delta), mode)
    np.testing.assert_array_equal(rn.view(np.int32), Z[case + "_num"].view(np.int32))
    np.testing.assert_array_equal(rd.view(np.int32), Z[case + "_den"].view(np.int32))


@pytest.mark.skipif(not os.path.exists(os.path.join(R.REF_DIR, "libpenalty_fast.so")),
                    reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("kind", ["huber", "tikhonov", "haar"])
def test_penalty_oracle_matches_live_reference(kind):
    rng = np.random.default_rng(3)
    for shape in ((2, 3), (40, 31)):
        img = (rng.standard_normal(shape) * 5).astype(np.float32)
        for mode, flags in ((P.FAST, "fast"), (P.IEEE, "strict")):
            a = P.penalty(kind, img, 0.37, mode)
            b = R.penalty(kind, img, 0.37, flags)
            for x, y in zip(a, b):
                np.testing.assert_array_equal(x.view(np.int32), y.view(np.int32))


def weight_sums(n):
    """Sum of the neighbour weights inside an n x n image, per pixel (centre excluded)."""
    w = np.array([[2 ** -0.5, 1, 2 ** -0.5], [1, 0, 1], [2 ** -0.5, 1, 2 ** -0.5]])
    out = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            for a in range(3):
                for b in range(3):
                    if 0 <= i + a - 1 < n and 0 <= j + b - 1 < n:
                        out[i, j] += w[a, b]
    return out


def test_constant_image_closed_forms():
    n = 9
    img = np.full((n, n), 3.25, np.float32)
    rn, rd = P.penalty("huber", img, 0.5)
    assert (rn == 0).all()
    np.testing.assert_allclose(rd, 2 * (weight_sums(n) + 4), rtol=1e-6)
    assert (P.penalty("haar", img)[0] == 0).all()
    assert (P.penalty("tikhonov", img)[1] == 2).all()


def test_tikhonov_of_a_ramp():
    # I = a*i + b*j: interior sum of w * 2 (I_c - I_nb) vanishes by symmetry
    n = 12
    i, j = np.mgrid[0:n, 0:n]
    img = (0.5 * i - 0.25 * j).astype(np.float32)
    rn = P.penalty("tikhonov", img, mode=P.IEEE)[0]
    assert np.abs(rn[1:-1, 1:-1]).max() < 1e-5
    # top-left corner: neighbours (0,1), (1,0), (1,1) with d = 0.25, -0.5, -0.25
    np.testing.assert_allclose(rn[0, 0], 2 * (0.25 - 0.5 - 0.25 * 2 ** -0.5), rtol=1e-6)


def _diag_system():
    rng = np.random.default_rng(4)
    a = rng.uniform(0.5, 2.0, (8, 8))
    b = rng.uniform(-0.2, 1.0, (8, 8)).astype(np.float32)
    return (lambda x: (a * x).astype(np.float32)), (lambda y: (a * y).astype(np.float32)), b


@pytest.mark.parametrize("kind", ["huber", "tikhonov", "haar"])
def test_sps_zero_beta_penalty_changes_nothing(kind):
    A, At, b = _diag_system()
    x0 = np.full((8, 8), 0.3, np.float32)
    x1, e1 = S.sps(x0, b, A, At, 3)
    x2, e2 = S.sps(x0, b, A, At, 3, penalty=kind, beta=0.0)
    np.testing.assert_array_equal(x1, x2)


def test_sps_clip_is_np_clip_of_the_step():
    A, At, b = _diag_system()
    x0 = np.full((8, 8), 0.3, np.float32)
    xu, _ = S.sps(x0, b, A, At, 1, clip=False)
    xc, _ = S.sps(x0, b, A, At, 1, clip=True)
    assert (xu < 0).any() and (xc >= 0).all()
    np.testing.assert_array_equal(xc, np.clip(xu, 0, np.max(xu)))
