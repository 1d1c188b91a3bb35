"""CPU checks that pin the slant-stacking oracle (oracle/radon_ss_oracle.c, restating
radon_ss.c:43-196) before it is trusted.

* golden vectors from the reference's own radon_ss.c, compiled from the reference
  sources with its flags and with IEEE flags (tests/golden/make_golden.py), methods 0
  and 1 -- bitwise on every bin whose samples stay inside the image (the reference
  indexes outside the image rows for lines that miss the square; those bins are marked
  in the fixture and the back-projection input is zero there);
* the live reference build in oracle/_ref when present -- bitwise, same rule;
* closed forms: theta = 0 gives column sums over rows 0..npix-2, theta = pi/2 row sums
  over columns 0..npix-2 (the loops stop one short of the last row/column);
* exact adjointness of the FP64 restatement for both methods;
* the edge-bin marker finds the out-of-square lines (first bin near theta = pi).
"""
import os

import numpy as np
import pytest

from oracle import radon as O
from oracle import ref as R
from iterative_tomographic_reconstruction_algorithms_amd.phantom import angles_deg

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "radon_ss_ref.npz")
Z = np.load(GOLDEN)
KEYS = sorted({k.split("/")[0] for k in Z.files})


def method_of(key):
    return int(key.rsplit("_m", 1)[1])


def masked_equal(got, want, edge):
    g = np.where(edge, 0, got).astype(np.float32)
    w = np.where(edge, 0, want).astype(np.float32)
    np.testing.assert_array_equal(g.view(np.int32), w.view(np.int32))


@pytest.mark.parametrize("key", KEYS)
def test_ss_oracle_matches_reference_golden(key):
    m = method_of(key)
    ang = Z[key + "/angles"]
    img, sino, edge = Z[key + "/image"], Z[key + "/sino"], Z[key + "/edge"]
    for mode, flags in ((O.FAST, "fast"), (O.IEEE, "strict")):
        masked_equal(O.ss_forward(img, ang, mode, m), Z[key + "/fwd_" + flags], edge)
        np.testing.assert_array_equal(O.ss_back(sino, ang, mode, m), Z[key + "/back_" + flags])


def test_edge_bins_are_few_and_present():
    key = "ss_grid_64_305_m1"
    edge = Z[key + "/edge"]
    assert 0 < edge.sum() < 0.01 * edge.size


def test_fast_and_ieee_forms_differ():
    key = "ss_grid_64_36_m1"
    a = O.ss_forward(Z[key + "/image"], Z[key + "/angles"], O.FAST, 1)
    b = O.ss_forward(Z[key + "/image"], Z[key + "/angles"], O.IEEE, 1)
    assert (a != b).mean() > 0.3


@pytest.mark.skipif(not R.available("ss"), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("n,v,seed", [(32, 17, 5), (48, 90, 6), (33, 12, 7)])
@pytest.mark.parametrize("method", [0, 1])
def test_ss_oracle_matches_live_reference(n, v, seed, method):
    rng = np.random.default_rng(seed)
    ang = np.sort(rng.uniform(0, np.pi, v)).astype(np.float32)
    img = (rng.random((n, n)) - 0.3).astype(np.float32)
    for flags, mode in (("fast", O.FAST), ("strict", O.IEEE)):
        edge = O.ss_edge_bins(n, ang, method, mode)[0]
        sino = np.where(edge, 0, rng.random((v, n))).astype(np.float32)
        masked_equal(O.ss_forward(img, ang, mode, method),
                     R.forward("ss", img, ang, flags, method), edge)
        np.testing.assert_array_equal(O.ss_back(sino, ang, mode, method),
                                      R.back("ss", sino, ang, flags, method))


@pytest.mark.parametrize("method", [0, 1])
def test_zero_angle_closed_form(method):
    n = 16
    img = np.random.default_rng(4).random((n, n))
    s = O.ss_forward(img, np.array([0.0], np.float32), O.F64, method)[0]
    # |sin| < 1e-5: rows y = -nh .. nh-2, sample column t exactly: bin k <- column k
    np.testing.assert_allclose(s, img[:n - 1, :].sum(axis=0), rtol=1e-13)


def test_right_angle_closed_form():
    n = 16
    img = np.random.default_rng(5).random((n, n))
    s = O.ss_forward(img, np.array([np.pi / 2], np.float32), O.F64, 1)[0]
    # |cos| < 1e-5: columns x = -nh .. nh-2, sample row ~ t: bin k <- row k (cos of
    # float(pi/2) is -4.4e-8, which tilts the line by that much)
    np.testing.assert_allclose(s, img[:, :n - 1].sum(axis=1), rtol=1e-6)


@pytest.mark.parametrize("method", [0, 1])
def test_fp64_oracle_is_exactly_adjoint(method):
    n, v = 24, 20
    rng = np.random.default_rng(7)
    ang = (angles_deg(v) * np.pi / 180).astype(np.float32)
    x = rng.random((n, n))
    y = rng.random((v, n))
    lhs = np.sum(O.ss_forward(x, ang, O.F64, method) * y)
    rhs = np.sum(x * O.ss_back(y, ang, O.F64, method))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_edge_marker_finds_missing_line():
    # theta slightly below pi: the line of bin 0 (t = -nh) is x ~ nh, outside [-nh, nh-1]
    n = 16
    ang = np.array([np.pi - 0.05], np.float32)
    edge, (lo, hi) = O.ss_edge_bins(n, ang, 1, O.FAST)
    assert edge[0, 0]
    assert lo < 0 or hi >= n * n
