"""The oracle against the reference's own known answers (test/test_distribution.cu) and the
published PCG32 outputs.  These pin the oracle before it is trusted as the checker."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "distribution_known_answers.json")


def _xi(s, eta):
    return np.float32(eval(s, {"eta": np.float32(eta)}))


@pytest.fixture(scope="module")
def ka():
    with open(GOLD) as f:
        return json.load(f)


def approx(a, b):
    # Catch2's Approx default: epsilon = FLT_EPSILON * 100, relative to the larger magnitude
    return abs(a - b) <= 1.1920929e-05 * (max(abs(a), abs(b)) + 1.0)


def test_pmf_1d(ka):
    c = ka["pmf_1d"]
    pmf, _, _, _ = O.dist1d(c["values"])
    assert all(approx(float(p), e) for p, e in zip(pmf, c["pmf"]))


def test_pmf_empty(ka):
    c = ka["pmf_empty"]
    pmf, tot, _, _ = O.dist1d(c["values"])
    assert list(pmf) == c["pmf"] and tot == 0.0


def test_total_mass(ka):
    c = ka["total_mass"]
    _, tot, _, _ = O.dist1d(c["values"])
    assert approx(tot, c["total"])


def test_sample_1d(ka):
    c = ka["sample_1d"]
    xis = [_xi(s, ka["eta"]) for s in c["xis"]]
    _, _, idx, pm = O.dist1d(c["values"], xis)
    n = c["asserted"]
    assert list(idx[:n]) == c["indices"][:n]
    assert all(approx(float(p), e) for p, e in zip(pm[:n], c["pmfs"][:n]))
    # the unasserted tail follows the same CMF rule
    assert list(idx) == c["indices"]


def test_pmf_2d(ka):
    c = ka["pmf_2d"]
    pmf, _, _ = O.dist2d(c["values"], c["width"], c["height"])
    assert all(approx(float(p), e) for p, e in zip(pmf.reshape(-1), c["pmf"]))


def test_sample_2d(ka):
    c = ka["sample_2d"]
    xis = [[_xi(a, ka["eta"]), _xi(b, ka["eta"])] for a, b in c["xis"]]
    _, xy, pm = O.dist2d(c["values"], c["width"], c["height"], xis)
    assert [list(map(int, r)) for r in xy] == c["samples"]
    assert all(approx(float(p), e) for p, e in zip(pm, c["pmfs"]))


def test_phitheta_canonical(ka):
    c = ka["phitheta_canonical"]
    m = c["multiplier"]
    pdf = O.phitheta_pdf_canonical(c["values"], c["width"], c["height"], c["width"] * m, c["height"] * m)
    # test_distribution.cu:447-453: sum of pdf/cells in float
    s = np.float32(0)
    cells = pdf.size
    for v in pdf.reshape(-1):
        s = np.float32(s + np.float32(v) / np.float32(cells))
    assert approx(float(s), c["integral"])


def test_phitheta_spherical(ka):
    """test_distribution.cu:456-494: E[pdf(w) * 4 pi] over uniform sphere points = 1 (1e-2).
    The reference draws its points with TEA-seeded LCG uniforms; the oracle's PCG32 stands
    in for that generator (only the uniformity of the points matters)."""
    c = ka["phitheta_spherical"]
    w, h = c["width"], c["height"]
    rgba = np.zeros((h, w, 4), np.float32)
    # the environment weight is r+g+b; put the value in r so the weight equals the value
    rgba[..., 0] = np.asarray(c["values"], np.float32).reshape(h, w)
    n = c["points"]
    u = O.pcg32(0x853C49E6748FEA9B, 7, 2 * n).astype(np.float64) / 2.0**32
    z = u[:n] * 2 - 1
    r = np.sqrt(np.maximum(0, 1 - z * z))
    phi = 2 * np.pi * u[n:]
    dirs = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    pdf = O.env_pdf(rgba, dirs, ref=True).astype(np.float64)
    assert abs(np.mean(pdf * 4 * np.pi) - c["integral"]) <= c["rel_tol"]


def test_pcg32_published_outputs():
    """pcg-c-basic pcg32-demo: pcg32_srandom_r(&rng, 42u, 54u) first six 32-bit outputs."""
    out = O.pcg32(42, 54, 6)
    assert [int(x) for x in out] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_path_uniforms_range_and_streams():
    a = O.path_uniforms(0, 5, 0, 4096)
    b = O.path_uniforms(0, 5, 1, 4096)
    c = O.path_uniforms(0, 6, 0, 4096)
    assert a.min() >= 0.0 and a.max() < 1.0
    assert not np.array_equal(a, b) and not np.array_equal(a, c)
    assert abs(a.mean() - 0.5) < 0.02
