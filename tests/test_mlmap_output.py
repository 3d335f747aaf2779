"""MLmap's outputs on the host (no GPU): the HEALPix FITS writer against the
files the reference's write_map (iofiles.c:19-53, compiled from source with
cfitsio) writes -- committed bytes in tests/golden/ref_write_map.npz and, where
oracle/_ref is built, a live comparison -- and x2map_pol
(mappraiser.c:836-853) against its definition."""
import os
import re

import numpy as np
import pytest

from midapack_amd import lib as G
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "ref_write_map.npz")
CASES = [(2, "f8", True, "C"), (2, "i4", True, "C"), (4, "f8", False, "G"), (1, "f8", True, "E")]
# the DATE value (19 characters) of the 7th card of the primary header
DATE = slice(6 * 80 + 11, 6 * 80 + 30)


def _same_but_date(a, b):
    a, b = np.frombuffer(a, np.uint8).copy(), np.frombuffer(b, np.uint8).copy()
    assert a.size == b.size
    assert re.fullmatch(rb"\d{4}-\d\d-\d\dT\d\d:\d\d:\d\d", a[DATE].tobytes())
    a[DATE] = 0
    b[DATE] = 0
    return np.array_equal(a, b)


@pytest.mark.parametrize("i", range(len(CASES)))
def test_write_map_matches_reference_bytes(i, tmp_path):
    z = np.load(GOLDEN)
    nside, _, nest, cs = CASES[i]
    fn = str(tmp_path / "m.fits")
    G.write_map(z[f"data{i}"], nside, fn, nest=nest, coordsys=cs)
    assert _same_but_date(open(fn, "rb").read(), z[f"fits{i}"].tobytes())
    vals, cards = G.read_map(fn)
    assert np.array_equal(vals, z[f"data{i}"].astype(np.float64))
    assert cards["PIXTYPE"] == "HEALPIX" and int(cards["NSIDE"]) == nside
    assert cards["ORDERING"] == ("NESTED" if nest else "RING")


@pytest.mark.skipif(not O.ref_iofiles_available(), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("nside", [8, 32])
def test_write_map_matches_live_reference(nside, tmp_path):
    rng = np.random.default_rng(nside)
    for data in (rng.standard_normal(12 * nside * nside) * 1e3,
                 rng.integers(0, 10 ** 6, 12 * nside * nside).astype(np.int32)):
        O.ref_write_map(data, nside, str(tmp_path / "r.fits"))
        G.write_map(data, nside, str(tmp_path / "g.fits"))
        assert _same_but_date(open(tmp_path / "g.fits", "rb").read(), open(tmp_path / "r.fits", "rb").read())


def test_x2map_pol_definition():
    rng = np.random.default_rng(4)
    nside = 4
    npix = 12 * nside * nside
    pixels = np.sort(rng.choice(npix, 50, replace=False)).astype(np.int32)
    lstid = (3 * pixels[:, None] + np.arange(3)[None, :]).ravel().astype(np.int32)
    x = rng.standard_normal(lstid.size)
    cond = rng.random(pixels.size)
    hits = rng.integers(1, 100, pixels.size).astype(np.int32)
    I, Q, U, C, H = G.x2map_pol(x, lstid, cond, hits, nside)
    seen = np.zeros(npix, bool)
    seen[pixels] = True
    assert np.all(I[~seen] == 0) and np.all(Q[~seen] == 0) and np.all(H[~seen] == 0) and np.all(C[~seen] == 0)
    assert np.array_equal(I[pixels], x[0::3]) and np.array_equal(Q[pixels], x[1::3])
    assert np.array_equal(U[pixels], x[2::3])
    assert np.array_equal(C[pixels], cond) and np.array_equal(H[pixels], hits)
    with pytest.raises(G.GPUError):
        G.x2map_pol(x, lstid + 3 * npix, cond, hits, nside)
