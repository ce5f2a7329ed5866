"""Pins of the Lorenzo front end's CPU restatement (oracle/lorenzo.py).  No
reference implementation exists (the reference reads ready-made codes), so
the definition is pinned by what the mathematics fixes: the 8-term residual
against a brute-force loop, closed-form residuals of a linear field, the
inverse as an exact round trip of q, the error bound of the reconstruction,
and the outlier order."""
import numpy as np

from oracle import lorenzo as L


def brute_residual(q):
    nz, ny, nx = q.shape

    def g(z, y, x):
        return int(q[z, y, x]) if z >= 0 and y >= 0 and x >= 0 else 0

    r = np.zeros(q.shape, dtype=np.int64)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                r[z, y, x] = (g(z, y, x) - g(z - 1, y, x) - g(z, y - 1, x) - g(z, y, x - 1)
                              + g(z - 1, y - 1, x) - g(z - 1, y - 1, x - 1) + g(z - 1, y, x - 1) + g(z, y - 1, x - 1))
    return r


def test_residual_equals_brute_force():
    rng = np.random.default_rng(1)
    q = rng.integers(-1000, 1000, (3, 4, 5))
    assert (L.residual(q) == brute_residual(q)).all()


def test_linear_field_closed_form():
    # q = 3x + 5y + 7z: every second difference vanishes, so r is q's value
    # at the origin corner's neighbours only: r(0,0,x>0) = 3, r(0,y>0,0) = 5,
    # r(z>0,0,0) = 7, r(0,0,0) = 0 and 0 elsewhere
    z, y, x = np.meshgrid(np.arange(4), np.arange(5), np.arange(6), indexing="ij")
    q = 3 * x + 5 * y + 7 * z
    r = L.residual(q)
    exp = np.zeros_like(r)
    exp[0, 0, 1:] = 3
    exp[0, 1:, 0] = 5
    exp[1:, 0, 0] = 7
    assert (r == exp).all()


def test_round_trip_and_error_bound():
    rng = np.random.default_rng(2)
    eb = 1e-3
    f = (np.sin(np.arange(7 * 40 * 300).reshape(7, 40, 300) / 37.0) * 3 + rng.normal(0, 0.01, (7, 40, 300)))
    f = f.astype(np.float32)
    f[3, 10, 100] = 50.0  # a spike: outliers around it
    codes, idx, val = L.lorenzo_quantize(f, eb, radius=512)
    assert idx.size >= 8 and (codes.reshape(-1)[idx] == 0).all()
    rec = L.lorenzo_reconstruct(codes, idx, val, f.shape, eb)
    q = L.quantize(f, eb)
    assert (rec == q.astype(np.float32) * np.float32(2 * eb)).all()
    # the error bound, up to the float32 roundings of f / (2 eb) and q * 2 eb
    assert np.abs(rec.astype(np.float64) - f).max() <= eb + 4 * np.finfo(np.float32).eps * np.abs(f).max()


def test_outliers_in_tile_order():
    # one outlier per tile in reverse index order lands in tile order
    f = np.zeros((32, 32, 512), dtype=np.float32)
    spots = [(17, 20, 300), (17, 3, 300), (2, 20, 10), (2, 3, 400), (2, 3, 10)]
    for z, y, x in spots:
        f[z, y, x] = 10.0
    codes, idx, val = L.lorenzo_quantize(f, 0.5, radius=2)
    # a spike of q = 10 makes residual +-10 at its 8-neighbourhood corners
    assert idx.size == 8 * len(spots)
    tiles = []
    for i in idx:
        z, rem = divmod(int(i), 32 * 512)
        y, x = divmod(rem, 512)
        tiles.append((z // 16) * 4 + (y // 16) * 2 + x // 256)
    assert tiles == sorted(tiles)
