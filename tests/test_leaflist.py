"""The leaf-list reader (mc_leaf_list_read, generic_amr_read of
src/read_generic_amr.f90:34-237): FITS binary tables and the text form
written here and read back exactly, the reference's column-name rules,
index fallback, origin keywords and file-name resolution, the reference's
own example leaf lists (when the checkout is present) against an
independent numpy decode of the same bytes and against the tiling
invariant (leaf volumes sum to the box), and the raster permutation of a
file 'car' grid (src/grid_mod_amr.f90:164-196).  Host only."""
import gzip
import os

import numpy as np
import pytest

from mochii_amd.leaflist import raster_order, read_leaf_list
from mochii_amd.lib import MochiiError

REF_EXAMPLES = "/root/reference/examples"


def _card(key, value):
    if isinstance(value, str):
        v = "'%-8s'" % value
        return ("%-8s= %-20s" % (key, v)).ljust(80)
    if isinstance(value, bool):
        return ("%-8s= %20s" % (key, "T" if value else "F")).ljust(80)
    return ("%-8s= %20s" % (key, repr(value))).ljust(80)


def _block(cards):
    s = "".join(cards) + "END".ljust(80)
    s += " " * (-len(s) % 2880)
    return s.encode("ascii")


def write_fits(path, cols, keys, gz=False):
    """cols: list of (name, TFORM letter, array); keys: extra header cards."""
    dt = {"D": ">f8", "E": ">f4", "J": ">i4", "K": ">i8", "I": ">i2"}
    n = len(cols[0][2])
    rec = np.zeros(n, dtype=[(f"c{i}", dt[t]) for i, (_, t, _) in enumerate(cols)])
    for i, (_, t, a) in enumerate(cols):
        rec[f"c{i}"] = a
    data = rec.tobytes()
    data += b"\0" * (-len(data) % 2880)
    prim = _block([_card("SIMPLE", True), _card("BITPIX", 8), _card("NAXIS", 0),
                   _card("EXTEND", True)])
    cards = [_card("XTENSION", "BINTABLE"), _card("BITPIX", 8), _card("NAXIS", 2),
             _card("NAXIS1", rec.dtype.itemsize), _card("NAXIS2", n), _card("PCOUNT", 0),
             _card("GCOUNT", 1), _card("TFIELDS", len(cols))]
    for i, (name, t, _) in enumerate(cols):
        cards += [_card(f"TTYPE{i + 1}", name), _card(f"TFORM{i + 1}", "1" + t)]
    cards += [_card(k, v) for k, v in keys.items()]
    blob = prim + _block(cards) + data
    if gz:
        with gzip.open(path, "wb") as f:
            f.write(blob)
    else:
        with open(path, "wb") as f:
            f.write(blob)


def _tree(seed=3, nbase=2):
    """A small random octree leaf list in [0, 8]^3 (box 8, origin 0)."""
    rng = np.random.default_rng(seed)
    leaves = []

    def split(cx, cy, cz, lev):
        if lev < 4 and (lev < nbase or rng.random() < 0.3):
            h = 8.0 / 2 ** (lev + 2)
            for dz in (-h, h):
                for dy in (-h, h):
                    for dx in (-h, h):
                        split(cx + dx, cy + dy, cz + dz, lev + 1)
        else:
            leaves.append((cx, cy, cz, lev))

    split(4.0, 4.0, 4.0, 0)
    a = np.array(leaves)
    return a[:, 0], a[:, 1], a[:, 2], a[:, 3].astype(np.int32), 10 ** rng.uniform(-1, 3, len(a))


@pytest.mark.parametrize("gz", [False, True])
def test_fits_round_trip_by_name_with_origin(tmp_path, gz):
    x, y, z, lev, nH = _tree()
    xHI = np.linspace(0, 1, x.size)
    # columns out of order, density under its alias, float32 xHI, an ignored T
    cols = [("level", "J", lev), ("dens", "D", nH), ("T", "D", np.full(x.size, 1e4)),
            ("z", "D", z), ("y", "D", y), ("x", "D", x), ("xHI", "E", xHI)]
    name = tmp_path / ("leaves.fits.gz" if gz else "leaves.fits")
    write_fits(name, cols, {"BOXLEN": 8.0, "ORIGINX": 0.0, "ORIGINY": 0.0, "ORIGINZ": 0.0},
               gz=gz)
    d = read_leaf_list(name)
    assert d["half"] == 4.0
    np.testing.assert_array_equal(d["x"], x - 4.0)
    np.testing.assert_array_equal(d["y"], y - 4.0)
    np.testing.assert_array_equal(d["z"], z - 4.0)
    np.testing.assert_array_equal(d["level"], lev)
    np.testing.assert_array_equal(d["nH"], nH)
    np.testing.assert_array_equal(d["xHI"], xHI.astype(np.float32).astype(np.float64))
    assert "metallicity" not in d and "ndust" not in d
    # the leaves tile the box
    assert np.sum((8.0 / 2.0 ** d["level"]) ** 3) == pytest.approx(512.0, rel=1e-14)


def test_fits_index_fallback_and_default_origin(tmp_path):
    """Unnamed columns fall back to positions 1..5; without ORIGIN keywords
    the box starts at -BOXLEN/2 (coordinates already centred)."""
    x, y, z, lev, nH = _tree(seed=5)
    cols = [("c1", "D", x - 4), ("c2", "D", y - 4), ("c3", "D", z - 4), ("c4", "I", lev),
            ("c5", "E", nH)]
    write_fits(tmp_path / "f.fits", cols, {"BOXLEN": 8.0})
    d = read_leaf_list(tmp_path / "f.fits")
    np.testing.assert_array_equal(d["x"], x - 4)
    np.testing.assert_array_equal(d["level"], lev)
    np.testing.assert_array_equal(d["nH"], nH.astype(np.float32).astype(np.float64))


def test_name_resolution_and_refusals(tmp_path):
    x, y, z, lev, nH = _tree(seed=7)
    write_fits(tmp_path / "grid.fits.gz", [("x", "D", x), ("y", "D", y), ("z", "D", z),
                                           ("level", "K", lev), ("nH", "D", nH)],
               {"BOXLEN": 8.0, "ORIGINX": 0.0, "ORIGINY": 0.0, "ORIGINZ": 0.0}, gz=True)
    d = read_leaf_list(tmp_path / "grid")            # extension resolved as the reference does
    np.testing.assert_array_equal(d["nH"], nH)
    (tmp_path / "g.h5").write_bytes(b"\x89HDF")
    with pytest.raises(MochiiError, match="HDF5"):
        read_leaf_list(tmp_path / "g.h5")
    with pytest.raises(MochiiError, match="not found"):
        read_leaf_list(tmp_path / "missing")
    write_fits(tmp_path / "nobox.fits", [("x", "D", x), ("y", "D", y), ("z", "D", z),
                                         ("level", "J", lev), ("nH", "D", nH)], {})
    with pytest.raises(MochiiError, match="BOXLEN"):
        read_leaf_list(tmp_path / "nobox.fits")


def test_text_form_five_and_nine_columns(tmp_path):
    x, y, z, lev, nH = _tree(seed=9)

    def row(a, b, c, l, n):
        return f"{float(a)!r} {float(b)!r} {float(c)!r} {int(l)} {float(n)!r}"

    rows9 = "\n".join(row(*r) + " 1e4 0 0 0" for r in zip(x - 4, y - 4, z - 4, lev, nH))
    (tmp_path / "l9.dat").write_text(f"{x.size} 8.0\n{rows9}\n")
    rows5 = "\n".join(row(*r) for r in zip(x - 4, y - 4, z - 4, lev, nH))
    (tmp_path / "l5.txt").write_text(f"{x.size} 8.0\n{rows5}\n")
    for f in ("l9.dat", "l5.txt"):
        d = read_leaf_list(tmp_path / f)
        assert d["half"] == 4.0
        np.testing.assert_array_equal(d["x"], x - 4)
        np.testing.assert_array_equal(d["z"], z - 4)
        np.testing.assert_array_equal(d["level"], lev)
        np.testing.assert_array_equal(d["nH"], nH)
    (tmp_path / "short.dat").write_text(f"{x.size + 1} 8.0\n{rows5}\n")
    with pytest.raises(MochiiError, match="ends after"):
        read_leaf_list(tmp_path / "short.dat")


def _numpy_decode(path):
    """An independent decode of a leaf-list FITS table: header cards parsed
    here, the rows as a big-endian numpy record array."""
    b = gzip.open(path).read() if str(path).endswith(".gz") else open(path, "rb").read()

    def header(at):
        cards, p = {}, at
        while True:
            c = b[p:p + 80].decode("ascii")
            p += 80
            if c.startswith("END"):
                return cards, -(-p // 2880) * 2880
            if c[8:10] == "= ":
                v = c[10:].split("/")[0].strip().strip("'").strip()
                cards[c[:8].strip()] = v

    _, p = header(0)
    h, p = header(p)
    nf = int(h["TFIELDS"])
    code = {"D": ">f8", "E": ">f4", "J": ">i4", "K": ">i8", "I": ">i2"}
    dt = np.dtype([(h[f"TTYPE{i}"].lower(), code[h[f"TFORM{i}"][-1]]) for i in range(1, nf + 1)])
    rec = np.frombuffer(b, dtype=dt, count=int(h["NAXIS2"]), offset=p)
    return h, rec


@pytest.mark.parametrize("name", ["adaptive_refinement/amr_L5.fits.gz",
                                  "pdr/amr_sphere_L5_nH100.fits.gz",
                                  "cosmological_cutout/tng_uniform.fits.gz"])
def test_reference_example_leaf_lists(name):
    path = os.path.join(REF_EXAMPLES, name)
    if not os.path.exists(path):
        pytest.skip("reference checkout not present")
    d = read_leaf_list(path)
    h, rec = _numpy_decode(path)
    box = float(h["BOXLEN"])
    org = [float(h.get(k, -0.5 * box)) for k in ("ORIGINX", "ORIGINY", "ORIGINZ")]
    assert d["half"] == 0.5 * box and d["x"].size == rec.size
    for k, o in zip("xyz", org):
        np.testing.assert_array_equal(d[k], rec[k].astype(np.float64) - (o + 0.5 * box))
    np.testing.assert_array_equal(d["level"], rec["level"])
    np.testing.assert_array_equal(d["nH"], rec["nh"].astype(np.float64))
    for opt in ("metallicity", "xHI", "ndust"):
        assert (opt in d) == (opt.lower() in rec.dtype.names)
    assert np.sum((box / 2.0 ** d["level"]) ** 3) == pytest.approx(box ** 3, rel=1e-12)
    assert np.all(np.abs(d["x"]) < 0.5 * box)


def test_raster_order_of_a_single_level_list():
    n = 4
    half = 2.0
    c = -half + (np.arange(n) + 0.5) * (2 * half / n)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    rng = np.random.default_rng(1)
    p = rng.permutation(n ** 3)
    x, y, z = X.ravel()[p], Y.ravel()[p], Z.ravel()[p]
    perm = raster_order(x, y, z, np.full(n ** 3, 2), half)
    k = np.arange(n ** 3)
    np.testing.assert_array_equal(x[perm], c[k % n])
    np.testing.assert_array_equal(y[perm], c[(k // n) % n])
    np.testing.assert_array_equal(z[perm], c[k // (n * n)])
    with pytest.raises(MochiiError, match="single-level"):
        raster_order(x, y, z, np.r_[np.full(n ** 3 - 1, 2), 3], half)
    with pytest.raises(MochiiError, match="expects nleaf"):
        raster_order(x[:-1], y[:-1], z[:-1], np.full(n ** 3 - 1, 2), half)
