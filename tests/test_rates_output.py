"""The rates-file writer (mochii_amd/output.py, gas_rates_write of
src/gas_rates_mod.f90:216-366) without a device: synthetic HDUs written and
read back by an independent FITS reader (tests/cases.py), header values of
every type, Fortran array layout (first index fastest), gzip, and the
reference's file naming (get_base_name + '_rates' + io_file_extension)."""
import numpy as np
import pytest

from mochii_amd.output import rates_filename, write_rates
from tests.cases import read_fits_images


def test_write_and_read_back(tmp_path):
    rng = np.random.default_rng(0)
    g = rng.uniform(size=37)
    jnu = rng.uniform(size=(37, 5))           # C (nleaf, nnu) = Fortran (nnu, nleaf)
    xyz = rng.uniform(size=(3, 37))
    head = [("EXTNAME", "Gamma_HI", "H I photoionization rate [s^-1]"),
            ("NNU_ION", 16, "ionizing energy-bin count"), ("EION_MAX", 100.0, "band upper edge"),
            ("TOT_LUM", 3.177837e38, "band luminosity"), ("CONVERGD", True, "converged"),
            ("FLDCONS", False, "x"), ("ENRGSAMP", "sobol", "sequence"),
            ("L_ESC", 1.2345678901234567e37, "escaped")]
    path = tmp_path / rates_filename("run/out/model.fits.gz")
    write_rates(path, [(g, head), (jnu, [("EXTNAME", "J_nu", "J")]),
                       (xyz, [("EXTNAME", "LeafXYZ", "centres")])])
    assert path.name == "model_rates.fits.gz"
    hdus = read_fits_images(path)
    assert [h["EXTNAME"] for h, _ in hdus] == ["Gamma_HI", "J_nu", "LeafXYZ"]
    h0, a0 = hdus[0]
    assert h0["SIMPLE"] is True and h0["NAXIS"] == 1 and h0["NAXIS1"] == 37
    np.testing.assert_array_equal(a0, g)
    assert h0["NNU_ION"] == 16 and h0["CONVERGD"] is True and h0["FLDCONS"] is False
    assert h0["ENRGSAMP"] == "sobol"
    assert h0["TOT_LUM"] == 3.177837e38 and h0["L_ESC"] == 1.2345678901234567e37
    assert h0["_order"][:3] == ["SIMPLE", "BITPIX", "NAXIS"]
    h1, a1 = hdus[1]
    assert h1["XTENSION"] == "IMAGE" and (h1["NAXIS1"], h1["NAXIS2"]) == (5, 37)
    np.testing.assert_array_equal(a1, jnu)
    h2, a2 = hdus[2]
    assert (h2["NAXIS1"], h2["NAXIS2"]) == (37, 3)
    np.testing.assert_array_equal(a2, xyz)


def test_rates_file_names():
    assert rates_filename("g1_stromgren.h5") == "g1_stromgren_rates.fits.gz"
    assert rates_filename("/a/b/c.fits") == "c_rates.fits.gz"
    assert rates_filename("plain") == "plain_rates.fits.gz"
    with pytest.raises(ValueError, match="FITS only"):
        rates_filename("x.h5", "hdf5")
