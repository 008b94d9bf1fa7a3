"""DM-trial input files (host logic): readData of PRESTO .dat files
(source/bpd_data.c:9-75) and the .inf header reader."""
import ctypes as C

import numpy as np
import pytest

INF_TEXT = """ Data file name without suffix          =  Lband_topo_DM0.00
 Telescope used                         =  Parkes
 Instrument used                        =  Multibeam
 Object being observed                  =  J1910-5959
 J2000 Right Ascension (hh:mm:ss.ssss)  =  19:10:00.0000
 J2000 Declination     (dd:mm:ss.ssss)  =  -59:59:00.0000
 Data observed by                       =  Someone
 Epoch of observation (MJD)             =  51000.123456789012
 Barycentered?           (1=yes, 0=no)  =  0
 Number of bins in the time series      =  1000
 Width of each time series bin (sec)    =  7.2e-05
 Any breaks in the data? (1=yes, 0=no)  =  0
 Type of observation (EM band)          =  Radio
 Beam diameter (arcsec)                 =  840
 Dispersion measure (cm-3 pc)           =  12.5
 Central freq of low channel (Mhz)      =  1182.1953125
 Total bandwidth (Mhz)                  =  288
 Number of channels                     =  96
 Channel bandwidth (Mhz)                =  3
 Data analyzed by                       =  Someone
 Any additional notes:
    Input filterbank samples have 2 bits.
"""


@pytest.fixture(scope="module")
def fd():
    import gpu_accelsearch_amd as fd
    fd.lib()
    return fd


def test_read_dat_and_padding(fd, tmp_path):
    x = np.random.default_rng(1).standard_normal(1000).astype(np.float32)
    path = tmp_path / "t.dat"
    x.tofile(path)
    np.testing.assert_array_equal(fd.readData(str(path)), x)
    p = C.POINTER(C.c_float)()
    n = fd.lib().readData(str(path).encode(), C.byref(p))
    try:
        assert n == 1000
        # 2000 zero floats before and after the samples (headers/bpd_common.h:35)
        before = np.array([p[k] for k in range(-2000, 0)])
        after = np.array([p[k] for k in range(1000, 3000)])
        assert not before.any() and not after.any()
    finally:
        fd.lib().fdas_free_data(p)


def test_read_dat_errors(fd, tmp_path):
    with pytest.raises(fd.FdasError, match="cannot open"):
        fd.readData(str(tmp_path / "missing.dat"))
    (tmp_path / "empty.dat").write_bytes(b"")
    with pytest.raises(fd.FdasError, match="no samples"):
        fd.readData(str(tmp_path / "empty.dat"))
    big = tmp_path / "big.dat"
    with open(big, "wb") as f:          # sparse: nothing is written or read
        f.truncate(4 * (67108864 + 1))
    with pytest.raises(fd.FdasError, match="at most"):
        fd.readData(str(big))


def test_read_inf(fd, tmp_path):
    (tmp_path / "a.inf").write_text(INF_TEXT)
    inf = fd.read_inf(str(tmp_path / "a.inf"))
    assert inf["basename"] == "Lband_topo_DM0.00" and inf["telescope"] == "Parkes"
    assert inf["n"] == 1000 and inf["dt"] == 7.2e-05 and inf["dm"] == 12.5
    assert inf["epoch_mjd"] == pytest.approx(51000.123456789012) and not inf["barycentred"]
    (tmp_path / "bad.inf").write_text(" Telescope used = X\n")
    with pytest.raises(fd.FdasError, match="sample count"):
        fd.read_inf(str(tmp_path / "bad.inf"))


def test_load_dm_trials_checks_lengths(fd, tmp_path):
    x = np.arange(1000, dtype=np.float32)
    x.tofile(tmp_path / "a.dat")
    (tmp_path / "a.inf").write_text(INF_TEXT)
    series, infos = fd.load_dm_trials([str(tmp_path / "a")])
    np.testing.assert_array_equal(series[0], x)
    x[:999].tofile(tmp_path / "b.dat")
    (tmp_path / "b.inf").write_text(INF_TEXT)
    with pytest.raises(ValueError, match="999"):
        fd.load_dm_trials([str(tmp_path / "b")])
