"""Input/output formats (acqTracking.cpp, processIsmrmrd.hpp:362-471,
processNIFTI.hpp:41-119): the acquisition table and coil-major gather on a
hand-built set of acquisitions with the expected vectors written out, and
the NIfTI-1 writer decoded at the standard's offsets.  CPU only (the
library's file writer runs on the host)."""
import numpy as np
import pytest

from powergrid_amd import io as PIO


def acq(k, nc=2, nro=3):
    c, s = np.meshgrid(np.arange(nc), np.arange(nro), indexing="ij")
    v = 100.0 * k + 10.0 * c + s
    data = (v - 1j * v).astype(np.complex64)          # (nc, nro)
    traj = np.stack([1000 + 10 * k + np.arange(nro), 2000 + 10 * k + np.arange(nro),
                     3000 + 10 * k + np.arange(nro), 0.5 * k + np.arange(nro)], axis=1)
    return data, traj.astype(np.float32)              # (nro, 4)


def test_acq_table_and_gather_order():
    # file order: (shot 1, par 0), (shot 0, par 1), (shot 0, par 0), (shot 1, par 1),
    # then (shot 0, par 0) again, which replaces acquisition 2
    keys = [(1, 0), (0, 1), (0, 0), (1, 1), (0, 0)]
    counters = [(sh, par, 0, 0, 0, 0, 0) for sh, par in keys]
    t = PIO.AcqTracking(counters, (2, 2, 1, 1, 1, 1, 1))
    assert t.acq_array[0, 0, 0, 0, 0, 0, 0] == 4
    assert t.acq_array[1, 0, 0, 0, 0, 0, 0] == 0
    acqs = [acq(k) for k in range(len(keys))]
    data, kx, ky, kz, tv = PIO.get_complete_acq_data([a[0] for a in acqs], [a[1] for a in acqs], t)
    # partition-major, shot-minor: acquisitions 4, 0, 1, 3
    order = [4, 0, 1, 3]
    exp = [100 * k + 10 * c + s for c in range(2) for k in order for s in range(3)]
    np.testing.assert_array_equal(data.real, np.array(exp, np.float32))
    np.testing.assert_array_equal(data.imag, -np.array(exp, np.float32))
    np.testing.assert_array_equal(kx, [1000 + 10 * k + s for k in order for s in range(3)])
    np.testing.assert_array_equal(kz, [3000 + 10 * k + s for k in order for s in range(3)])
    np.testing.assert_allclose(tv, [0.5 * k + s for k in order for s in range(3)])


def test_acq_gather_selects_slice_and_skips_missing():
    counters = [(0, 0, 1, 0, 0, 0, 0), (1, 0, 0, 0, 0, 0, 0), (0, 0, 0, 0, 0, 0, 0)]
    t = PIO.AcqTracking(counters, (3, 1, 2, 1, 1, 1, 1))
    acqs = [acq(k) for k in range(3)]
    d0, kx0, *_ = PIO.get_complete_acq_data([a[0] for a in acqs], [a[1] for a in acqs], t, NSlice=0)
    assert d0.size == 2 * 2 * 3  # shots 0 and 1 of slice 0; shot 2 absent
    np.testing.assert_array_equal(kx0[:3], [1020, 1021, 1022])
    d1, *_ = PIO.get_complete_acq_data([a[0] for a in acqs], [a[1] for a in acqs], t, NSlice=1)
    assert d1.size == 2 * 3
    with pytest.raises(ValueError):
        PIO.AcqTracking([(3, 0, 0, 0, 0, 0, 0)], (3, 1, 1, 1, 1, 1, 1))


def test_nifti_real_header_and_data(tmp_path):
    Nx, Ny, Nz, Nslab = 4, 3, 2, 2
    img = np.arange(Nx * Ny * Nz * Nslab, dtype=np.float32) * 0.5 - 3
    PIO.write_nifti_real(tmp_path / "im", img, Nx, Ny, Nz, Nslab)
    raw = open(tmp_path / "im.nii", "rb").read()
    assert len(raw) == 352 + 4 * img.size
    hdr, data = PIO.read_nifti(tmp_path / "im.nii")
    assert hdr["sizeof_hdr"] == 348
    assert hdr["dim"] == [3, Nx, Ny, Nz, Nslab, 1, 1, 1]
    assert hdr["datatype"] == 16 and hdr["bitpix"] == 32
    assert hdr["pixdim"][:4] == [1.0, 1.0, 1.0, 1.0]
    assert hdr["vox_offset"] == 352.0 and hdr["scl_slope"] == 100.0
    assert hdr["xyzt_units"] == 2 and hdr["magic"] == b"n+1\0"
    assert raw[348:352] == b"\0\0\0\0"
    np.testing.assert_array_equal(data, img)
