"""The formats on either side of the reconstruction: ISMRMRD acquisitions in
(acqTracking.cpp, processIsmrmrd.hpp) and NIfTI-1 images out
(processNIFTI.hpp).

ISMRMRD datasets are HDF5 files; no HDF5 reader is available here, so the
acquisitions come in already read: per acquisition its encoding counters,
its data (channels x samples, the ISMRMRD in-memory order) and its
trajectory (samples x trajectory dimensions, kx, ky, kz, t).  What the
reference does with them -- the acquisition table indexed by the encoding
counters and the gather into PowerGrid's coil-major data vector -- is done
here on the host, since it is a one-off reordering of the input.  NIfTI
output goes through the library (pg_write_nifti_*), with the magnitude and
phase formed on the device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import DeviceArray, _as_device, _sh
from ._lib import check, lib

# acqTracking.cpp:57-65: the encoding counters that index the table, in the
# table's dimension order
COUNTERS = ("kspace_encode_step_1", "kspace_encode_step_2", "slice", "repetition", "average",
            "contrast", "phase")


class AcqTracking:
    """acqTracking (acqTracking.cpp:14-80): acq_array[shot, par, slice, rep,
    avg, echo, phase] = acquisition index, -1 where no acquisition has those
    counters; a later acquisition with the same counters replaces an
    earlier one.

    counters: (num_acq, 7) integers in COUNTERS order; maxima: the seven
    table sizes (the header's encoding limits + 1, acqTracking.cpp:18-26)."""

    def __init__(self, counters, maxima):
        c = np.asarray(counters, dtype=np.int64).reshape(-1, 7)
        self.maxima = tuple(int(m) for m in maxima)
        if len(self.maxima) != 7 or min(self.maxima) < 1:
            raise ValueError("seven positive table sizes expected")
        if np.any(c < 0) or np.any(c >= np.array(self.maxima)):
            raise ValueError("an encoding counter exceeds its limit")
        self.acq_array = np.full(self.maxima, -1, dtype=np.int64)
        for idx, key in enumerate(c):
            self.acq_array[tuple(key)] = idx

    @property
    def NShotMax(self):
        return self.maxima[0]

    @property
    def NParMax(self):
        return self.maxima[1]


def get_complete_acq_data(acq_data, acq_traj, tracking, NSlice=0, NRep=0, NAve=0, NEcho=0,
                          NPhase=0):
    """getCompleteISMRMRDAcqData (processIsmrmrd.hpp:362-471).

    The acquisitions of one (slice, repetition, average, echo, phase), taken
    partition-major and shot-minor (:391-395), are stacked as
    data[sample, coil, acquisition], permuted to [sample, acquisition, coil]
    and vectorised column-major (:455-462): sample i of acquisition a of coil
    c lands at i + nro*a + nro*nacq*c.  kx, ky, kz, t come from trajectory
    columns 0..3 (:437-442), in the same acquisition order.

    acq_data[k]: (nc, nro) complex per acquisition; acq_traj[k]: (nro, >=4).
    Returns (data complex64, kx, ky, kz, tvec float32)."""
    order = []
    for par in range(tracking.NParMax):
        for shot in range(tracking.NShotMax):
            idx = tracking.acq_array[shot, par, NSlice, NRep, NAve, NEcho, NPhase]
            if idx != -1:
                order.append(int(idx))
    if not order:
        raise ValueError("no acquisition has these counters")
    first = np.asarray(acq_data[order[0]])
    nc, nro = first.shape
    nacq = len(order)
    data = np.empty((nc, nacq, nro), np.complex64)
    traj = np.empty((4, nacq, nro), np.float32)
    for a, idx in enumerate(order):
        d = np.asarray(acq_data[idx])
        t = np.asarray(acq_traj[idx])
        if d.shape != (nc, nro) or t.shape[0] != nro or t.shape[1] < 4:
            raise ValueError(f"acquisition {idx} does not match the first one's shape")
        data[:, a, :] = d
        traj[:, a, :] = t[:, :4].T
    return (data.ravel(), traj[0].ravel(), traj[1].ravel(), traj[2].ravel(), traj[3].ravel())


def write_nifti_real(filename, image, Nx, Ny, Nz=1, Nslab=1, N_TR=1):
    """writeNiftiRealImage (processNIFTI.hpp:41-101): <filename>.nii, float32."""
    a = np.ascontiguousarray(np.asarray(image, dtype=np.float32).ravel())
    if a.size != Nx * Ny * Nz * Nslab * N_TR:
        raise ValueError("image size does not match the dimensions")
    check(lib().pg_write_nifti_real(str(filename).encode(), a.ctypes.data_as(C.c_void_p), Nx, Ny,
                                    Nz, Nslab, N_TR))


def write_nifti_mag_phase(filename, image, Nx, Ny, Nz=1, Nslab=1, N_TR=1, stream=None):
    """writeNiftiMagPhsImage (processNIFTI.hpp:104-119): <filename>_mag.nii and
    <filename>_phs.nii from a complex image (numpy or DeviceArray)."""
    d = _as_device(image, np.complex64)
    if d.size != Nx * Ny * Nz * Nslab * N_TR:
        raise ValueError("image size does not match the dimensions")
    check(lib().pg_write_nifti_mag_phase(str(filename).encode(), d.ptr, Nx, Ny, Nz, Nslab, N_TR,
                                         _sh(stream)))


def read_nifti(path):
    """Header fields and float32 data of a NIfTI-1 single file written above
    (offsets of the NIfTI-1 standard)."""
    raw = open(path, "rb").read()
    hdr = {
        "sizeof_hdr": int(np.frombuffer(raw, "<i4", 1, 0)[0]),
        "dim": np.frombuffer(raw, "<i2", 8, 40).tolist(),
        "datatype": int(np.frombuffer(raw, "<i2", 1, 70)[0]),
        "bitpix": int(np.frombuffer(raw, "<i2", 1, 72)[0]),
        "pixdim": np.frombuffer(raw, "<f4", 8, 76).tolist(),
        "vox_offset": float(np.frombuffer(raw, "<f4", 1, 108)[0]),
        "scl_slope": float(np.frombuffer(raw, "<f4", 1, 112)[0]),
        "xyzt_units": raw[123],
        "magic": raw[344:348],
    }
    off = int(hdr["vox_offset"])
    return hdr, np.frombuffer(raw, "<f4", offset=off)


__all__ = ["AcqTracking", "get_complete_acq_data", "write_nifti_real", "write_nifti_mag_phase",
           "read_nifti", "COUNTERS", "DeviceArray"]
