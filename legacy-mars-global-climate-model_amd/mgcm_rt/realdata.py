"""Radiation tables from the reference's own data files.

This is the start-up RADSETUP performs (code/radsetup.f90:26-46 -> SETSPV,
SETSPI, SETRAD with LAGINTERP), with the library doing the file reads and the
table builders on the device:

    tables = realdata.load_tables(data_dir, nlay)

* k-coefficients: data_dir/CO2H2O_V_12_95_INTEL and CO2H2O_IR_12_95_INTEL read
  by mgcm_rt_read_ktable (LAGINTERP's reads, code/laginterp.f90:59-69), then
  refined from the 11 reference pressures to the 51 PINT points by
  mgcm_rt_laginterp on the device (code/laginterp.f90:30-236); FZEROV/FZEROI
  come from the files' second records.
* grids: TGASREF = 50..350 K (code/setrad.f90:124-130); PGASREF as the REAL(4)
  literals 1e-6..1e4 mbar (code/setrad.f90:112-122); PFGASREF = the PINT points
  LAGINTERP returns; WREFH2O (code/modules.f90:353-355).
* Gauss weights: the single GWEIGHT set for every band (code/modules.f90:337-345).
* Planck table and DWNI: mgcm_rt_planck_table on the device (code/setspi.f90).
* solar fluxes, Rayleigh tau and the dust constants: SETSPV/SETRAD's values,
  the same as the synthetic tables use (synth.base_tables).
"""
from __future__ import annotations

import os

import numpy as np

from . import laginterp, planck_table, read_ktable, synth

VIS_FILE = "CO2H2O_V_12_95_INTEL"
IR_FILE = "CO2H2O_IR_12_95_INTEL"

# code/setrad.f90:112-122 (REAL(4) literals)
PGREF11 = np.array([1.0e-6, 1.0e-5, 1.0e-4, 1.0e-3, 1.0e-2, 1.0e-1, 1.0, 1.0e1, 1.0e2, 1.0e3, 1.0e4],
                   dtype=np.float32).astype(np.float64)
# code/setrad.f90:124-130
TGASREF = np.array([50.0, 100.0, 150.0, 200.0, 250.0, 300.0, 350.0])
# code/modules.f90:353-355
WREFH2O = np.array([1.0e-7, 1.0e-6, 1.0e-5, 1.0e-4, 1.0e-3, 1.0e-2, 1.0e-1, 2.0e-1, 3.0e-1, 4.0e-1])


def write_ktable(path: str, raw: np.ndarray, fzero: np.ndarray) -> None:
    """Writes raw CO2(7,11,10,nb,17) and FZERO(nb) as the two-record sequential
    unformatted file LAGINTERP reads (little-endian, 4-byte length markers)."""
    recs = [np.ascontiguousarray(raw, dtype="<f8").tobytes(), np.ascontiguousarray(fzero, dtype="<f8").tobytes()]
    with open(path, "wb") as f:
        for r in recs:
            n = np.array([len(r)], dtype="<u4").tobytes()
            f.write(n + r + n)


def load_tables(data_dir: str, nlay: int) -> dict:
    """The table dict of mgcm_rt.Radiation for the reference's k-distribution."""
    raw_v, fzv = read_ktable(os.path.join(data_dir, VIS_FILE), synth.NV)
    raw_i, fzi = read_ktable(os.path.join(data_dir, IR_FILE), synth.NI)
    pint, co2v = laginterp(PGREF11, raw_v)
    _, co2i = laginterp(PGREF11, raw_i)
    gw = synth.GWEIGHT_REF
    t = synth.base_tables(nlay, np.tile(gw, (synth.NV, 1)), np.tile(gw, (synth.NI, 1)))
    planck, dwni, _ = planck_table()
    t.update(co2v=co2v, co2i=co2i, pfgasref=pint, tgasref=TGASREF.copy(), wrefh2o=WREFH2O.copy(),
             fzerov=fzv, fzeroi=fzi, planckir=planck, dwni=dwni)
    return t
