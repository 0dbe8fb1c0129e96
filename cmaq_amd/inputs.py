"""Seeded synthetic cell inputs for the ROS3 workloads.

Every value is a pure function of (seed, field, cell index) through a
splitmix64 hash, so any subset of cells reproduces exactly (the CPU baseline
times a subset of the same cells the GPU integrates).

Distributions (per cell):
  T ~ U(250, 310) K;  P ~ U(500, 1013) hPa, passed in atm (hPa*100/101325);
  H2O log-uniform on [1e3, 3e4] ppm;
  every species log-uniform on [1e-6, 1e2] ppb, except O3 ~ U(20, 80) ppb and
  CO ~ U(80, 300) ppb; concentrations are passed in ppm;
  short-lived species (lifetime below 0.6 ms: O1D, O3P, CB6's ROR, SAPRC07's
  TBUO, CRACMM1-AMORE's ISON; see FAST_SPECIES) start at 1e-6 ppb -- with the
  plain distribution RBSOLVER rejects down to DTMIN and stops
  (rbsolver.F:528-583) in ~58 % of cells (O1D) or ~0.1 % (ROR);
  J = clear-sky profile (CLEAR_SKY_J, 1/min) * cos(zenith), cos(zenith) ~ U(0,1),
  with 30 % of cells at night (J = 0);
  heterogeneous rates and the sea-water fraction are 0.
Workload variants:
  regional  - cells ordered layer-major over nlay layers, J * exp(-k/20);
  mixed     - 40 % clean day, 30 % night, 30 % near-source with NO and NO2
              log-uniform on [10, 500] ppb and the VOC species x10.
"""

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
    return z ^ (z >> np.uint64(31))


def uniform(seed, field, cells):
    """U[0,1) from (seed, field, cell index); cells is an int array."""
    with np.errstate(over="ignore"):
        key = _splitmix(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(field))
        h = _splitmix(key ^ (np.asarray(cells, dtype=np.uint64) * np.uint64(0xD6E8FEB86659FD93)))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _logu(u, lo, hi):
    return np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))


# Clear-sky, overhead-sun surface photolysis frequencies (1/s) assumed for the
# synthetic workloads; the reference repository holds no J table, these are
# representative mid-latitude summer values per CSQY table name.
_CLEAR_SKY_J_PER_S = {
    # CB6r5 tables
    "NO2_IUPAC10": 8.0e-3, "O3_O3P_IUPAC10": 4.0e-4, "O3_O1D_IUPAC10": 3.0e-5,
    "H2O2_IUPAC10": 7.0e-6, "NO3NO2_06": 1.8e-1, "NO3NO_06": 2.2e-2,
    "N2O5_IUPAC10": 3.0e-5, "HONO_IUPAC10": 1.5e-3, "HNO3_IUPAC10": 6.0e-7,
    "PNA_IUPAC10": 5.0e-6, "PAN_IUPAC10": 4.0e-7, "MEPX_IUPAC10": 5.0e-6,
    "NTR_IUPAC10": 1.0e-6, "FORM_R_IUPAC13": 3.0e-5, "FORM_M_IUPAC13": 4.5e-5,
    "ALD2_R_IUPAC13": 4.0e-6, "ALDX_R_IUPAC13": 1.5e-5, "GLYD_IUPAC13": 3.0e-6,
    "GLY_R_IUPAC13": 3.0e-5, "MGLY_IUPAC10": 1.2e-4, "KET_IUPAC10": 1.0e-6,
    "ACET_IUPAC10": 5.0e-7, "ISPD": 1.0e-5, "HPALD": 2.0e-4, "CL2_IUPAC04": 2.0e-3,
    "HOCL_IUPAC04": 2.5e-4, "FMCL_IUPAC04": 1.0e-5, "CLNO2_IUPAC13": 3.0e-4,
    "CLONO2_1": 3.0e-5, "CLONO2_2": 1.0e-5, "IC3ONO2": 5.0e-6, "ACRO_09": 2.0e-6,
    # CRACMM2 tables
    "O3O3P_NASA06": 4.0e-4, "O3O1D_NASA06": 3.0e-5, "H2O2_RACM2": 7.0e-6,
    "NO2_RACM2": 8.0e-3, "NO3NO_RACM2": 2.2e-2, "NO3NO2_RACM2": 1.8e-1,
    "HONO_RACM2": 1.5e-3, "HNO3_RACM2": 6.0e-7, "HNO4_RACM2": 5.0e-6,
    "HCHO_MOL_JPL19": 4.5e-5, "HCHO_RAD_JPL19": 3.0e-5, "CH3CHO_RACM2": 4.0e-6,
    "ALD_JPL19": 1.5e-5, "CH3COCH3A_JPL19": 5.0e-7, "CH3COCH3B_JPL19": 2.0e-7,
    "UALD_RACM2": 1.0e-5, "MEK_JGR19": 1.0e-6, "KET_JGR19": 1.0e-6, "HKET_RACM2": 2.0e-6,
    "MACR_RACM2": 3.0e-6, "MVK_RACM2": 5.0e-6, "GLYH2_RACM2": 1.0e-5, "GLYF_RACM2": 2.0e-5,
    "GLYHX_RACM2": 3.0e-5, "MGLY_RACM2": 1.2e-4, "BALD_RACM2": 5.0e-5, "OP1_RACM2": 5.0e-6,
    "PAA_RACM2": 1.0e-6, "ONIT_RACM2": 1.0e-6, "PAN1_RACM2": 3.0e-7, "PAN2_RACM2": 1.0e-7,
    # CB6r3, CB6r5m, CRACMM1, RACM2 and SAPRC07 tables
    "NO2_06": 8.0e-3, "O3O1D_06": 3.0e-5, "O3O3P_06": 4.0e-4, "NO3NO2_6": 1.8e-1,
    "HONO_06": 1.5e-3, "HNO3": 6.0e-7, "HNO4_06": 5.0e-6, "H2O2": 7.0e-6, "PAN": 4.0e-7,
    "HCHOR_06": 3.0e-5, "HCHOM_06": 4.5e-5, "CCHO_R": 4.0e-6, "C2CHO": 2.0e-5, "ACET_06": 5.0e-7,
    "MEK_06": 1.0e-6, "COOH": 5.0e-6, "GLY_07R": 3.0e-5, "GLY_07M": 1.0e-5, "MGLY_06": 1.2e-4,
    "BACL_07": 1.0e-4, "BALD_06": 5.0e-5, "AFG1": 1.0e-4, "MACR_06": 3.0e-6, "MVK_06": 5.0e-6,
    "NOA": 1.0e-5, "PAA": 1.0e-6, "HOCCHO_IUPAC": 3.0e-6, "CL2": 2.0e-3, "CLNO_06": 2.0e-3,
    "CLONO": 5.0e-4, "CLNO2": 3.0e-4, "HOCL_06": 2.5e-4, "CLCCHO": 1.0e-5, "CLACET": 1.0e-6,
    "FORM_R_IUPAC10": 3.0e-5, "FORM_M_IUPAC10": 4.5e-5, "ALD2_R_IUPAC10": 4.0e-6,
    "ALDX_R_IUPAC10": 1.5e-5, "GLYD_IUPAC10": 3.0e-6, "GLY_R_IUPAC10": 3.0e-5,
    "HCHO_RAD_RACM2": 3.0e-5, "ALD_RACM2": 1.5e-5, "CH3COCH3_RACM2": 5.0e-7, "KET_RACM2": 1.0e-6,
    "MEK_RACM2": 1.0e-6, "BR2_IUPAC10": 3.0e-2, "BRO_IUPAC10": 3.0e-2, "HOBR_IUPAC10": 2.0e-3,
    "BRNO2_IUPAC10": 2.0e-3, "BRONO2_M_IUPAC10": 1.0e-4, "BRONO2_R_IUPAC10": 1.0e-3,
    "BRCL_IUPAC10": 1.0e-2, "MB3_IUPAC10": 1.0e-7, "MB2C_BLIDE98": 1.0e-7, "MBC2_BLIDE98": 1.0e-7,
    "COHBR_JPL2010": 1.0e-5, "I2_IUPAC10": 1.0e-1, "HOI_IUPAC10": 1.0e-2, "IO_IUPAC10": 1.0e-1,
    "OIO_06": 1.0e-1, "INO_06": 1.0e-2, "INO2_06": 1.0e-3, "IONO2_06": 1.0e-2, "IBR_IUPAC10": 3.0e-2,
    "ICL_IUPAC10": 1.0e-2, "CH3I_IUPAC10": 5.0e-6, "MI2_IUPAC10": 3.0e-3, "MIB_IUPAC10": 1.0e-3,
    "MIC_IUPAC10": 3.0e-4,
}

VOC_SPECIES = {
    "CB6R5_AE7_AQ": ["PAR", "OLE", "IOLE", "ETH", "ETHA", "ETOH", "MEOH", "FORM", "ALD2", "ALDX",
                     "ISOP", "TERP", "APIN", "TOL", "XYLMN", "BENZENE", "NAPH", "PRPA", "ETHY", "KET",
                     "ACET", "SESQ", "SOAALK", "FORM_PRIMARY", "ALD2_PRIMARY", "BUTADIENE13",
                     "ACROLEIN", "TOLU"],
    "CRACMM2": ["ETH", "HC3", "HC5", "HC10", "ETE", "OLT", "OLI", "ACE", "BEN", "TOL", "XYL", "EBZ",
                "ISO", "API", "LIM", "HCHO", "ACD", "ALD", "ACT", "MEK", "KET", "MOH", "EOH", "ROH",
                "SESQ", "STY", "NAPH", "FURAN", "BDE13", "ACRO", "PROG"],
}

# Short-lived species start at 1e-6 ppb (the bottom of the log-uniform range)
# instead of the draw, as they would after any earlier step:
# * every species whose first-order loss frequency (unimolecular and fixed-
#   species reactions, photolysis at the clear-sky profile) exceeds 1e5 /min,
#   i.e. a lifetime below 0.6 ms, anywhere in T in [250, 310] K, P in
#   [500, 1013] hPa: O(1D) (~6.6e10 /min, below DTMIN = 1e-8 min), O(3P)
#   (~1e7 /min), the CB6 alkoxy radical ROR (~5e6 /min) and SAPRC07's TBUO
#   (~1.7e5 /min);
# * CRACMM1-AMORE's ISON, whose bimolecular loss reaches 3.5e8 /min.
# With O(1D) drawn up to 100 ppb ROS3's embedded error estimate tends to y0/2
# (it is not damped as h*lambda -> -inf) and RBSOLVER rejects down to DTMIN
# and stops (rbsolver.F:528-583) in ~58 % of cells.  With ROR drawn up to
# 100 ppb its decomposition (R133, k ~ 1.5e6 /min) removes 2.7 PAR per ROR
# through CB6's negative PAR coefficient within microseconds; where that
# exceeds the PAR present, PAR's error term takes > 95 % of ERR and the same
# failure follows in ~0.1 % of cells.  The list is re-derived from the rate
# constants by tests/test_inputs.py.
FAST_SPECIES = {
    "CB6R5_AE7_AQ": ("O1D", "O", "ROR"),
    "CB6R3_AE7_AQ": ("O1D", "O", "ROR"),
    "CB6R5M_AE7_AQ": ("O1D", "O", "ROR"),
    "CRACMM1AMORE_AQ": ("O1D", "O3P", "ISON"),
    "SAPRC07TC_AE6_AQ": ("O1D", "O3P", "TBUO"),
    "SAPRC07TIC_AE7I_AQ": ("O1D", "O3P", "TBUO"),
}
FAST_DEFAULT = ("O1D", "O3P")
FAST_LOSS_PER_MIN = 1.0e5


def sub_dtmin_species(mech):
    """Species started at 1e-6 ppb (the short-lived set above)."""
    return set(FAST_SPECIES.get(mech["mechname"], FAST_DEFAULT))

F_T, F_P, F_H2O, F_COSZ, F_NIGHT, F_REGIME, F_NO, F_NO2, F_SPC = 1, 2, 3, 4, 5, 6, 7, 8, 100


def clear_sky_j(mech):
    """Clear-sky J profile in 1/min, one value per PHOTAB entry."""
    return np.array([_CLEAR_SKY_J_PER_S[n] * 60.0 for n in mech["photab"]])


def make_cells(mech, cells, seed=0, workload="clean", nlay_cells=None, svoc_ppb=None):
    """Inputs for the given cell indices (int array).

    workload: 'clean' (config-0 distributions), 'regional' (plus J decay with
    layer; nlay_cells = cells per layer) or 'mixed' (three regimes).
    svoc_ppb: optional (lo, hi) log-uniform range for species whose name
    starts with 'VROC' (the larger-mechanism workload).
    Returns a dict of float64 arrays in the C-ABI layouts.
    """
    cells = np.asarray(cells, dtype=np.int64)
    n = cells.size
    names = mech["chemistry_spc"]
    ns = len(names)
    temp = 250.0 + 60.0 * uniform(seed, F_T, cells)
    pres = (500.0 + 513.0 * uniform(seed, F_P, cells)) * 100.0 / 101325.0
    h2o = _logu(uniform(seed, F_H2O, cells), 1.0e3, 3.0e4)
    y = np.empty((ns, n))
    fast = sub_dtmin_species(mech)
    for i, nm in enumerate(names):
        u = uniform(seed, F_SPC + i, cells)
        if nm in fast:
            v = np.full(n, 1.0e-6)
        elif nm == "O3":
            v = 20.0 + 60.0 * u
        elif nm == "CO":
            v = 80.0 + 220.0 * u
        elif svoc_ppb is not None and nm.startswith("VROC"):
            v = _logu(u, *svoc_ppb)
        else:
            v = _logu(u, 1.0e-6, 1.0e2)
        y[i] = v * 1.0e-3
    cosz = uniform(seed, F_COSZ, cells)
    night = uniform(seed, F_NIGHT, cells) < 0.30
    if workload == "mixed":
        reg = uniform(seed, F_REGIME, cells)
        night = (reg >= 0.40) & (reg < 0.70)
        source = reg >= 0.70
        idx = {nm: i for i, nm in enumerate(names)}
        for nm, fld in (("NO", F_NO), ("NO2", F_NO2)):
            y[idx[nm], source] = _logu(uniform(seed, fld, cells[source]), 10.0, 500.0) * 1.0e-3
        for nm in VOC_SPECIES.get(mech["mechname"], []):
            if nm in idx:
                y[idx[nm], source] *= 10.0
    scale = np.where(night, 0.0, cosz)
    if workload == "regional":
        k = cells // int(nlay_cells)
        scale = scale * np.exp(-k / 20.0)
    rj = clear_sky_j(mech)[:, None] * scale[None, :]
    return {"temp": temp, "pres": pres, "h2o": h2o, "rj": np.ascontiguousarray(rj),
            "y": np.ascontiguousarray(y), "night": night}


# ---- the same generator on the GPU (bench-sized workloads) ----------------
def _i64(c):
    """uint64 constant as the int64 with the same bits."""
    return c - (1 << 64) if c >= (1 << 63) else c


def _t_splitmix(x):
    import torch  # noqa: F401
    x = x + _i64(0x9E3779B97F4A7C15)
    z = x
    z = (z ^ ((z >> 30) & ((1 << 34) - 1))) * _i64(0xBF58476D1CE4E5B9)
    z = (z ^ ((z >> 27) & ((1 << 37) - 1))) * _i64(0x94D049BB133111EB)
    return z ^ ((z >> 31) & ((1 << 33) - 1))


def uniform_torch(seed, field, cells):
    """Bit-identical to uniform() for int64 torch tensors of cell indices."""
    import torch
    with np.errstate(over="ignore"):
        key = int(_splitmix(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(field)))
    h = _t_splitmix(torch.bitwise_xor(cells * _i64(0xD6E8FEB86659FD93), _i64(key)))
    return ((h >> 11) & ((1 << 53) - 1)).to(torch.float64) * (1.0 / 9007199254740992.0)


def make_cells_torch(mech, ncells, seed=0, workload="clean", nlay_cells=None, svoc_ppb=None,
                     device="cuda"):
    """make_cells for cells 0..ncells-1, generated directly in device memory."""
    import torch
    cells = torch.arange(ncells, dtype=torch.int64, device=device)
    names = mech["chemistry_spc"]
    ns = len(names)
    logu = lambda u, lo, hi: torch.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))
    out = {
        "temp": 250.0 + 60.0 * uniform_torch(seed, F_T, cells),
        "pres": (500.0 + 513.0 * uniform_torch(seed, F_P, cells)) * 100.0 / 101325.0,
        "h2o": logu(uniform_torch(seed, F_H2O, cells), 1.0e3, 3.0e4),
    }
    y = torch.empty((ns, ncells), dtype=torch.float64, device=device)
    fast = sub_dtmin_species(mech)
    for i, nm in enumerate(names):
        u = uniform_torch(seed, F_SPC + i, cells)
        if nm in fast:
            y[i] = 1.0e-6 * 1.0e-3
            continue
        if nm == "O3":
            v = 20.0 + 60.0 * u
        elif nm == "CO":
            v = 80.0 + 220.0 * u
        elif svoc_ppb is not None and nm.startswith("VROC"):
            v = logu(u, *svoc_ppb)
        else:
            v = logu(u, 1.0e-6, 1.0e2)
        y[i] = v * 1.0e-3
        del u, v
    cosz = uniform_torch(seed, F_COSZ, cells)
    night = uniform_torch(seed, F_NIGHT, cells) < 0.30
    if workload == "mixed":
        reg = uniform_torch(seed, F_REGIME, cells)
        night = (reg >= 0.40) & (reg < 0.70)
        source = reg >= 0.70
        idx = {nm: i for i, nm in enumerate(names)}
        for nm, fld in (("NO", F_NO), ("NO2", F_NO2)):
            v = logu(uniform_torch(seed, fld, cells), 10.0, 500.0) * 1.0e-3
            y[idx[nm]] = torch.where(source, v, y[idx[nm]])
        for nm in VOC_SPECIES.get(mech["mechname"], []):
            if nm in idx:
                y[idx[nm]] = torch.where(source, y[idx[nm]] * 10.0, y[idx[nm]])
    scale = torch.where(night, torch.zeros_like(cosz), cosz)
    if workload == "regional":
        k = torch.div(cells, int(nlay_cells), rounding_mode="floor").to(torch.float64)
        scale = scale * torch.exp(-k / 20.0)
    jprof = torch.as_tensor(clear_sky_j(mech), device=device)
    out["rj"] = (jprof[:, None] * scale[None, :]).contiguous()
    out["y"] = y
    out["night"] = night
    return out


# ---- CCTM-side arrays of one CHEM call -----------------------------------
F_MAP, F_ZONE, F_ICE, F_HET = 9, 10, 11, 12
_MWAIR, _MWWAT, _RGAS = 28.9628, 18.0153, 8.314510   # CONST.EXT: MWAIR, MWWAT, RGASUNIV


def make_grid(mech, ncols, nrows, nlays, seed=0, workload="clean", extra_species=64, ld_het_pad=1):
    """The REAL(4) arrays CHEM holds for a (ncols, nrows, nlays) grid, from the
    same per-cell draws as make_cells (cell = col + ncols*(row + nrows*lay)).

    CGRID has the mechanism species at a seeded permutation of CGRID indices
    among numb_mech_spc + extra_species planes (the other planes hold a
    sentinel); aerosol species (CONVERT_CONC) are stored in ug/m3 and gas
    species in ppm.  DENS = P*MWAIR/(R*T); QV from the H2O mixing ratio;
    SEAWATER_ZONE and SEAICE are seeded 2-D fields; KHETERO is
    [cell][nhetero + ld_het_pad] with seeded rates.  Returns (grid dict,
    cgrid_index (1-based), convert_conc, species_molwt, cell dict of
    make_cells)."""
    n = ncols * nrows * nlays
    cells = np.arange(n, dtype=np.int64)
    d = make_cells(mech, cells, seed=seed, workload=workload, nlay_cells=ncols * nrows)
    ns = mech["numb_mech_spc"]
    nspcsd = ns + extra_species
    perm = np.argsort(uniform(seed, F_MAP, np.arange(nspcsd, dtype=np.int64)))
    cgrid_index = (perm[:ns] + 1).astype(np.int32)
    conv = np.array([1 if c else 0 for c in mech["convert_conc"]], np.int32)
    molwt = np.array(mech["species_molwt"], np.float64)
    pres_pa = d["pres"] * 101325.0
    dens = pres_pa * _MWAIR * 1.0e-3 / (_RGAS * d["temp"])
    qv = d["h2o"] * 1.0e-6 * _MWWAT / _MWAIR
    cgrid = np.full((nspcsd, n), -1.0, np.float32)
    to_ug = molwt[:, None] * dens[None, :] / (1.0e-3 * _MWAIR)     # ppm -> ug/m3
    cgrid[cgrid_index - 1] = np.where(conv[:, None] == 1, d["y"] * to_ug, d["y"]).astype(np.float32)
    c2 = np.arange(ncols * nrows, dtype=np.int64)
    zone = np.where(uniform(seed, F_ZONE, c2) < 0.5, 0.0, uniform(seed, F_ZONE + 100, c2))
    ice = np.where(uniform(seed, F_ICE, c2) < 0.8, 0.0, 0.5 * uniform(seed, F_ICE + 100, c2))
    nhet = mech["nhetero"]
    ld = nhet + ld_het_pad
    kh = np.zeros((n, ld))
    for k in range(nhet):
        kh[:, k] = _logu(uniform(seed, F_HET + k, cells), 1.0e-6, 1.0e-1)
    grid = {"ncols": ncols, "nrows": nrows, "nlays": nlays, "nspcsd": nspcsd,
            "cgrid": cgrid, "ta": d["temp"].astype(np.float32), "dens": dens.astype(np.float32),
            "pres": pres_pa.astype(np.float32), "qv": qv.astype(np.float32),
            "seawater_zone": zone.astype(np.float32), "seaice": ice.astype(np.float32),
            "rj": d["rj"].astype(np.float32), "khetero": kh}
    return grid, cgrid_index, conv, molwt, d
