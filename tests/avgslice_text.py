"""Expected text of write_avgslice's records (011_scalars2.f90:393-442,
fmt_5168 = "(1400(E14.5))"), for the tests of the slice writers.  The E14.5
edit descriptor is restated from the Fortran standard's E editing and
pinned by hand-written fields in tests/test_oracle_restart.py."""
import math

# file names as flow_slice / theta_slice pass them (011_scalars2.f90:203-244, 305-311)
NAMES = {
    "U": "u", "V": "v", "W": "w", "P": "p", "DUDT": "dudt", "DVDT": "dvdt", "DWDT": "dwdt",
    "DUDZ": "dudz", "DVDZ": "dvdz", "U2": "u2", "V2": "v2", "W2": "w2", "P2": "p2", "UV": "uv",
    "UW": "uw", "VW": "vw", "WP": "wp", "TXX": "txx", "TXY": "txy", "TXZ": "txz", "TYY": "tyy",
    "TYZ": "tyz", "TZZ": "tzz", "U3": "u3", "V3": "v3", "W3": "w3", "U2_FLUC": "u2fluc",
    "V2_FLUC": "v2fluc", "UV_FLUC": "uvfluc", "U2W_FLUC": "u2wfluc", "V2W_FLUC": "v2wfluc",
    "DISSIP": "dissip", "UTXZ_FLUC": "utxz_fluc", "VTYZ_FLUC": "vtyz_fluc", "WTZZ": "wtzz", "CS": "Cs",
    "THETA": "theta", "THETA2": "theta2", "T2_FLUC": "Tfluc", "DTDZ": "dTdz", "SGS_T3": "sgs_t3",
    "WT": "wt", "NU_T": "Nu_t",
}
THETA_SLICES = ("THETA", "THETA2", "T2_FLUC", "DTDZ", "SGS_T3", "WT", "NU_T")


def e14_5(x):
    """Fortran E14.5: [-]0.ddddd then E+ee (|e| <= 99) or +eee, right-justified."""
    if math.isnan(x):
        s = "NaN"
    elif math.isinf(x):
        s = "-Infinity" if x < 0 else "Infinity"
    elif x == 0.0:
        s = ("-" if math.copysign(1.0, x) < 0 else "") + "0.00000E+00"
    else:
        m, e = f"{abs(x):.4e}".split("e")
        ex = int(e) + 1
        digits = m.replace(".", "")
        tail = f"E{'-' if ex < 0 else '+'}{abs(ex):02d}" if abs(ex) <= 99 else f"{'-' if ex < 0 else '+'}{abs(ex):03d}"
        s = ("-" if x < 0 else "") + "0." + digits + tail
    return s.rjust(14) if len(s) <= 14 else "*" * 14


def record_lines(values):
    """One write of fmt_5168: format reversion starts a new line after every
    1400 items."""
    return ["".join(e14_5(v) for v in values[i:i + 1400]) for i in range(0, len(values), 1400)]


def records(slice2d, time, average_dim_num):
    """The lines write_avgslice appends for one (nzt-1, nxt) window slice."""
    if average_dim_num == 1:
        return [line for row in slice2d for line in record_lines([time] + list(row))]
    means = []
    for row in slice2d:
        s = 0.0
        for v in row:      # sum(arr2d, dim=1) in x order
            s += v
        means.append(s / len(row))
    return record_lines([time] + means)
