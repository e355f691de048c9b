"""Text digest of per-stream output, in the format of the reference's
regression tool (bin/epremdigest.py:33-113): for every variable its shape,
dtype, nan-aware min/max/mean/sum/rms rounded to 6 significant figures, and
a SHA-256 of the little-endian bytes.  Values are the physical (scaled)
values a netCDF reader returns."""
import hashlib
from math import floor, log10

import numpy as np


def sigfig(x, n=6):
    if x is None:
        return "nan"
    x = float(x)
    if np.isnan(x):
        return "nan"
    if np.isinf(x):
        return "inf" if x > 0 else "-inf"
    if x == 0.0:
        return "0.0"
    d = n - int(floor(log10(abs(x)))) - 1
    return "%.6g" % round(x, d)


def var_hash(arr):
    a = np.asarray(arr)
    if a.dtype.kind == "f":
        a = a.astype("<" + a.dtype.str[1:])
    return hashlib.sha256(a.tobytes()).hexdigest()


def stats_line(arr):
    a = np.asarray(arr, dtype="float64")
    if a.size == 0:
        return "min=nan max=nan mean=nan sum=0.0 rms=nan"
    with np.errstate(invalid="ignore", divide="ignore"):
        vals = (np.nanmin(a), np.nanmax(a), np.nanmean(a), np.nansum(a), np.sqrt(np.nanmean(a * a)))
    return "min=%s max=%s mean=%s sum=%s rms=%s" % tuple(sigfig(float(v)) for v in vals)


def digest_variable(name, arr):
    arr = np.asarray(arr)
    return ["var %s" % name,
            "  shape=%s dtype=%s" % (str(arr.shape), str(arr.dtype)),
            "  " + stats_line(arr),
            "  sha256=%s" % var_hash(arr)]


def parse_digest(text):
    """{file: {var: {"shape":..., "stats":..., "sha256":...}}} from digest text."""
    out, cur_file, cur_var = {}, None, None
    for line in text.splitlines():
        if line.startswith("file "):
            cur_file = line[5:].strip()
            out[cur_file] = {}
        elif line.startswith("var "):
            cur_var = line[4:].strip()
            out[cur_file][cur_var] = {}
        elif line.startswith("  shape="):
            out[cur_file][cur_var]["shape"] = line.strip()
        elif line.startswith("  min="):
            out[cur_file][cur_var]["stats"] = line.strip()
        elif line.startswith("  sha256="):
            out[cur_file][cur_var]["sha256"] = line.strip()[7:]
    return out


def digest_file(path):
    """Digest lines of one netCDF file, variables in sorted order, scalar
    variables skipped, values scaled by their scale_factor as a netCDF
    reader returns them (bin/epremdigest.py:100-113)."""
    import os
    from scipy.io import netcdf_file

    lines = ["file %s" % os.path.basename(path)]
    with netcdf_file(path, "r", mmap=False, maskandscale=True) as ds:
        for k in sorted(ds.variables.keys()):
            v = ds.variables[k]
            if len(v.shape) == 0:
                continue
            arr = np.asarray(v[:])
            arr = arr.astype(arr.dtype.newbyteorder("="))
            lines.extend(digest_variable(k, arr))
    return lines


def digest_dir(rundir):
    """Digest text of every obs*.nc file of a run directory."""
    import glob
    import os

    lines = ["# eprem_digest granularity=global"]
    for f in sorted(glob.glob(os.path.join(rundir, "obs*.nc"))):
        lines.extend(digest_file(f))
    lines.append("")
    return "\n".join(lines)
