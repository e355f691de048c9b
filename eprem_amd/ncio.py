"""Per-stream observer files in the reference's unified-output layout.

The reference writes one netCDF file per stream, obs%06d.nc (or
flux%06d.nc with streamFluxOutput), through libnetcdf
(initObserverDataNetCDF / writeObserverDataNetCDF,
src/unifiedOutput.c:213-549).  libnetcdf is not available here, so this
module writes the same dimensions, variables, attributes and values in the
netCDF classic format with 64-bit offsets (CDF-2), which any netCDF reader
opens.  Values are stored in code units with a `scale_factor` attribute,
exactly as the reference stores them; a reader that applies scale_factor
(netCDF4, scipy) returns the physical values the reference's digest tool
hashes.

Format (netCDF classic, big-endian): magic "CDF\\x02", record count,
dimension list, global attribute list, variable list (name, dimension ids,
attributes, type, slab size, 64-bit data offset), then the fixed-size
variables one after another, then the records, each holding one slab of
every record variable in definition order.
"""
import struct

import numpy as np

from . import constants as K

_NC_DIMENSION, _NC_VARIABLE, _NC_ATTRIBUTE = 0x0A, 0x0B, 0x0C
_NC_CHAR, _NC_INT, _NC_DOUBLE = 2, 4, 6
_TYPES = {_NC_INT: (">i4", 4), _NC_DOUBLE: (">f8", 8)}


def _pad4(b):
    return b + b"\0" * ((4 - len(b) % 4) % 4)


def _name(s):
    b = s.encode()
    return struct.pack(">I", len(b)) + _pad4(b)


def _attrs(attrs):
    if not attrs:
        return struct.pack(">II", 0, 0)
    out = struct.pack(">II", _NC_ATTRIBUTE, len(attrs))
    for key, val in attrs:
        out += _name(key)
        if isinstance(val, str):
            b = val.encode()
            out += struct.pack(">II", _NC_CHAR, len(b)) + _pad4(b)
        else:
            out += struct.pack(">II", _NC_DOUBLE, 1) + struct.pack(">d", float(val))
    return out


class _Var:
    def __init__(self, name, dims, nctype, attrs, data):
        self.name, self.dims, self.nctype, self.attrs = name, dims, nctype, attrs
        self.data = np.asarray(data, dtype=_TYPES[nctype][0])


def write_classic(path, dims, variables):
    """dims: [(name, length or None for the record dimension)];
    variables: [_Var] in definition order.  A record variable's data has the
    record count as its first axis."""
    dim_index = {d[0]: i for i, d in enumerate(dims)}
    rec_dim = [d[0] for d in dims if d[1] is None]
    rec_dim = rec_dim[0] if rec_dim else None
    nrec = 0
    for v in variables:
        if v.dims and v.dims[0] == rec_dim:
            nrec = max(nrec, v.data.shape[0])

    def slab_bytes(v):
        shape = v.data.shape[1:] if (v.dims and v.dims[0] == rec_dim) else v.data.shape
        n = int(np.prod(shape, dtype=np.int64)) * _TYPES[v.nctype][1]
        return n + (4 - n % 4) % 4

    def header(begins):
        h = b"CDF\x02" + struct.pack(">I", nrec)
        h += struct.pack(">II", _NC_DIMENSION, len(dims))
        for name, length in dims:
            h += _name(name) + struct.pack(">I", 0 if length is None else length)
        h += _attrs([])
        h += struct.pack(">II", _NC_VARIABLE, len(variables))
        for v, b in zip(variables, begins):
            h += _name(v.name) + struct.pack(">I", len(v.dims))
            for d in v.dims:
                h += struct.pack(">I", dim_index[d])
            h += _attrs(v.attrs)
            h += struct.pack(">II", v.nctype, slab_bytes(v)) + struct.pack(">Q", b)
        return h

    hlen = len(header([0] * len(variables)))
    fixed = [v for v in variables if not (v.dims and v.dims[0] == rec_dim)]
    record = [v for v in variables if v.dims and v.dims[0] == rec_dim]
    begins, off = {}, hlen
    for v in fixed:
        begins[v.name] = off
        off += slab_bytes(v)
    rec_start = off
    for v in record:
        begins[v.name] = off
        off += slab_bytes(v)
    recsize = off - rec_start
    with open(path, "wb") as fh:
        fh.write(header([begins[v.name] for v in variables]))
        for v in fixed:
            b = v.data.tobytes()
            fh.write(b + b"\0" * (slab_bytes(v) - len(b)))
        for r in range(nrec):
            for v in record:
                # a slice keeps the big-endian dtype (a scalar would not)
                b = v.data[r:r + 1].tobytes() if r < v.data.shape[0] else b""
                fh.write(b + b"\0" * (slab_bytes(v) - len(b)))
    return recsize


def observer_file_vars(raw, r_scale, stream_flux=False, pre_eruption=0.0):
    """The reference's variable set for one stream file, from code-unit
    arrays: raw["time"] [T], raw["phiOffset"] [T] (default zeros), raw["mu"]
    [M], raw["mass"], raw["charge"] [species], raw["egrid"], raw["vgrid"] [E],
    raw[R, T, P, Br, Bt, Bp, Vr, Vt, Vp, Rho] [T, shell], and raw["Dist"]
    [T, shell, species, E, M] (or raw["flux"] [T, shell, species, E])."""
    nt = len(raw["time"])
    nshell = raw["R"].shape[1]
    dims = [("time", None), ("shell", nshell), ("species", len(raw["mass"])),
            ("energy", len(raw["egrid"])), ("mu", len(raw["mu"]))]
    day = K.DAY
    b_scale = K.MHD_B_NORM * 1.0e5
    v_scale = K.C / 1.0e5
    V = _Var
    out = [
        V("preEruption", [], _NC_DOUBLE, [("units", "julian date"), ("scale_factor", day)], pre_eruption),
        V("time", ["time"], _NC_DOUBLE, [("units", "julian date"), ("scale_factor", day)], raw["time"]),
        V("phiOffset", ["time"], _NC_DOUBLE, [("units", "julian date"), ("scale_factor", day)],
          raw.get("phiOffset", np.zeros(nt))),
        V("shell", ["shell"], _NC_INT, [("units", "shell")], np.arange(nshell)),
        V("mu", ["mu"], _NC_DOUBLE, [("units", "cos(mu)")], raw["mu"]),
        V("mass", ["species"], _NC_DOUBLE, [("units", "nucleon")], raw["mass"]),
        V("charge", ["species"], _NC_DOUBLE, [("units", "e-")], raw["charge"]),
        V("egrid", ["energy"], _NC_DOUBLE, [("units", "MeV"), ("scale_factor", K.MP * K.C * K.C / K.MEV)],
          raw["egrid"]),
        V("vgrid", ["energy"], _NC_DOUBLE, [("units", "km/s"), ("scale_factor", v_scale)], raw["vgrid"]),
        V("R", ["time", "shell"], _NC_DOUBLE, [("units", "au"), ("scale_factor", r_scale)], raw["R"]),
        V("T", ["time", "shell"], _NC_DOUBLE, [("units", "radian")], raw["T"]),
        V("P", ["time", "shell"], _NC_DOUBLE, [("units", "radian")], raw["P"]),
    ]
    for k in ("Br", "Bt", "Bp"):
        out.append(V(k, ["time", "shell"], _NC_DOUBLE, [("units", "nT"), ("scale_factor", b_scale)], raw[k]))
    for k in ("Vr", "Vt", "Vp"):
        out.append(V(k, ["time", "shell"], _NC_DOUBLE, [("units", "km/s"), ("scale_factor", v_scale)], raw[k]))
    out.append(V("Rho", ["time", "shell"], _NC_DOUBLE,
                 [("units", "cm^-3"), ("scale_factor", K.MHD_DENSITY_NORM)], raw["Rho"]))
    if stream_flux:
        scale = ((K.MHD_DENSITY_NORM * K.C) / (K.MP * K.C * K.C)) * K.MEV
        out.append(V("flux", ["time", "shell", "species", "energy"], _NC_DOUBLE,
                     [("units", "# / cm^2 s sr MeV"), ("scale_factor", scale)], raw["flux"]))
    else:
        out.append(V("Dist", ["time", "shell", "species", "energy", "mu"], _NC_DOUBLE,
                     [("units", "s^3/km^6"), ("scale_factor", 1.0 / 27.0)], raw["Dist"]))
    return dims, out


def write_observer_file(path, raw, r_scale, stream_flux=False, pre_eruption=0.0):
    """Write one stream's file (obs%06d.nc layout)."""
    dims, variables = observer_file_vars(raw, r_scale, stream_flux, pre_eruption)
    write_classic(path, dims, variables)


def stream_file_name(stream, stream_flux=False):
    """The reference's per-stream file name (src/unifiedOutput.c:112-121)."""
    return ("flux%06d.nc" if stream_flux else "obs%06d.nc") % stream
