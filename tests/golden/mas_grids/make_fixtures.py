"""Writes the MAS fixtures of this folder from the reference's test inputs.

The reference's mas_wind and mas_cme cases read MAS output as HDF5 files,
one per field and time level (testsuite/mas_*/input/data/{bp,bt,br,vp,vt,vr,
rho}NNN.h5; format in src/masIO.c:34-240, file names in
src/readMAS.c:1005-1069).  This script dumps them with mas_h5dump.c, built
against the HDF5 C library in /opt/conda, into

  mas_wind_levels.npz  meshes and data of time levels 1 and 2 of mas_wind
                       (seven fields, each on its own staggered mesh)
  mas_cme_meshes.npz   the seven staggered meshes of mas_cme level 1

Keys: "<level>_<field>_<r|t|p|data>", field in bp bt br vp vt vr rho, data
in the files' [phi][theta][r] order (r fastest).  Run it in a container
that holds the reference at /root/reference; the tests read only the .npz
files.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/testsuite"
FIELDS = ("bp", "bt", "br", "vp", "vt", "vr", "rho")


def build(tmp):
    exe = os.path.join(tmp, "mas_h5dump")
    subprocess.check_call(["gcc", "-O1", os.path.join(HERE, "mas_h5dump.c"), "-I/opt/conda/include",
                           "-L/opt/conda/lib", "-lhdf5", "-Wl,-rpath,/opt/conda/lib", "-o", exe])
    return exe


def dump(exe, path, tmp):
    out = os.path.join(tmp, "dump.bin")
    subprocess.check_call([exe, path, out])
    raw = open(out, "rb").read()
    n1, n2, n3 = np.frombuffer(raw[:12], np.int32)
    v = np.frombuffer(raw[12:], np.float32)
    r, t, p = v[:n1], v[n1:n1 + n2], v[n1 + n2:n1 + n2 + n3]
    data = v[n1 + n2 + n3:]
    assert data.size == n1 * n2 * n3
    return r.copy(), t.copy(), p.copy(), data.copy()


def main():
    with tempfile.TemporaryDirectory() as tmp:
        exe = build(tmp)
        out = {}
        for level in (1, 2):
            for f in FIELDS:
                r, t, p, d = dump(exe, os.path.join(REF, "mas_wind/input/data/%s%03d.h5" % (f, level)), tmp)
                for k, a in zip("rtp", (r, t, p)):
                    out["%d_%s_%s" % (level, f, k)] = a
                out["%d_%s_data" % (level, f)] = d
        np.savez_compressed(os.path.join(HERE, "mas_wind_levels.npz"), **out)
        out = {}
        for f in FIELDS:
            r, t, p, _ = dump(exe, os.path.join(REF, "mas_cme/input/data/%s001.h5" % f), tmp)
            for k, a in zip("rtp", (r, t, p)):
                out["1_%s_%s" % (f, k)] = a
        np.savez_compressed(os.path.join(HERE, "mas_cme_meshes.npz"), **out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
