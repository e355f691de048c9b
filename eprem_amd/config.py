"""EPREM input-file reader and parameter conversion (host side).

The reference reads its input with libconfig (src/configuration.c:42-219)
and then converts to code units (setRuntimeConstants,
src/configuration.c:324-348).  libconfig is not available here, so this
module parses the subset of its syntax that EPREM input files use:
``key = value`` settings with optional ``;``/``,`` terminators, ``#``,
``//`` and ``/* */`` comments, integers, floats, strings and ``[..]``
arrays.

libconfig types matter: ``config_lookup_float`` fails on an integer literal
and ``config_lookup_int`` fails on a float literal (auto-conversion is off
by default), in which case the reference silently falls back to the
default.  `read_config` mirrors that and records such keys in
``cfg["_type_mismatch"]``.
"""
import re

from . import constants as K

BADINT = 2147483647
BADVALUE = 1.0e33
VERYSMALL = 1.0e-33


class ConfigError(ValueError):
    pass


_TOKEN = re.compile(
    r"""\s*(?:
        (?P<str>"(?:[^"\\]|\\.)*")
      | (?P<num>[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?L?)
      | (?P<lb>\[) | (?P<rb>\]) | (?P<comma>,)
      | (?P<bool>true|false|TRUE|FALSE)
    )""",
    re.VERBOSE,
)


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    out = []
    for line in text.splitlines():
        # '#' and '//' start comments outside strings
        res, in_str = [], False
        i = 0
        while i < len(line):
            ch = line[i]
            if ch == '"':
                in_str = not in_str
            if not in_str and (ch == "#" or line.startswith("//", i)):
                break
            res.append(ch)
            i += 1
        out.append("".join(res))
    return "\n".join(out)


def _parse_scalar(tok):
    if tok.startswith('"'):
        return bytes(tok[1:-1], "utf-8").decode("unicode_escape")
    if tok in ("true", "TRUE"):
        return True
    if tok in ("false", "FALSE"):
        return False
    t = tok.rstrip("L")
    if re.fullmatch(r"[-+]?\d+", t):
        return int(t)
    return float(t)


def parse_libconfig(text):
    """Return {key: value} for an EPREM input file (flat settings only)."""
    text = _strip_comments(text)
    settings = {}
    pos = 0
    key_re = re.compile(r"\s*([A-Za-z*][-A-Za-z0-9_*]*)\s*[=:]\s*")
    while True:
        m = re.compile(r"\s*").match(text, pos)
        pos = m.end()
        if pos >= len(text):
            break
        km = key_re.match(text, pos)
        if not km:
            raise ConfigError("syntax error near: %r" % text[pos:pos + 40])
        key = km.group(1)
        pos = km.end()
        if text.startswith("[", pos):
            pos += 1
            vals = []
            while True:
                tm = _TOKEN.match(text, pos)
                if not tm:
                    raise ConfigError("bad array for %s" % key)
                pos = tm.end()
                if tm.group("rb"):
                    break
                if tm.group("comma"):
                    continue
                vals.append(_parse_scalar(tm.group(0).strip()))
            value = vals
        else:
            tm = _TOKEN.match(text, pos)
            if not tm or tm.group("lb") or tm.group("rb") or tm.group("comma"):
                raise ConfigError("bad value for %s" % key)
            value = _parse_scalar(tm.group(0).strip())
            pos = tm.end()
        tm = re.compile(r"\s*[;,]?").match(text, pos)
        pos = tm.end()
        settings[key] = value
    return settings


# (kind, default, min, max) in the order of getParams
# (src/configuration.c:58-218).  Defaults that depend on other keys are
# callables evaluated on the partially built config.
_THIRD = 1.0 / 3.0
SPEC = [
    ("numNodesPerStream", "int", 1, 1, BADINT),
    ("numRowsPerFace", "int", 2, 1, BADINT),
    ("numColumnsPerFace", "int", 2, 1, BADINT),
    ("numEnergySteps", "int", 20, 2, BADINT),
    ("numMuSteps", "int", 20, 2, BADINT),
    ("rScale", "float", 0.005, VERYSMALL, BADVALUE),
    ("flowMag", "float", 400.0e5, VERYSMALL, BADVALUE),
    ("mhdDensityAu", "float", 8.30, VERYSMALL, BADVALUE),
    ("mhdBAu", "float", 1.60e-5, VERYSMALL, BADVALUE),
    ("simStartTime", "float", 0.0, 0.0, BADVALUE),
    ("tDel", "float", 0.01041666666667, VERYSMALL, BADVALUE),
    ("simStopTime", "float", lambda c: c["simStartTime"] + c["tDel"], lambda c: c["simStartTime"], BADVALUE),
    ("numEpSteps", "int", 30, 1, BADINT),
    ("aziSunStart", "float", 0.0, 0.0, BADVALUE),
    ("omegaSun", "float", (0.004144 / K.MAS_TIME_NORM) * K.TAU, 0.0, BADVALUE),
    ("dsh_min", "float", 5.0e-5, VERYSMALL, BADVALUE),
    ("dsh_hel_min", "float", 2.5e-4, VERYSMALL, BADVALUE),
    ("kperxkpar", "float", 0.01, VERYSMALL, BADVALUE),
    ("lamo", "float", 1.0, VERYSMALL, BADVALUE),
    ("mfpType", "int", 1, 1, 2),
    ("mfpRadialPower", "float", 2.0, VERYSMALL, BADVALUE),
    ("mfpBPower", "float", 0.75, VERYSMALL, BADVALUE),
    ("rigidityPower", "float", _THIRD, VERYSMALL, BADVALUE),
    ("focusingLimit", "float", 1.0, 0.0, 1.0),
    ("eMin", "float", 1.0, VERYSMALL, BADVALUE),
    ("eMax", "float", 1000.0, lambda c: c["eMin"], BADVALUE),
    ("useStochastic", "int", 0, 0, 1),
    ("useEPBoundary", "int", 1, 0, 1),
    ("checkSeedPopulation", "int", 1, 0, 1),
    ("seedFunctionTest", "int", 0, 0, 1),
    ("fluxLimiter", "int", 1, 0, 1),
    ("gammaEhigh", "float", 0.0, -BADVALUE, BADVALUE),
    ("gammaElow", "float", 0.0, -BADVALUE, BADVALUE),
    ("FailModeDump", "int", 1, 0, 1),
    ("outputFloat", "int", 0, 0, 1),
    ("unifiedOutput", "int", 1, 0, 1),
    ("unifiedOutputTime", "float", 0.0, 0.0, BADVALUE),
    ("pointObserverOutput", "int", 0, 0, 1),
    ("pointObserverOutputTime", "float", 0.0, 0.0, BADVALUE),
    ("streamFluxOutput", "int", 0, 0, 1),
    ("streamFluxOutputTime", "float", 0.0, 0.0, BADVALUE),
    ("subTimeCouple", "int", 0, 0, 1),
    ("epremDomain", "int", 0, 0, 1),
    ("epremDomainOutputTime", "float", 0.0, 0.0, BADVALUE),
    ("unstructuredDomain", "int", 0, 0, 1),
    ("unstructuredDomainOutputTime", "float", 0.0, 0.0, BADVALUE),
    ("useAdiabaticChange", "int", 1, 0, 1),
    ("dampCooling", "float", 0.0, 0.0, 1.0),
    ("useAdiabaticFocus", "int", 1, 0, 1),
    ("useShellDiffusion", "int", 0, 0, 1),
    ("useParallelDiffusion", "int", 1, 0, 1),
    ("useDrift", "int", 0, 0, 1),
    ("numSpecies", "int", 1, 0, 100),
    ("numObservers", "int", 0, 0, 1000),
    ("idw_p", "float", 3.0, VERYSMALL, BADVALUE),
    ("masTriLinear", "int", 1, 0, 1),
    ("masCouple", "int", 0, 0, 1),
    ("masCorRotateFake", "int", 0, 0, 1),
    ("masHelCouple", "int", 0, 0, 1),
    ("masNumFiles", "int", 0, 0, 32767),
    ("masHelNumFiles", "int", 0, 0, 32767),
    ("useMasSteadyStateDt", "int", 1, 0, 1),
    ("masSteadyState", "int", 1, 0, 1),
    ("masDigits", "int", 3, 0, 32767),
    ("masHelDigits", "int", 3, 0, 32767),
    ("masCoupledTime", "int", 1, 0, 1),
    ("masStartTime", "float", 0.0, 0.0, BADVALUE),
    ("epEquilibriumCalcDuration", "float", 0.0, 0.0, BADVALUE),
    ("preEruptionDuration", "float", 0.0, 0.0, BADVALUE),
    ("masRadialMin", "float", 0.0, 0.0, BADVALUE),
    ("masRadialMax", "float", 0.0, 0.0, BADVALUE),
    ("masHelRadialMin", "float", 0.0, 0.0, BADVALUE),
    ("masHelRadialMax", "float", 0.0, 0.0, BADVALUE),
    ("masVmin", "float", 50.0e5, 0.0, BADVALUE),
    ("masInitFromOuterBoundary", "int", 2, 0, 2),
    ("masInitMonteCarlo", "int", 0, 0, 1),
    ("masInitRadius", "float", 0.0, 0.0, BADVALUE),
    ("masInitTimeStep", "float", 0.000011574074074, 0.0, BADVALUE),
    ("useManualStreamSpawnLoc", "int", 0, 0, 1),
    ("parallelFlow", "float", 0.0, 0.0, BADVALUE),
    ("fieldAligned", "int", 0, 0, 1),
    ("epCalcStartTime", "float", lambda c: c["simStartTime"], 0.0, BADVALUE),
    ("masRotateSolution", "int", 1, 0, 1),
    ("useBoundaryFunction", "int", 1, 0, 1),
    ("boundaryFunctionInitDomain", "int", 1, 0, 1),
    ("boundaryFunctAmplitude", "float", 1.0, VERYSMALL, BADVALUE),
    ("boundaryFunctXi", "float", 1.0, 0.0, BADVALUE),
    ("boundaryFunctGamma", "float", 2.0, 0.0, BADVALUE),
    ("boundaryFunctBeta", "float", 1.7, 0.0, BADVALUE),
    ("boundaryFunctEcutoff", "float", 1.0, 0.0, BADVALUE),
    ("boundaryFunctR0", "float", 1.0, lambda c: c["rScale"], BADVALUE),
    ("shockSolver", "int", 0, 0, 1),
    ("shockDetectPercent", "float", 1.0, 0.0, BADVALUE),
    ("minInjectionEnergy", "float", 0.01, VERYSMALL, BADVALUE),
    ("shockInjectionFactor", "float", 1.0, 0.0, BADVALUE),
    ("idealShock", "int", 0, 0, 1),
    ("dumpFreq", "int", 1, 0, 1000000),
    ("outputRestart", "int", 0, 0, 1000000),
    ("dumpOnAbort", "int", 0, 0, 1),
    ("saveRestartFile", "int", 0, 0, 1),
    ("adiabaticChangeAlg", "int", 1, 1, 3),
    ("adiabaticFocusAlg", "int", 1, 1, 3),
]


def read_config(text, n_procs=1):
    """Build the resolved config (input-file units) like getParams().

    Raises ConfigError where the reference panics on a range check
    (readInt/readDouble, src/configuration.c:222-263).
    """
    raw = parse_libconfig(text)
    cfg = {"_type_mismatch": []}
    for key, kind, default, lo, hi in SPEC:
        if key == "numNodesPerStream":
            default, lo = n_procs, n_procs
        val = raw.get(key, None)
        if val is not None and kind == "float" and not isinstance(val, float):
            cfg["_type_mismatch"].append(key)
            val = None
        if val is not None and kind == "int" and (isinstance(val, bool) or not isinstance(val, int)):
            cfg["_type_mismatch"].append(key)
            val = None
        if val is None:
            val = default(cfg) if callable(default) else default
        lo_v = lo(cfg) if callable(lo) else lo
        if val < lo_v or val > hi:
            raise ConfigError("%s is out of the acceptable range: %r <= %r <= %r" % (key, lo_v, val, hi))
        cfg[key] = val
    ns = cfg["numSpecies"]
    for key in ("mass", "charge"):
        arr = raw.get(key)
        if ns > 0 and arr is not None:
            cfg[key] = [float(v) for v in (arr if isinstance(arr, list) else [arr])][:ns]
        else:
            cfg[key] = [1.0]
    for key in ("obsR", "obsTheta", "obsPhi"):
        arr = raw.get(key)
        cfg[key] = [float(v) for v in arr] if (cfg["numObservers"] > 0 and arr) else [0.0]
    cfg["masDirectory"] = raw.get("masDirectory", " ")
    return cfg


def ep_params(cfg):
    """EpParams in code units (setRuntimeConstants, src/configuration.c:324-348)."""
    from .abi import EpParams

    if cfg["numSpecies"] != 1:
        raise ConfigError("only numSpecies = 1 is supported")
    p = EpParams()
    p.rScale = cfg["rScale"]
    p.tDel = cfg["tDel"] / K.DAY
    p.numEpSteps = cfg["numEpSteps"]
    p.mfpType = cfg["mfpType"]
    p.lamo = cfg["lamo"]
    p.mfpRadialPower = cfg["mfpRadialPower"]
    p.mfpBPower = cfg["mfpBPower"]
    p.mhdBsAu = cfg["mhdBAu"] / K.MHD_B_NORM
    p.rigidityPower = cfg["rigidityPower"]
    p.eMin = cfg["eMin"]
    p.eMax = cfg["eMax"]
    p.mass = cfg["mass"][0]
    p.charge = cfg["charge"][0]
    p.dsh_min = cfg["dsh_min"]
    p.dampCooling = cfg["dampCooling"]
    p.boundaryFunctAmplitude = cfg["boundaryFunctAmplitude"]
    p.boundaryFunctXi = cfg["boundaryFunctXi"]
    p.boundaryFunctGamma = cfg["boundaryFunctGamma"]
    p.boundaryFunctBeta = cfg["boundaryFunctBeta"]
    p.boundaryFunctEcutoff = cfg["boundaryFunctEcutoff"]
    p.boundaryFunctR0 = cfg["boundaryFunctR0"]
    p.useParallelDiffusion = cfg["useParallelDiffusion"]
    p.useAdiabaticFocus = cfg["useAdiabaticFocus"]
    p.useAdiabaticChange = cfg["useAdiabaticChange"]
    p.checkSeedPopulation = cfg["checkSeedPopulation"]
    p.useEPBoundary = cfg["useEPBoundary"]
    p.useBoundaryFunction = cfg["useBoundaryFunction"]
    p.adiabaticFocusAlg = cfg["adiabaticFocusAlg"]
    p.adiabaticChangeAlg = cfg["adiabaticChangeAlg"]
    return p


# Options of the reference that change the analytic-wind host loop
# (eprem_amd.driver.ParkerRun) and that it does not carry out: with one of
# them set, a run would silently differ from the reference's.
#   masCouple          MAS background instead of the analytic wind
#                      (src/eprem.c:472-480, src/flow.c:48-77)
#   idealShock         the cone test's shock added to the wind (src/flow.c:
#                      280-403 with config.idealShock)
#   useShellDiffusion  DiffuseShellData / DriftShellData: the library has
#   useDrift           them (ep_shell_operators) but the host loop does not
#                      call them
#   masCorRotateFake   the fake-corotation B_phi term of masWind
#                      (src/masInterp.c:1039-1047), which the library's
#                      masWind does not add
UNSUPPORTED_BY_HOST_LOOP = ("masCouple", "idealShock", "useShellDiffusion", "useDrift", "masCorRotateFake")


def check_host_loop_options(cfg):
    """Raise ConfigError for every option the analytic-wind host loop does
    not implement (UNSUPPORTED_BY_HOST_LOOP), naming them all."""
    bad = [k for k in UNSUPPORTED_BY_HOST_LOOP if cfg.get(k, 0)]
    if bad:
        raise ConfigError("options not implemented by the analytic-wind host loop: " +
                          ", ".join("%s=%r" % (k, cfg[k]) for k in bad))


def default_params(**over):
    """EpParams from the reference defaults, overridden by keyword."""
    cfg = read_config("")
    for k, v in over.items():
        cfg[k] = v
    return ep_params(cfg)


def stream_count(cfg):
    return 6 * cfg["numRowsPerFace"] * cfg["numColumnsPerFace"]


__all__ = ["ConfigError", "parse_libconfig", "read_config", "ep_params", "default_params",
           "stream_count", "check_host_loop_options", "UNSUPPORTED_BY_HOST_LOOP"]
