"""ctypes mirror of include/eprem_hip.h.

The structures here must match the C header field for field; the
`test_abi_layout` test checks their sizes against the compiled library.
"""
import ctypes as C

EP_OK = 0
EP_ERR_INVALID_ARG = -1
EP_ERR_NOT_INIT = -2
EP_ERR_HIP = -3
EP_ERR_NAN = -4
EP_ERR_INF = -5
EP_ERR_DS_UNDERFLOW = -6
EP_ERR_UNSUPPORTED = -7
EP_ERR_STREAM_SHORT = -8
EP_ERR_NO_DEVICE = -9

ERR_NAMES = {
    EP_OK: "EP_OK",
    EP_ERR_INVALID_ARG: "EP_ERR_INVALID_ARG",
    EP_ERR_NOT_INIT: "EP_ERR_NOT_INIT",
    EP_ERR_HIP: "EP_ERR_HIP",
    EP_ERR_NAN: "EP_ERR_NAN",
    EP_ERR_INF: "EP_ERR_INF",
    EP_ERR_DS_UNDERFLOW: "EP_ERR_DS_UNDERFLOW",
    EP_ERR_UNSUPPORTED: "EP_ERR_UNSUPPORTED",
    EP_ERR_STREAM_SHORT: "EP_ERR_STREAM_SHORT",
    EP_ERR_NO_DEVICE: "EP_ERR_NO_DEVICE",
}

EP_LAYOUT_REFERENCE = 0

EP_OP_STREAM = 0x1
EP_OP_FOCUS = 0x2
EP_OP_CHANGE = 0x4
EP_OP_SEEDCHECK = 0x8
EP_OP_SHELL = 0x10
EP_OP_DRIFT = 0x20
EP_OP_ALL = 0x3F


class EpDims(C.Structure):
    _fields_ = [
        ("num_streams", C.c_int32),
        ("num_nodes", C.c_int32),
        ("num_energies", C.c_int32),
        ("num_mu", C.c_int32),
    ]


class EpParams(C.Structure):
    _fields_ = [
        ("rScale", C.c_double),
        ("tDel", C.c_double),
        ("numEpSteps", C.c_int32),
        ("mfpType", C.c_int32),
        ("lamo", C.c_double),
        ("mfpRadialPower", C.c_double),
        ("mfpBPower", C.c_double),
        ("mhdBsAu", C.c_double),
        ("rigidityPower", C.c_double),
        ("eMin", C.c_double),
        ("eMax", C.c_double),
        ("mass", C.c_double),
        ("charge", C.c_double),
        ("dsh_min", C.c_double),
        ("dampCooling", C.c_double),
        ("boundaryFunctAmplitude", C.c_double),
        ("boundaryFunctXi", C.c_double),
        ("boundaryFunctGamma", C.c_double),
        ("boundaryFunctBeta", C.c_double),
        ("boundaryFunctEcutoff", C.c_double),
        ("boundaryFunctR0", C.c_double),
        ("useParallelDiffusion", C.c_int32),
        ("useAdiabaticFocus", C.c_int32),
        ("useAdiabaticChange", C.c_int32),
        ("checkSeedPopulation", C.c_int32),
        ("useEPBoundary", C.c_int32),
        ("useBoundaryFunction", C.c_int32),
        ("adiabaticFocusAlg", C.c_int32),
        ("adiabaticChangeAlg", C.c_int32),
    ]


_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)


class EpNodeSoa(C.Structure):
    _fields_ = [(n, _DP) for n in ("x", "y", "z", "rmag", "bmag", "dlnb", "dlnn", "dupar")]


class EpDiag(C.Structure):
    _fields_ = [
        ("max_subcycles_change", C.c_int32),
        ("max_subcycles_focusing", C.c_int32),
        ("max_nsteps_stream", C.c_int32),
        ("err_code", C.c_int32),
        ("min_tau", C.c_double),
        ("err_stream", C.c_int32),
        ("err_node", C.c_int32),
        ("err_energy", C.c_int32),
        ("err_mu", C.c_int32),
        ("cell_updates", C.c_int64),
        ("device_ms", C.c_double),
        ("stream_ms", C.c_double),
        ("focus_ms", C.c_double),
        ("change_ms", C.c_double),
        ("prep_ms", C.c_double),
        ("stream_sweeps", C.c_int64),
        ("focus_sweeps", C.c_int64),
        ("change_sweeps", C.c_int64),
        ("stream_launches", C.c_int32),
        ("heavy_nodes", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class EpSpawn(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in ("x", "y", "z", "rmag", "f")]


class EpObservers(C.Structure):
    _fields_ = [
        ("num", C.c_int32),
        ("r_au", _DP),
        ("theta", _DP),
        ("phi", _DP),
        ("idw_p", C.c_double),
    ]


class EpMasParams(C.Structure):
    _fields_ = [
        ("s_cor", C.c_double),
        ("phi_offset", C.c_double),
        ("mas_radial_max", C.c_double),
        ("mas_vmin", C.c_double),
    ]


class EpMhdParams(C.Structure):
    _fields_ = [
        ("mode", C.c_int32),
        ("mhdUs", C.c_double),
        ("mhdNsAu", C.c_double),
        ("mhdBsAu", C.c_double),
        ("omegaSun", C.c_double),
        ("mas", EpMasParams),
    ]


class EpMasPrev(C.Structure):
    _fields_ = [(n, _DP) for n in ("rold_x", "rold_y", "rold_z", "br", "btheta", "bphi",
                                   "vr", "vtheta", "vphi", "density")]


class EpSurrogateStats(C.Structure):
    _fields_ = [
        ("network_slabs", C.c_int64),
        ("fallback_slabs", C.c_int64),
        ("plain_slabs", C.c_int64),
        ("network_cell_subcycles", C.c_int64),
        ("rejected_stream_substeps", C.c_int32),
        ("pad_", C.c_int32),
        ("max_residual", C.c_double),
    ]


class EpMasOut(C.Structure):
    _fields_ = [(n, _DP) for n in ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density")]


def dptr(a):
    """Pointer to the data of a C-contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype.kind == "f" and a.dtype.itemsize == 8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


def fptr(a):
    if a is None:
        return None
    assert a.dtype.kind == "f" and a.dtype.itemsize == 4 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_FP)


def iptr(a):
    if a is None:
        return None
    assert a.dtype.kind == "i" and a.dtype.itemsize == 4 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_int32))


# Every extern "C" symbol declared in include/eprem_hip.h.
EXPORTED_SYMBOLS = (
    "ep_init",
    "ep_set_params",
    "ep_get_grids",
    "ep_set_nodes",
    "ep_upload_f",
    "ep_download_f",
    "ep_upload_f_streams",
    "ep_download_f_streams",
    "ep_init_seed_population",
    "ep_step",
    "ep_get_node_outputs",
    "ep_mas_set_level",
    "ep_interp_mhd",
    "ep_last_error",
    "ep_finalize",
    "ep_set_tuning",
    "ep_set_schedule",
    "ep_get_schedule",
    "ep_selftest_dpp",
    "ep_observe",
    "ep_observe_noalloc",
    "ep_stream_flux",
    "ep_set_positions",
    "ep_update_mhd",
    "ep_autotune",
    "ep_set_graph",
    "ep_set_node_fusion",
    "ep_spawn_shell",
    "ep_move_nodes",
    "ep_move_nodes_mas",
    "ep_get_positions",
    "ep_set_precision",
    "ep_selftest_copy",
    "ep_fastmath_host",
    "ep_selftest_fastmath",
    "ep_mlp_init",
    "ep_mlp_num_params",
    "ep_mlp_get_params",
    "ep_mlp_set_params",
    "ep_mlp_forward",
    "ep_mlp_loss_grad",
    "ep_mlp_train",
    "ep_mlp_finalize",
    "ep_mlp_bench",
    "ep_surrogate_samples",
    "ep_set_surrogate",
    "ep_get_surrogate_stats",
    "ep_surrogate_probe",
    "ep_shell_operators",
)
