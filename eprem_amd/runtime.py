"""ctypes binding of libeprem_hip.so (the C ABI in include/eprem_hip.h).

This is the product path: there is no CPU fallback.  Loading fails loudly
if the library has not been built, and `EpSolver` raises on every non-zero
return code with the library's own message (the reference panics at the
same points, src/error.c:39-49).
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import (EP_LAYOUT_REFERENCE, EP_OK, EP_OP_ALL, EpDiag, EpDims, EpMasOut, EpMasParams,
                  EpMasPrev, EpNodeSoa, EpParams, dptr, fptr, iptr)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libeprem_hip.so")

_lib = None


class EpError(RuntimeError):
    def __init__(self, code, msg, diag=None):
        super().__init__("%s (%d): %s" % (abi.ERR_NAMES.get(code, "?"), code, msg))
        self.code = code
        self.diag = diag


def lib():
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libeprem_hip.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        DP, IP = C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.ep_init.argtypes = [C.POINTER(EpDims), C.POINTER(EpParams)]
        L.ep_set_params.argtypes = [C.POINTER(EpParams)]
        L.ep_set_tuning.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        L.ep_set_schedule.argtypes = [C.c_int32]
        L.ep_get_schedule.argtypes = [C.POINTER(C.c_int32)]
        L.ep_get_grids.argtypes = [DP, DP, DP, DP, DP, DP]
        L.ep_set_nodes.argtypes = [C.POINTER(EpNodeSoa), IP]
        L.ep_upload_f.argtypes = [DP, C.c_int32]
        L.ep_download_f.argtypes = [DP, C.c_int32]
        L.ep_upload_f_streams.argtypes = [C.c_int32, C.c_int32, DP, C.c_int32]
        L.ep_download_f_streams.argtypes = [C.c_int32, C.c_int32, DP, C.c_int32]
        L.ep_init_seed_population.argtypes = []
        L.ep_step.argtypes = [C.c_uint32, C.POINTER(EpDiag)]
        L.ep_get_node_outputs.argtypes = [DP, DP, DP]
        L.ep_mas_set_level.argtypes = [C.c_int32, C.POINTER(C.POINTER(C.c_float)),
                                       C.POINTER(C.POINTER(C.c_float)), IP]
        L.ep_interp_mhd.argtypes = [C.POINTER(EpMasParams), C.POINTER(EpMasPrev),
                                    C.POINTER(EpMasOut), DP]
        L.ep_last_error.argtypes = [C.c_char_p, C.c_size_t]
        L.ep_finalize.argtypes = []
        L.ep_selftest_dpp.argtypes = [DP, DP, DP]
        L.ep_observe.argtypes = [C.POINTER(abi.EpObservers), DP, DP]
        L.ep_observe_noalloc.argtypes = [C.POINTER(abi.EpObservers), DP, DP]
        L.ep_stream_flux.argtypes = [DP]
        L.ep_set_positions.argtypes = [DP, DP, DP, DP, IP]
        L.ep_update_mhd.argtypes = [C.POINTER(abi.EpMhdParams), C.POINTER(EpMasPrev), C.POINTER(EpMasOut), DP]
        L.ep_autotune.argtypes = [IP, IP, IP, IP, DP]
        L.ep_spawn_shell.argtypes = [C.POINTER(abi.EpSpawn)]
        L.ep_move_nodes.argtypes = [C.c_double]
        L.ep_move_nodes_mas.argtypes = [C.POINTER(EpMasParams), C.c_double]
        L.ep_get_positions.argtypes = [DP, DP, DP, DP, IP]
        L.ep_set_graph.argtypes = [C.c_int32]
        L.ep_set_node_fusion.argtypes = [C.c_int32]
        L.ep_set_precision.argtypes = [C.c_int32]
        L.ep_selftest_copy.argtypes = [C.c_size_t, C.c_int32, DP]
        FP = C.POINTER(C.c_float)
        L.ep_mlp_init.argtypes = [IP, C.c_int32, C.c_uint64]
        L.ep_mlp_num_params.argtypes = [C.POINTER(C.c_int64)]
        L.ep_mlp_get_params.argtypes = [FP]
        L.ep_mlp_set_params.argtypes = [FP]
        L.ep_mlp_forward.argtypes = [FP, C.c_int32, FP]
        L.ep_mlp_loss_grad.argtypes = [FP, FP, C.c_int32, DP, FP]
        L.ep_mlp_train.argtypes = [FP, FP, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_uint64, FP]
        L.ep_mlp_finalize.argtypes = []
        L.ep_mlp_bench.argtypes = [C.c_int32, C.c_int32, C.c_int32, DP]
        L.ep_surrogate_samples.argtypes = [C.c_int32, C.c_int32, C.c_uint64, FP, FP, IP, IP]
        L.ep_set_surrogate.argtypes = [C.c_int32, C.c_double]
        L.ep_get_surrogate_stats.argtypes = [C.POINTER(abi.EpSurrogateStats)]
        L.ep_surrogate_probe.argtypes = [C.c_int32, DP]
        L.ep_shell_operators.argtypes = [IP, C.c_double, C.c_int32]
        for name in abi.EXPORTED_SYMBOLS:
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def last_error():
    buf = C.create_string_buffer(512)
    lib().ep_last_error(buf, 512)
    return buf.value.decode()


def _check(rc, diag=None):
    if rc != EP_OK:
        raise EpError(rc, last_error(), diag)


NODE_FIELDS = ("x", "y", "z", "rmag", "bmag", "dlnb", "dlnn", "dupar")


class EpSolver:
    """Owns the library's device state for one problem size.

    Mirrors the reference's driver interface: `step()` is one call of
    updateEnergeticParticles() (src/energeticParticles.c:69-203).
    """

    def __init__(self, dims, params, stream_block=None, stream_halo=None):
        self.dims = dims if isinstance(dims, EpDims) else EpDims(*dims)
        self.params = params
        _check(lib().ep_init(C.byref(self.dims), C.byref(params)))
        if stream_block or stream_halo:
            self.set_tuning(stream_block or 256, stream_halo or 32)
        self._keep = None

    @property
    def shape_f(self):
        d = self.dims
        return (d.num_streams, d.num_nodes, d.num_energies, d.num_mu)

    def set_tuning(self, stream_block, stream_halo, lane_nodes=1, fuse_focus=1):
        """Streaming tiles of stream_block kept nodes (halo included),
        stream_halo subcycles per pass, lane_nodes (1, 2 or 3) nodes per lane;
        fuse_focus: focusing done by the last streaming pass (1) or by the
        separate kernel (0)."""
        _check(lib().ep_set_tuning(int(stream_block), int(stream_halo), int(lane_nodes), int(fuse_focus)))

    def set_schedule(self, schedule):
        """Streaming launches: 1 persistent per substep, 0 one per pass,
        2 one per stream (the unbatched form)."""
        _check(lib().ep_set_schedule(int(schedule)))

    def schedule(self):
        v = C.c_int32()
        _check(lib().ep_get_schedule(C.byref(v)))
        return v.value

    def autotune(self):
        """Tile-size search for the streaming kernel; returns (block, halo,
        lane_nodes, fuse_focus, ms)."""
        b, h, r, fz, ms = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32(), C.c_double()
        _check(lib().ep_autotune(C.byref(b), C.byref(h), C.byref(r), C.byref(fz), C.byref(ms)))
        return b.value, h.value, r.value, fz.value, ms.value

    def set_precision(self, mode):
        """0 / False: FP64 in the reference's operation order; 1 / True: FP32
        storage, streaming, focusing and interpolation with FP64 relaxation
        and adiabatic change; 2: FP64 with fused multiply-adds in streaming;
        3: mode 1 with ln f in the float storage (streams with the slab
        launch, two nodes per lane, focusing in its own kernel; setting it
        moves the tiling there)."""
        _check(lib().ep_set_precision(int(mode)))

    def set_graph(self, enable):
        _check(lib().ep_set_graph(1 if enable else 0))

    def set_node_fusion(self, fused):
        """1: focusing and adiabatic change of each node in one pass over f
        whenever every node needs at most one change subcycle (default);
        0: always the two separate kernels.  Results are bit-identical."""
        _check(lib().ep_set_node_fusion(1 if fused else 0))

    def set_params(self, params):
        self.params = params
        _check(lib().ep_set_params(C.byref(params)))

    def grids(self):
        E, M = self.dims.num_energies, self.dims.num_mu
        out = {"egrid": np.zeros(E), "vgrid": np.zeros(E), "mugrid": np.zeros(M), "rigidity": np.zeros(E)}
        dl, dm = C.c_double(), C.c_double()
        _check(lib().ep_get_grids(dptr(out["egrid"]), dptr(out["vgrid"]), dptr(out["mugrid"]),
                                  dptr(out["rigidity"]), C.byref(dl), C.byref(dm)))
        out["dlnp"], out["dmu"] = dl.value, dm.value
        return out

    def set_nodes(self, nodes, n_active=None):
        S, N = self.dims.num_streams, self.dims.num_nodes
        arrs = {k: np.ascontiguousarray(nodes[k], dtype=np.float64).reshape(S, N) for k in NODE_FIELDS}
        soa = EpNodeSoa(*[dptr(arrs[k]) for k in NODE_FIELDS])
        na = None if n_active is None else np.ascontiguousarray(n_active, dtype=np.int32)
        _check(lib().ep_set_nodes(C.byref(soa), iptr(na)))

    def upload_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == self.shape_f, (f.shape, self.shape_f)
        _check(lib().ep_upload_f(dptr(f), EP_LAYOUT_REFERENCE))

    def download_f(self):
        f = np.empty(self.shape_f, dtype=np.float64)
        _check(lib().ep_download_f(dptr(f), EP_LAYOUT_REFERENCE))
        return f

    def upload_f_streams(self, s0, f):
        """f [count][N][E][M] for streams s0 .. s0+count-1."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape[1:] == self.shape_f[1:], (f.shape, self.shape_f)
        _check(lib().ep_upload_f_streams(int(s0), int(f.shape[0]), dptr(f), EP_LAYOUT_REFERENCE))

    def download_f_streams(self, s0, count):
        f = np.empty((count,) + tuple(self.shape_f[1:]), dtype=np.float64)
        _check(lib().ep_download_f_streams(int(s0), int(count), dptr(f), EP_LAYOUT_REFERENCE))
        return f

    def init_seed_population(self):
        _check(lib().ep_init_seed_population())

    def step(self, op_flags=EP_OP_ALL, raise_on_error=True):
        dg = EpDiag()
        rc = lib().ep_step(op_flags, C.byref(dg))
        if raise_on_error:
            _check(rc, dg.as_dict())
        return rc, dg.as_dict()

    def node_outputs(self):
        S, N = self.dims.num_streams, self.dims.num_nodes
        ds, bp, bm = (np.zeros((S, N)) for _ in range(3))
        _check(lib().ep_get_node_outputs(dptr(ds), dptr(bp), dptr(bm)))
        return ds, bp, bm

    def mas_set_level(self, level, fields, meshes):
        fields = [np.ascontiguousarray(a, dtype=np.float32) for a in fields]
        meshes = [np.ascontiguousarray(a, dtype=np.float32) for a in meshes]
        dims = np.array([len(m) for m in meshes], dtype=np.int32)
        FP = C.POINTER(C.c_float)
        fa = (FP * 7)(*[fptr(a) for a in fields])
        ma = (FP * 21)(*[fptr(a) for a in meshes])
        _check(lib().ep_mas_set_level(int(level), fa, ma, iptr(dims)))

    def interp_mhd(self, s_cor, phi_offset, mas_radial_max, mas_vmin, prev=None):
        S, N = self.dims.num_streams, self.dims.num_nodes
        mp = EpMasParams(s_cor, phi_offset, mas_radial_max, mas_vmin)
        keep = []
        pv = None
        if prev is not None:
            names = ("rold_x", "rold_y", "rold_z", "br", "btheta", "bphi", "vr", "vtheta", "vphi", "density")
            arrs = [np.ascontiguousarray(prev[k], dtype=np.float64).reshape(S, N) for k in names]
            keep.extend(arrs)
            pv = EpMasPrev(*[dptr(a) for a in arrs])
        outs = {k: np.zeros((S, N)) for k in ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density")}
        mo = EpMasOut(*[dptr(outs[k]) for k in ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density")])
        ms = C.c_double()
        _check(lib().ep_interp_mhd(C.byref(mp), C.byref(pv) if pv is not None else None, C.byref(mo),
                                   C.byref(ms)))
        outs["device_ms"] = ms.value
        return outs

    def set_positions(self, nodes, n_active=None):
        S, N = self.dims.num_streams, self.dims.num_nodes
        arrs = [np.ascontiguousarray(nodes[k], dtype=np.float64).reshape(S, N) for k in ("x", "y", "z", "rmag")]
        na = None if n_active is None else np.ascontiguousarray(n_active, dtype=np.int32)
        _check(lib().ep_set_positions(*[dptr(a) for a in arrs], iptr(na)))

    def spawn_shell(self, x, y, z, rmag, f=None):
        """Ripple every node one shell outward and make node 0 the new shell
        (positions [S]; f [S, E, M] or None for the seed spectrum)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (x, y, z, rmag)]
        fa = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
        sp = abi.EpSpawn(*[dptr(a) for a in arrs], dptr(fa))
        _check(lib().ep_spawn_shell(C.byref(sp)))

    def move_nodes(self, tdel):
        _check(lib().ep_move_nodes(float(tdel)))

    def move_nodes_mas(self, tdel, s_cor, phi_offset, mas_radial_max, mas_vmin=0.0):
        """masMoveNodes on the device: RK4 through the MAS levels set with
        set_mas_level, masWind of the device-held node state outside."""
        mp = EpMasParams(s_cor, phi_offset, mas_radial_max, mas_vmin)
        _check(lib().ep_move_nodes_mas(C.byref(mp), float(tdel)))

    def positions(self):
        S, N = self.dims.num_streams, self.dims.num_nodes
        out = {k: np.zeros((S, N)) for k in ("x", "y", "z", "rmag")}
        na = np.zeros(S, dtype=np.int32)
        _check(lib().ep_get_positions(*[dptr(out[k]) for k in ("x", "y", "z", "rmag")], iptr(na)))
        out["n_active"] = na
        return out

    def update_mhd(self, prev, mode=0, mhdUs=0.0, mhdNsAu=0.0, mhdBsAu=0.0, omegaSun=0.0, s_cor=0.0,
                   phi_offset=0.0, mas_radial_max=0.0, mas_vmin=0.0):
        """updateMhd on the device; returns host copies of the new B, V, n.
        prev=None uses the MHD state the device holds from the last call."""
        S, N = self.dims.num_streams, self.dims.num_nodes
        names = ("rold_x", "rold_y", "rold_z", "br", "btheta", "bphi", "vr", "vtheta", "vphi", "density")
        pv = None
        if prev is not None:
            arrs = [np.ascontiguousarray(prev[k], dtype=np.float64).reshape(S, N) for k in names]
            pv = EpMasPrev(*[dptr(a) for a in arrs])
        outs = {k: np.zeros((S, N)) for k in ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density")}
        mo = EpMasOut(*[dptr(outs[k]) for k in ("br", "btheta", "bphi", "vr", "vtheta", "vphi", "density")])
        mp = abi.EpMhdParams(mode, mhdUs, mhdNsAu, mhdBsAu, omegaSun,
                             EpMasParams(s_cor, phi_offset, mas_radial_max, mas_vmin))
        ms = C.c_double()
        _check(lib().ep_update_mhd(C.byref(mp), C.byref(pv) if pv is not None else None, C.byref(mo),
                                   C.byref(ms)))
        outs["device_ms"] = ms.value
        return outs

    def observe(self, r_au, theta, phi, idw_p=3.0, noalloc=False):
        """Distribution [O, E, M] and omnidirectional flux [O, E] at point
        observers (code units); noalloc: through ep_observe_noalloc (fixed
        scratch, no allocation)."""
        r, t, ph = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (r_au, theta, phi))
        o = len(r)
        ob = abi.EpObservers(o, dptr(r), dptr(t), dptr(ph), float(idw_p))
        E, M = self.dims.num_energies, self.dims.num_mu
        dist, flux = np.zeros((o, E, M)), np.zeros((o, E))
        fn = lib().ep_observe_noalloc if noalloc else lib().ep_observe
        _check(fn(C.byref(ob), dptr(dist), dptr(flux)))
        return dist, flux

    def stream_flux(self):
        d = self.dims
        flux = np.zeros((d.num_streams, d.num_nodes, d.num_energies))
        _check(lib().ep_stream_flux(dptr(flux)))
        return flux

    def shell_operators(self, nbr, kperxkpar=0.0, drift=False):
        """Perpendicular diffusion (kperxkpar > 0) and drift between
        neighbouring streams (nbr [S, 4] from eprem_amd.cube.neighbours)."""
        nb = None if nbr is None else np.ascontiguousarray(nbr, dtype=np.int32)
        _check(lib().ep_shell_operators(iptr(nb), float(kperxkpar), 1 if drift else 0))

    def surrogate_samples(self, K, n, seed=1, where=False):
        """Training pairs (x [k, 3M+6], y [k, M]) of K streaming subcycles and
        the focusing of the same time on the current state; only the
        qualifying rows of n candidates (with where: also their (stream,
        energy, node) [k, 3])."""
        M = self.dims.num_mu
        x = np.zeros((n, 3 * M + 6), np.float32)
        y = np.zeros((n, M), np.float32)
        v = np.zeros(n, np.int32)
        w = np.zeros((n, 3), np.int32)
        _check(lib().ep_surrogate_samples(int(K), int(n), int(seed), fptr(x), fptr(y), iptr(v), iptr(w)))
        ok = v != 0
        if where:
            return x[ok], y[ok], w[ok]
        return x[ok], y[ok]

    def set_surrogate(self, K, tol=1e-3):
        """Learned-surrogate form of step(): the network of the current Mlp
        replaces K streaming + focusing subcycles of every slab with at least
        K, with a per-stream plain fallback above the residual tol; K = 0
        turns it off."""
        _check(lib().ep_set_surrogate(int(K), float(tol)))

    def surrogate_stats(self):
        st = abi.EpSurrogateStats()
        _check(lib().ep_get_surrogate_stats(C.byref(st)))
        return {n: getattr(st, n) for n, _ in abi.EpSurrogateStats._fields_ if n != "pad_"}

    def surrogate_probe(self, mode):
        """mode 0: one network application made current; 1: the plain K
        subcycles into the spare buffer.  Returns the largest residual."""
        r = C.c_double()
        _check(lib().ep_surrogate_probe(int(mode), C.byref(r)))
        return r.value

    def close(self):
        lib().ep_finalize()


def selftest_dpp():
    x = np.arange(64, dtype=np.float64) + 0.5
    a, b = np.zeros(64), np.zeros(64)
    _check(lib().ep_selftest_dpp(dptr(x), dptr(a), dptr(b)))
    return x, a, b


def fastmath(x, device=False):
    """(log(x), exp(x)) through the table-driven functions of the fused
    precision modes, on the host or on the device (device needs an
    initialised EpSolver)."""
    xs = np.ascontiguousarray(x, dtype=np.float64).ravel()
    lg, ex = np.empty_like(xs), np.empty_like(xs)
    L = lib()
    fn = L.ep_selftest_fastmath if device else L.ep_fastmath_host
    DPt = C.POINTER(C.c_double)
    fn.argtypes = [DPt, C.c_int64, DPt, DPt]
    _check(fn(dptr(xs), len(xs), dptr(lg), dptr(ex)))
    return lg, ex


def copy_probe(nbytes=1 << 30, reps=20):
    """Mean time (ms) and bandwidth (GB/s, read + write) of a plain device
    copy of nbytes bytes."""
    ms = C.c_double()
    _check(lib().ep_selftest_copy(int(nbytes), int(reps), C.byref(ms)))
    return ms.value, 2.0 * nbytes / (ms.value * 1e-3) / 1e9


if __name__ == "__main__":
    import sys
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 1 << 30
    t, bw = copy_probe(n)
    print("copy %d bytes: %.4f ms, %.1f GB/s (read + write)" % (n, t, bw))


class Mlp:
    """The device multilayer perceptron (matrix-core products, bf16 operands,
    fp32 accumulation and master weights)."""

    def __init__(self, sizes, seed=1):
        self.sizes = list(sizes)
        sz = np.ascontiguousarray(self.sizes, dtype=np.int32)
        _check(lib().ep_mlp_init(iptr(sz), len(sz), int(seed)))
        n = C.c_int64()
        _check(lib().ep_mlp_num_params(C.byref(n)))
        self.nparam = n.value

    def params(self):
        p = np.zeros(self.nparam, np.float32)
        _check(lib().ep_mlp_get_params(fptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        _check(lib().ep_mlp_set_params(fptr(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.sizes[-1]), np.float32)
        _check(lib().ep_mlp_forward(fptr(x), x.shape[0], fptr(y)))
        return y

    def loss_grad(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(self.nparam, np.float32)
        loss = C.c_double()
        _check(lib().ep_mlp_loss_grad(fptr(x), fptr(y), x.shape[0], C.byref(loss), fptr(g)))
        return loss.value, g

    def train(self, x, y, epochs, batch, lr, seed=7):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        hist = np.zeros(epochs, np.float32)
        _check(lib().ep_mlp_train(fptr(x), fptr(y), x.shape[0], int(epochs), int(batch), float(lr), int(seed),
                                  fptr(hist)))
        return hist

    def bench(self, n, reps=5, layerwise=False):
        """Mean ms of one forward pass over n device-resident rows (the fused
        kernel, or with layerwise=True one product kernel per layer)."""
        ms = C.c_double()
        _check(lib().ep_mlp_bench(int(n), int(reps), 1 if layerwise else 0, C.byref(ms)))
        return ms.value

    def close(self):
        lib().ep_mlp_finalize()
