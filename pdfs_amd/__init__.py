"""Python host-side mirror of the MI355X-native FSM C-ABI (include/pdfs.h).

Loads pdfs_amd/lib/libpdfs.so (built by pdfs_amd/Makefile; run
`__graft_entry__.build()`); there is no fallback: if the library is missing
every entry point raises. Device arrays are torch tensors (HIP memory);
their data_ptr() is handed to the C-ABI.

Reference interface mirrored:
  run_fsm(Phi*, int)  single_gpu_cuda/src/phi3D_fsm.h:69-71
  runFSM(Phi*, int)   mpi_openacc/src/phi3D_fsm.h:73
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpdfs.so")

PDFS_OK, PDFS_EINVAL, PDFS_EHIP, PDFS_ENOMEM, PDFS_ETIMEOUT, PDFS_ENOTCONV = 0, -1, -2, -3, -4, -5
PDFS_F64_REF, PDFS_F32 = 0, 1
PDFS_STORE_F32, PDFS_STORE_FIXED = 0, 1
PDFS_OVERLAP_OFF, PDFS_OVERLAP_AUTO, PDFS_OVERLAP_ALWAYS = 0, 1, 2
PDFS_OV_AUTO_FIELDS, PDFS_OV_AUTO_DEPTH = 2, 6
GEN_POINT, GEN_TERRAIN, GEN_SPHERES = 0, 1, 2
TILE_16x16, TILE_32x8, TILE_16x8, TILE_AUTO = 0, 1, 2, -2
TILE_32x16, TILE_16x16_1CU, TILE_16x8_2CU, TILE_8x8, TILE_16x4 = 3, 4, 5, 6, 7
TILES = (TILE_16x16, TILE_32x8, TILE_16x8, TILE_32x16, TILE_16x16_1CU, TILE_16x8_2CU, TILE_8x8,
         TILE_16x4)

EXPORTED = [
    "run_fsm", "runFSM", "pdfs_default_opts", "pdfs_solve", "pdfs_ctx_create",
    "pdfs_ctx_destroy", "pdfs_last_error", "pdfs_solve_f32_dev", "pdfs_sweep_f32_dev",
    "pdfs_solve_ref64_dev", "pdfs_solve_speed64_dev", "pdfs_judge_f32_dev",
    "pdfs_judge_fixed_dev",
    "pdfs_selftest_sqrt_f32", "pdfs_generate_dev",
    "pdfs_generate_ex_dev",
    "pdfs_pack_dev",
    "pdfs_unpack_dev", "pdfs_solve_batch_f32_dev", "pdfs_calc_dist_field", "pdfs_prepost", "pdfs_vti_read",
    "pdfs_vti_write", "pdfs_vti_free", "pdfs_ctx_stream", "pdfs_net_create", "pdfs_net_destroy",
    "pdfs_net_load", "pdfs_net_save", "pdfs_net_num_params", "pdfs_net_get", "pdfs_net_set",
    "pdfs_net_get_calibration", "pdfs_net_set_calibration", "pdfs_net_get_cut", "pdfs_net_set_cut",
    "pdfs_net_grid",
    "pdfs_net_features_dev", "pdfs_net_target_dev", "pdfs_net_forward_dev",
    "pdfs_net_train_step_dev", "pdfs_warm_start_f32_dev", "pdfs_warm_start_batch_f32_dev",
    "pdfs_solve_warm_f32_dev",
    "pdfs_solve_batch_warm_f32_dev", "pdfs_sweep_f32_noalloc", "pdfs_solve_speed_f32_dev",
    "pdfs_solve_batch_speed_f32_dev", "pdfs_judge_speed_f32_dev",
]
WEIGHTS_PATH = os.path.join(HERE, "weights", "warm_unet.bin")
CLI_PATH = os.path.join(HERE, "bin", "PDFS")
NC_LIB_PATH = os.path.join(HERE, "lib", "libpdfs_nc.so")
NC_EXPORTED = ["pdfs_nc_read", "pdfs_nc_write", "pdfs_nc_free"]


class NcField(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int), ("dx", C.c_double),
                ("dy", C.c_double), ("dz", C.c_double), ("flags", C.POINTER(C.c_int)),
                ("distance", C.POINTER(C.c_double))]


_NC = None


def nc_lib():
    """lib/libpdfs_nc.so: NetCDF-4 field files over HDF5 (host only)."""
    global _NC
    if _NC is None:
        if not os.path.exists(NC_LIB_PATH):
            raise PdfsError(PDFS_EINVAL, f"{NC_LIB_PATH} is missing (built where HDF5 is present)")
        L = C.CDLL(NC_LIB_PATH)
        L.pdfs_nc_read.argtypes = [C.c_char_p, C.POINTER(NcField)]
        L.pdfs_nc_read.restype = C.c_int
        L.pdfs_nc_free.argtypes = [C.POINTER(NcField)]
        L.pdfs_nc_free.restype = None
        L.pdfs_nc_write.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                    C.c_double, C.c_double, C.c_double]
        L.pdfs_nc_write.restype = C.c_int
        _NC = L
    return _NC


def nc_write(path, flags, distance, h=(1.0, 1.0, 1.0)):
    """Write Flags (int32) and Distance (float64) [z][y][x] as a NetCDF-4
    file of the MPI CLI's layout; returns the library's code (0 = ok)."""
    import numpy as np
    fl = np.ascontiguousarray(flags, dtype=np.int32)
    di = np.ascontiguousarray(distance, dtype=np.float64)
    assert fl.shape == di.shape and fl.ndim == 3
    Z, Y, X = fl.shape
    return nc_lib().pdfs_nc_write(str(path).encode(), fl.ctypes.data, di.ctypes.data, X, Y, Z, *h)


def nc_read(path):
    """(code, dict(flags, distance [z][y][x], h=(dx, dy, dz)) or None)."""
    import numpy as np
    f = NcField()
    rc = nc_lib().pdfs_nc_read(str(path).encode(), C.byref(f))
    if rc:
        return rc, None
    n = f.x * f.y * f.z
    out = {"flags": np.ctypeslib.as_array(f.flags, (n,)).copy().reshape(f.z, f.y, f.x),
           "distance": np.ctypeslib.as_array(f.distance, (n,)).copy().reshape(f.z, f.y, f.x),
           "h": (f.dx, f.dy, f.dz)}
    nc_lib().pdfs_nc_free(C.byref(f))
    return 0, out


class Phi(C.Structure):
    """single_gpu_cuda/src/phi3D.h:45-55 layout (64 bytes)."""
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("location", C.POINTER(C.c_int)), ("distance", C.POINTER(C.c_double)),
                ("F", C.c_double)]


class PhiACC(C.Structure):
    """mpi_openacc/src/phi3D.h:45-54 layout (56 bytes)."""
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("distance", C.POINTER(C.c_double)), ("F", C.c_double)]


class Opts(C.Structure):
    _fields_ = [("precision", C.c_int), ("max_passes", C.c_int), ("fixed_passes", C.c_int),
                ("tol", C.c_double), ("judge", C.c_int), ("tile", C.c_int),
                ("quiet", C.c_int), ("profile", C.c_int), ("graph", C.c_int),
                ("certify", C.c_int), ("certify_from", C.c_int), ("polish", C.c_int),
                ("polish_tol", C.c_double), ("tile64", C.c_int), ("storage", C.c_int),
                ("judge_tol", C.c_double), ("stop_quanta", C.c_int), ("overlap", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("passes", C.c_int), ("orderings", C.c_int), ("point_updates", C.c_longlong),
                ("last_changed", C.c_longlong), ("last_maxdiff", C.c_double),
                ("judge_flagged", C.c_longlong), ("ms_sweeps", C.c_double),
                ("ms_judge", C.c_double), ("ms_total", C.c_double),
                ("ms_sweep_kernels", C.c_double), ("tile", C.c_int), ("judge_runs", C.c_int),
                ("fallbacks", C.c_int), ("ms_warm", C.c_double), ("polish_passes", C.c_int),
                ("polish_orderings", C.c_int), ("polish_maxdiff", C.c_double),
                ("ms_polish", C.c_double), ("polish_skipped", C.c_int),
                ("polish_point_updates", C.c_longlong), ("ms_polish_kernels", C.c_double),
                ("storage", C.c_int), ("sweep_launches", C.c_int), ("ms_convert", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PdfsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pdfs error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """The loaded C-ABI library. Raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: the HIP extension must be built "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        vp, I, D, LL = C.c_void_p, C.c_int, C.c_double, C.c_longlong
        L.run_fsm.argtypes = [C.POINTER(Phi), I]
        L.run_fsm.restype = None
        L.runFSM.argtypes = [C.POINTER(PhiACC), I]
        L.runFSM.restype = None
        L.pdfs_default_opts.argtypes = [C.POINTER(Opts)]
        L.pdfs_solve.argtypes = [vp, I, I, I, D, D, D, C.POINTER(Opts), C.POINTER(Stats)]
        L.pdfs_solve.restype = I
        L.pdfs_ctx_create.argtypes = []
        L.pdfs_ctx_create.restype = vp
        L.pdfs_ctx_destroy.argtypes = [vp]
        L.pdfsi_set_overlap64.argtypes = [vp, I]
        L.pdfsi_set_overlap64.restype = I
        L.pdfs_last_error.restype = C.c_char_p
        L.pdfs_solve_f32_dev.argtypes = [vp, vp, vp, I, I, I, D, D, D, C.POINTER(Opts),
                                         C.POINTER(Stats)]
        L.pdfs_solve_f32_dev.restype = I
        L.pdfs_sweep_f32_dev.argtypes = [vp, vp, I, I, I, D, D, D, I, I, C.POINTER(LL)]
        L.pdfs_sweep_f32_dev.restype = I
        L.pdfs_solve_ref64_dev.argtypes = [vp, vp, I, I, I, D, D, D, I, D, C.POINTER(Stats)]
        L.pdfs_solve_ref64_dev.restype = I
        L.pdfs_solve_speed64_dev.argtypes = [vp, vp, vp, I, I, I, D, D, D, I, D, C.POINTER(Stats)]
        L.pdfs_solve_speed64_dev.restype = I
        L.pdfs_judge_f32_dev.argtypes = [vp, vp, vp, I, I, I, D, D, D, vp, C.POINTER(LL)]
        L.pdfs_judge_f32_dev.restype = I
        L.pdfs_judge_fixed_dev.argtypes = [vp, vp, vp, I, I, I, D, D, D, D, vp, C.POINTER(LL)]
        L.pdfs_judge_fixed_dev.restype = I
        L.pdfs_selftest_sqrt_f32.argtypes = [vp, C.POINTER(C.c_ulonglong)]
        L.pdfs_selftest_sqrt_f32.restype = I
        L.pdfs_generate_dev.argtypes = [vp, I, C.c_uint64, I, I, I, D, vp, vp]
        L.pdfs_generate_dev.restype = I
        L.pdfs_generate_ex_dev.argtypes = [vp, I, C.c_uint64, I, I, I, D, I, D, D, D, D, vp, vp]
        L.pdfs_generate_ex_dev.restype = I
        L.pdfs_pack_dev.argtypes = [vp, vp, vp, I, I, I]
        L.pdfs_pack_dev.restype = I
        L.pdfs_unpack_dev.argtypes = [vp, vp, vp, I, I, I]
        L.pdfs_unpack_dev.restype = I
        L.pdfs_solve_batch_f32_dev.argtypes = [vp, I, vp, vp, vp, D, D, D, C.POINTER(Opts),
                                               C.POINTER(Stats), vp, vp]
        L.pdfs_solve_batch_f32_dev.restype = I
        L.pdfs_solve_speed_f32_dev.argtypes = [vp, vp, vp, vp, I, I, I, D, D, D,
                                               C.POINTER(Opts), C.POINTER(Stats)]
        L.pdfs_solve_speed_f32_dev.restype = I
        L.pdfs_solve_batch_speed_f32_dev.argtypes = [vp, I, vp, vp, vp, vp, D, D, D,
                                                     C.POINTER(Opts), C.POINTER(Stats), vp, vp]
        L.pdfs_solve_batch_speed_f32_dev.restype = I
        L.pdfs_judge_speed_f32_dev.argtypes = [vp, vp, vp, vp, I, I, I, D, D, D, vp,
                                               C.POINTER(LL)]
        L.pdfs_judge_speed_f32_dev.restype = I
        L.pdfs_calc_dist_field.argtypes = [C.POINTER(Phi), I, I]
        L.pdfs_calc_dist_field.restype = I
        L.pdfs_prepost.argtypes = [I, vp, vp, I, I, I]
        L.pdfs_prepost.restype = I
        F = C.c_float
        L.pdfs_ctx_stream.argtypes = [vp]
        L.pdfs_ctx_stream.restype = vp
        L.pdfs_net_create.argtypes = [vp, C.c_uint64]
        L.pdfs_net_create.restype = vp
        L.pdfs_net_destroy.argtypes = [vp]
        L.pdfs_net_destroy.restype = None
        L.pdfs_net_load.argtypes = [vp, C.c_char_p]
        L.pdfs_net_load.restype = I
        L.pdfs_net_save.argtypes = [vp, C.c_char_p]
        L.pdfs_net_save.restype = I
        L.pdfs_net_num_params.argtypes = [vp]
        L.pdfs_net_num_params.restype = LL
        L.pdfs_net_get.argtypes = [vp, I, vp]
        L.pdfs_net_get.restype = I
        L.pdfs_net_set.argtypes = [vp, vp]
        L.pdfs_net_set.restype = I
        L.pdfs_net_get_calibration.argtypes = [vp, C.POINTER(F), C.POINTER(F), C.POINTER(F)]
        L.pdfs_net_get_calibration.restype = None
        L.pdfs_net_set_calibration.argtypes = [vp, F, F, F]
        L.pdfs_net_set_calibration.restype = None
        L.pdfs_net_get_cut.argtypes = [vp]
        L.pdfs_net_get_cut.restype = F
        L.pdfs_net_set_cut.argtypes = [vp, F]
        L.pdfs_net_set_cut.restype = None
        L.pdfs_net_grid.argtypes = [I, I, I, C.POINTER(I), C.POINTER(I), C.POINTER(I)]
        L.pdfs_net_grid.restype = None
        L.pdfs_net_features_dev.argtypes = [vp, vp, I, I, I, vp]
        L.pdfs_net_features_dev.restype = I
        L.pdfs_net_target_dev.argtypes = [vp, vp, I, I, I, vp]
        L.pdfs_net_target_dev.restype = I
        L.pdfs_net_forward_dev.argtypes = [vp, vp, I, I, I, I, vp]
        L.pdfs_net_forward_dev.restype = I
        L.pdfs_net_train_step_dev.argtypes = [vp, vp, vp, I, I, I, I, F, C.POINTER(F)]
        L.pdfs_net_train_step_dev.restype = I
        L.pdfs_warm_start_f32_dev.argtypes = [vp, vp, vp, I, I, I, D]
        L.pdfs_warm_start_f32_dev.restype = I
        L.pdfs_warm_start_batch_f32_dev.argtypes = [vp, I, vp, vp, vp, D]
        L.pdfs_warm_start_batch_f32_dev.restype = I
        L.pdfs_solve_warm_f32_dev.argtypes = [vp, vp, vp, vp, I, I, I, D, D, D, C.POINTER(Opts),
                                              C.POINTER(Stats)]
        L.pdfs_solve_warm_f32_dev.restype = I
        L.pdfs_solve_batch_warm_f32_dev.argtypes = [vp, vp, I, vp, vp, vp, D, D, D,
                                                    C.POINTER(Opts), C.POINTER(Stats), vp, vp]
        L.pdfs_solve_batch_warm_f32_dev.restype = I
        L.pdfs_sweep_f32_noalloc.argtypes = [vp, I, I, I, D, D, D, I, I, vp, vp]
        L.pdfs_sweep_f32_noalloc.restype = I
        _lib = L
    return _lib


def _sync_inputs():
    """Wait for the work torch queued on its current stream (the inputs of a
    library call). Only that stream: other threads' contexts keep running."""
    import torch
    torch.cuda.current_stream().synchronize()


def _check(rc, ok=(PDFS_OK,)):
    if rc not in ok:
        raise PdfsError(rc, lib().pdfs_last_error().decode(errors="replace"))
    return rc


def default_opts(**kw):
    o = Opts()
    lib().pdfs_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _dims(a):
    z, y, x = a.shape
    return x - 2, y - 2, z - 2


def run_fsm(distance, iterations=1, h=(1.0, 1.0, 1.0)):
    """Reference solver entry on a padded FP64 numpy array (in place)."""
    assert distance.dtype == np.float64 and distance.flags.c_contiguous
    X, Y, Z = _dims(distance)
    pf = Phi(X, Y, Z, h[0], h[1], h[2], None,
             distance.ctypes.data_as(C.POINTER(C.c_double)), 1.0)
    lib().run_fsm(C.byref(pf), iterations)
    return distance


def runFSM(distance, sn, h=(1.0, 1.0, 1.0)):
    """One reference ordering sn (1..8), OpenACC-build entry (in place)."""
    assert distance.dtype == np.float64 and distance.flags.c_contiguous
    X, Y, Z = _dims(distance)
    pf = PhiACC(X, Y, Z, h[0], h[1], h[2], distance.ctypes.data_as(C.POINTER(C.c_double)), 1.0)
    lib().runFSM(C.byref(pf), sn)
    return distance


def calc_dist_field(location, distance, h=(1.0, 1.0, 1.0), passes=1, precision=PDFS_F64_REF):
    """single_gpu_cuda/src/phi3D.c:107-126 on padded host arrays (in place):
    update_distance, `passes` passes (<= 0: until converged), location 1 ->
    -1, ghosts <- nearest interior."""
    assert distance.dtype == np.float64 and location.dtype == np.int32
    X, Y, Z = _dims(distance)
    pf = Phi(X, Y, Z, h[0], h[1], h[2], location.ctypes.data_as(C.POINTER(C.c_int)),
             distance.ctypes.data_as(C.POINTER(C.c_double)), 1.0)
    _check(lib().pdfs_calc_dist_field(C.byref(pf), passes, precision))
    return distance


def prepost(op, location, distance):
    """One CLI pre/post step on padded host arrays (in place): 0
    update_distance, 1 set_distance_negative_inside, 2 adjust_boundary."""
    X, Y, Z = _dims(distance)
    _check(lib().pdfs_prepost(op, None if location is None else location.ctypes.data,
                              distance.ctypes.data, X, Y, Z))
    return distance


def solve(distance, h=(1.0, 1.0, 1.0), precision=PDFS_F32, allow_notconv=False, **kw):
    """pdfs_solve on a host padded FP64 array (in place). Returns stats dict."""
    assert distance.dtype == np.float64 and distance.flags.c_contiguous
    X, Y, Z = _dims(distance)
    o = default_opts(precision=precision, **kw)
    st = Stats()
    rc = lib().pdfs_solve(distance.ctypes.data, X, Y, Z, h[0], h[1], h[2], C.byref(o),
                          C.byref(st))
    _check(rc, (PDFS_OK, PDFS_ENOTCONV) if allow_notconv else (PDFS_OK,))
    return st.as_dict()


class Context:
    """Owns a pdfs_ctx (device scratch, stream). Device arrays are torch tensors."""

    def __init__(self):
        import weakref
        self.c = lib().pdfs_ctx_create()
        if not self.c:
            raise PdfsError(PDFS_EHIP, lib().pdfs_last_error().decode())
        self._nets = weakref.WeakSet()

    def close(self):
        if self.c:
            for net in list(self._nets):
                net.close()
            lib().pdfs_ctx_destroy(self.c)
            self.c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _ptr(t):
        if t is None:
            return None
        assert t.is_cuda and t.is_contiguous()
        return t.data_ptr()

    def judge_fixed(self, q, init, h, base, d, tol, flags=None):
        """The FP64 judge of a fixed-point field: q an int32/uint32 cuda tensor
        [Z, Y, X] of quanta (value = base + q d, all ones = +inf), init float32
        of the same shape. Returns the number of failing points."""
        import torch
        assert q.dtype in (torch.int32, torch.uint32) and q.dim() == 3
        assert init.dtype == torch.float32 and init.shape == q.shape
        Z, Y, X = q.shape
        n = C.c_longlong()
        _sync_inputs()
        _check(lib().pdfs_judge_fixed_dev(self.c, self._ptr(q), self._ptr(init), X, Y, Z, h,
                                          base, d, tol, self._ptr(flags), C.byref(n)))
        return n.value

    def solve_f32(self, field, init=None, h=(1.0, 1.0, 1.0), allow_notconv=False, **kw):
        """field: dense float32 cuda tensor [Z, Y, X], solved in place."""
        import torch
        assert field.dtype == torch.float32 and field.dim() == 3
        Z, Y, X = field.shape
        o = default_opts(precision=PDFS_F32, **kw)
        st = Stats()
        _sync_inputs()
        rc = lib().pdfs_solve_f32_dev(self.c, self._ptr(field), self._ptr(init), X, Y, Z, *h,
                                      C.byref(o), C.byref(st))
        _check(rc, (PDFS_OK, PDFS_ENOTCONV) if allow_notconv else (PDFS_OK,))
        return st.as_dict()

    def solve_batch_f32(self, fields, inits=None, h=(1.0, 1.0, 1.0), allow_notconv=False,
                        **kw):
        """fields: list of dense float32 cuda tensors [Z, Y, X] (uneven
        sizes), solved in place in shared launches. Returns (stats,
        per-field passes, per-field completion ms)."""
        import torch
        n = len(fields)
        ptrs = (C.c_void_p * n)(*[self._ptr(f) for f in fields])
        iptrs = (C.c_void_p * n)(*[self._ptr(t) for t in inits]) if inits else None
        dims = (C.c_int * (3 * n))()
        for i, f in enumerate(fields):
            assert f.dtype == torch.float32 and f.dim() == 3
            Z, Y, X = f.shape
            dims[3 * i], dims[3 * i + 1], dims[3 * i + 2] = X, Y, Z
        passes = (C.c_int * n)()
        ms = (C.c_double * n)()
        o = default_opts(precision=PDFS_F32, **kw)
        st = Stats()
        _sync_inputs()
        rc = lib().pdfs_solve_batch_f32_dev(self.c, n, ptrs, iptrs, dims, *h, C.byref(o),
                                            C.byref(st), passes, ms)
        _check(rc, (PDFS_OK, PDFS_ENOTCONV) if allow_notconv else (PDFS_OK,))
        return st.as_dict(), list(passes), list(ms)

    def solve_speed_f32(self, field, slowness, init=None, h=(1.0, 1.0, 1.0),
                        allow_notconv=False, **kw):
        """|grad u| = slowness (1/F per point) on a dense float32 field
        [Z, Y, X], in place; slowness: float32 cuda tensor of the same
        shape. Options and statistics as solve_f32."""
        import torch
        assert field.dtype == torch.float32 and field.dim() == 3
        assert slowness.dtype == torch.float32 and slowness.shape == field.shape
        Z, Y, X = field.shape
        o = default_opts(precision=PDFS_F32, **kw)
        st = Stats()
        _sync_inputs()
        rc = lib().pdfs_solve_speed_f32_dev(self.c, self._ptr(field), self._ptr(init),
                                            self._ptr(slowness), X, Y, Z, *h, C.byref(o),
                                            C.byref(st))
        _check(rc, (PDFS_OK, PDFS_ENOTCONV) if allow_notconv else (PDFS_OK,))
        return st.as_dict()

    def solve_batch_speed_f32(self, fields, slownesses, inits=None, h=(1.0, 1.0, 1.0),
                              allow_notconv=False, **kw):
        """solve_batch_f32 of |grad u| = s with slownesses[i] for fields[i]."""
        import torch
        n = len(fields)
        assert len(slownesses) == n
        ptrs = (C.c_void_p * n)(*[self._ptr(f) for f in fields])
        iptrs = (C.c_void_p * n)(*[self._ptr(t) for t in inits]) if inits else None
        sptrs = (C.c_void_p * n)(*[self._ptr(t) for t in slownesses])
        dims = (C.c_int * (3 * n))()
        for i, f in enumerate(fields):
            assert f.dtype == torch.float32 and f.dim() == 3
            assert slownesses[i].dtype == torch.float32 and slownesses[i].shape == f.shape
            Z, Y, X = f.shape
            dims[3 * i], dims[3 * i + 1], dims[3 * i + 2] = X, Y, Z
        passes = (C.c_int * n)()
        ms = (C.c_double * n)()
        o = default_opts(precision=PDFS_F32, **kw)
        st = Stats()
        _sync_inputs()
        rc = lib().pdfs_solve_batch_speed_f32_dev(self.c, n, ptrs, iptrs, sptrs, dims, *h,
                                                  C.byref(o), C.byref(st), passes, ms)
        _check(rc, (PDFS_OK, PDFS_ENOTCONV) if allow_notconv else (PDFS_OK,))
        return st.as_dict(), list(passes), list(ms)

    def judge_speed_f32(self, field, init, slowness, h=(1.0, 1.0, 1.0), flags=None):
        Z, Y, X = field.shape
        n = C.c_longlong(0)
        _sync_inputs()
        _check(lib().pdfs_judge_speed_f32_dev(self.c, self._ptr(field), self._ptr(init),
                                              self._ptr(slowness), X, Y, Z, *h,
                                              self._ptr(flags), C.byref(n)))
        return n.value

    def sweep_f32(self, field, sn, h=(1.0, 1.0, 1.0), tile=-1):
        import torch
        Z, Y, X = field.shape
        ch = C.c_longlong(0)
        _sync_inputs()
        _check(lib().pdfs_sweep_f32_dev(self.c, self._ptr(field), X, Y, Z, *h, sn, tile,
                                        C.byref(ch)))
        return ch.value

    def solve_ref64(self, padded, passes=0, tol=1e-12, h=(1.0, 1.0, 1.0)):
        import torch
        assert padded.dtype == torch.float64 and padded.dim() == 3
        Z, Y, X = (s - 2 for s in padded.shape)
        st = Stats()
        _sync_inputs()
        _check(lib().pdfs_solve_ref64_dev(self.c, self._ptr(padded), X, Y, Z, *h, passes, tol,
                                          C.byref(st)))
        return st.as_dict()

    def solve_speed64(self, padded, slowness, passes=0, tol=1e-12, h=(1.0, 1.0, 1.0)):
        """Eikonal |grad u| = slowness (= 1/F per point) on a padded FP64
        field, reference arithmetic and schedule; both float64 cuda tensors
        of the padded shape."""
        import torch
        assert padded.dtype == torch.float64 and padded.dim() == 3
        assert slowness.dtype == torch.float64 and slowness.shape == padded.shape
        Z, Y, X = (s - 2 for s in padded.shape)
        st = Stats()
        _sync_inputs()
        _check(lib().pdfs_solve_speed64_dev(self.c, self._ptr(padded), self._ptr(slowness), X, Y,
                                            Z, *h, passes, tol, C.byref(st)))
        return st.as_dict()

    def selftest_sqrt_f32(self):
        """Floats in [2^-96, inf) where the sweep's FP32 square root differs
        from the correctly rounded one (0 expected)."""
        n = C.c_ulonglong(0)
        _check(lib().pdfs_selftest_sqrt_f32(self.c, C.byref(n)))
        return n.value

    def judge_f32(self, field, init, h=(1.0, 1.0, 1.0), flags=None):
        import torch
        Z, Y, X = field.shape
        n = C.c_longlong(0)
        _sync_inputs()
        _check(lib().pdfs_judge_f32_dev(self.c, self._ptr(field), self._ptr(init), X, Y, Z, *h,
                                        self._ptr(flags), C.byref(n)))
        return n.value

    def generate(self, kind, seed, X, Y, Z, h=1.0, dense=None, padded=None):
        import torch
        _sync_inputs()
        _check(lib().pdfs_generate_dev(self.c, kind, seed, X, Y, Z, h, self._ptr(dense),
                                       self._ptr(padded)))

    def generate_ex(self, kind, seed, X, Y, Z, h=1.0, nobj=0, p=(0.0, 0.0), q=(0.0, 0.0),
                    dense=None, padded=None):
        import torch
        _sync_inputs()
        _check(lib().pdfs_generate_ex_dev(self.c, kind, seed, X, Y, Z, h, nobj, p[0], p[1],
                                          q[0], q[1], self._ptr(dense), self._ptr(padded)))

    def pack(self, padded, dense):
        import torch
        Z, Y, X = dense.shape
        _sync_inputs()
        _check(lib().pdfs_pack_dev(self.c, self._ptr(padded), self._ptr(dense), X, Y, Z))
        self._sync()

    def unpack(self, dense, padded):
        import torch
        Z, Y, X = dense.shape
        _sync_inputs()
        _check(lib().pdfs_unpack_dev(self.c, self._ptr(dense), self._ptr(padded), X, Y, Z))
        self._sync()

    def _sync(self):
        """Wait for the context's own stream (calls that return before their
        kernels finish: pack, unpack, the network's feature and target
        kernels, warm starts)."""
        import torch
        torch.cuda.ExternalStream(lib().pdfs_ctx_stream(self.c)).synchronize()

    def solve_warm_f32(self, net, field, init, h=(1.0, 1.0, 1.0), allow_notconv=False, **kw):
        """Warm-started certified solve (in place); falls back to a plain
        solve from init when the judge flags the warm-started result."""
        import torch
        assert field.dtype == torch.float32 and field.dim() == 3
        Z, Y, X = field.shape
        o = default_opts(precision=PDFS_F32, **kw)
        st = Stats()
        _sync_inputs()
        rc = lib().pdfs_solve_warm_f32_dev(self.c, net.n, self._ptr(field), self._ptr(init), X, Y,
                                           Z, *h, C.byref(o), C.byref(st))
        _check(rc, (PDFS_OK, PDFS_ENOTCONV) if allow_notconv else (PDFS_OK,))
        return st.as_dict()


    def solve_batch_warm_f32(self, net, fields, inits, h=(1.0, 1.0, 1.0), allow_notconv=False,
                             **kw):
        """Batched warm-started certified solve (in place). Returns (stats,
        per-field passes, per-field fallback flags)."""
        import torch
        n = len(fields)
        ptrs = (C.c_void_p * n)(*[self._ptr(f) for f in fields])
        iptrs = (C.c_void_p * n)(*[self._ptr(t) for t in inits])
        dims = (C.c_int * (3 * n))()
        for i, f in enumerate(fields):
            assert f.dtype == torch.float32 and f.dim() == 3
            Z, Y, X = f.shape
            dims[3 * i], dims[3 * i + 1], dims[3 * i + 2] = X, Y, Z
        passes = (C.c_int * n)()
        fb = (C.c_int * n)()
        o = default_opts(precision=PDFS_F32, **kw)
        st = Stats()
        _sync_inputs()
        rc = lib().pdfs_solve_batch_warm_f32_dev(self.c, net.n, n, ptrs, iptrs, dims, *h,
                                                 C.byref(o), C.byref(st), passes, fb)
        _check(rc, (PDFS_OK, PDFS_ENOTCONV) if allow_notconv else (PDFS_OK,))
        return st.as_dict(), list(passes), list(fb)


def sweep_f32_noalloc(field, sn, result, h=(1.0, 1.0, 1.0), tile=-1, stream=None):
    """One FP32 ordering sn in place with no device allocation
    (pdfs_sweep_f32_noalloc), on `stream` (a torch.cuda.Stream, default the
    current one). result: int64 cuda tensor of 2 elements, [0] += points
    changed, [1] |= the edge-wait timeout flag; both written on the stream."""
    import torch
    assert field.dtype == torch.float32 and field.dim() == 3 and field.is_contiguous()
    assert result.dtype == torch.int64 and result.numel() >= 2 and result.is_cuda
    Z, Y, X = field.shape
    st = stream if stream is not None else torch.cuda.current_stream()
    _check(lib().pdfs_sweep_f32_noalloc(field.data_ptr(), X, Y, Z, *h, sn, tile, st.cuda_stream,
                                        result.data_ptr()))


def net_grid(X, Y, Z):
    """Coarse grid (Dz, Dy, Dx) of the warm-start network for a field."""
    d = [C.c_int(0) for _ in range(3)]
    lib().pdfs_net_grid(X, Y, Z, *(C.byref(v) for v in d))
    return tuple(v.value for v in d)


class Net:
    """The warm-start network (pdfs_net) of a Context. Tensors are torch
    cuda tensors: masks BF16 [nb, Dz, Dy, Dx, 16] (as int16 or bfloat16),
    targets FP32 [nb, Dz, Dy, Dx]."""

    def __init__(self, ctx, seed=1, path=None):
        self.ctx = ctx
        self.n = lib().pdfs_net_create(ctx.c, seed)
        if not self.n:
            raise PdfsError(PDFS_EHIP, lib().pdfs_last_error().decode())
        ctx._nets.add(self)  # closed before its context
        if path is not None:
            self.load(path)

    def close(self):
        if self.n:
            lib().pdfs_net_destroy(self.n)
            self.n = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, path=None):
        _check(lib().pdfs_net_load(self.n, (path or WEIGHTS_PATH).encode()))

    def save(self, path):
        _check(lib().pdfs_net_save(self.n, path.encode()))

    @property
    def num_params(self):
        return lib().pdfs_net_num_params(self.n)

    def params(self, which=0):
        import numpy as np
        out = np.empty(self.num_params, np.float32)
        _check(lib().pdfs_net_get(self.n, which, out.ctypes.data))
        return out

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        _check(lib().pdfs_net_set(self.n, p.ctypes.data))

    @property
    def calibration(self):
        v = [C.c_float(0) for _ in range(3)]
        lib().pdfs_net_get_calibration(self.n, *(C.byref(x) for x in v))
        return tuple(x.value for x in v)

    def set_calibration(self, scale, a, b):
        lib().pdfs_net_set_calibration(self.n, scale, a, b)

    @property
    def cut(self):
        """Cutoff in h: the warm start keeps the init where the raw
        prediction reaches cut * hmax (0: none)."""
        return lib().pdfs_net_get_cut(self.n)

    def set_cut(self, cut):
        lib().pdfs_net_set_cut(self.n, cut)

    def features(self, init, mask):
        import torch
        Z, Y, X = init.shape
        _sync_inputs()
        _check(lib().pdfs_net_features_dev(self.n, init.data_ptr(), X, Y, Z, mask.data_ptr()))
        self.ctx._sync()

    def target(self, u, t):
        import torch
        Z, Y, X = u.shape
        _sync_inputs()
        _check(lib().pdfs_net_target_dev(self.n, u.data_ptr(), X, Y, Z, t.data_ptr()))
        self.ctx._sync()

    def forward(self, mask, pred=None):
        import torch
        nb, Dz, Dy, Dx, ch = mask.shape
        assert ch == 16 and mask.is_contiguous()
        _sync_inputs()
        _check(lib().pdfs_net_forward_dev(self.n, mask.data_ptr(), nb, Dz, Dy, Dx,
                                          None if pred is None else pred.data_ptr()))
        return pred

    def train_step(self, mask, target, lr):
        import torch
        nb, Dz, Dy, Dx, ch = mask.shape
        assert ch == 16 and mask.is_contiguous() and target.is_contiguous()
        assert tuple(target.shape) == (nb, Dz, Dy, Dx) and target.dtype == torch.float32
        loss = C.c_float(0)
        _sync_inputs()
        _check(lib().pdfs_net_train_step_dev(self.n, mask.data_ptr(), target.data_ptr(), nb, Dz,
                                             Dy, Dx, lr, C.byref(loss)))
        return loss.value

    def warm_start(self, init, w, hmax=1.0):
        import torch
        Z, Y, X = init.shape
        _sync_inputs()
        _check(lib().pdfs_warm_start_f32_dev(self.n, init.data_ptr(), w.data_ptr(), X, Y, Z, hmax))
        self.ctx._sync()

    def warm_start_batch(self, inits, ws, hmax=1.0):
        """Warm starts of uneven fields in one launch per layer (ws[i] gets
        the warm start of inits[i]; dense [Z][Y][X] FP32 CUDA tensors)."""
        import torch
        nf = len(inits)
        assert nf == len(ws) and nf > 0
        ip = (C.c_void_p * nf)(*[t.data_ptr() for t in inits])
        wp = (C.c_void_p * nf)(*[t.data_ptr() for t in ws])
        dims = (C.c_int * (3 * nf))()
        for i, t in enumerate(inits):
            assert t.dtype == torch.float32 and t.is_contiguous() and ws[i].shape == t.shape
            dims[3 * i], dims[3 * i + 1], dims[3 * i + 2] = t.shape[2], t.shape[1], t.shape[0]
        _sync_inputs()
        _check(lib().pdfs_warm_start_batch_f32_dev(self.n, nf, ip, wp, dims, hmax))
        self.ctx._sync()
