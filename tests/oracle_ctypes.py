"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and to the reference
FSM core built from source (oracle/_ref/libpdfs_ref.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg. Arrays are padded (X+2)(Y+2)(Z+2), x fastest,
numpy shape (Z+2, Y+2, X+2).
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libpdfs_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_I, _D = C.c_int, C.c_double

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        G = [_D] * 7
        L.orc_godunov_ref.argtypes = G
        L.orc_godunov_ref.restype = _D
        L.orc_godunov_d64.argtypes = G
        L.orc_godunov_d64.restype = _D
        L.orc_godunov_f32.argtypes = [C.c_float] * 4 + [_D] * 3
        L.orc_godunov_f32.restype = C.c_float
        for nm, p in (("orc_sweep_ref64", _dp), ("orc_sweep_lex64", _dp),
                      ("orc_sweep_f32", _fp), ("orc_sweep_d64", _dp)):
            getattr(L, nm).argtypes = [p, _I, _I, _I, _D, _D, _D, _I]
            getattr(L, nm).restype = None
        L.orc_solve_ref64.argtypes = [_dp, _I, _I, _I, _D, _D, _D, _D, _I, C.POINTER(_D)]
        L.orc_solve_ref64.restype = _I
        L.orc_sweep_ref64_speed.argtypes = [_dp, _dp, _I, _I, _I, _D, _D, _D, _I]
        L.orc_sweep_ref64_speed.restype = None
        L.orc_solve_ref64_speed.argtypes = [_dp, _dp, _I, _I, _I, _D, _D, _D, _D, _I,
                                            C.POINTER(_D)]
        L.orc_solve_ref64_speed.restype = _I
        L.orc_solve_d64.argtypes = [_dp, _I, _I, _I, _D, _D, _D, _D, _I]
        L.orc_solve_d64.restype = _I
        L.orc_solve_f32.argtypes = [_fp, _I, _I, _I, _D, _D, _D, _I, C.POINTER(C.c_long)]
        L.orc_solve_f32.restype = _I
        L.orc_update_distance.argtypes = [_ip, _dp, C.c_size_t]
        L.orc_set_negative_inside.argtypes = [_ip, _dp, C.c_size_t]
        L.orc_adjust_boundary.argtypes = [_dp, _I, _I, _I]
        L.orc_fill_ghost.argtypes = [_dp, _I, _I, _I, _D]
        L.orc_gen_point.argtypes = [_dp, _I, _I, _I, C.c_uint64]
        L.orc_gen_terrain.argtypes = [_dp, _I, _I, _I, _D, C.c_uint64, _I, _D, _D, _D, _D]
        L.orc_gen_spheres.argtypes = [_dp, _I, _I, _I, _D, C.c_uint64, _I, _D, _D]
        L.orc_judge_f32.argtypes = [_fp, _fp, _I, _I, _I, _D, _D, _D, C.c_void_p]
        L.orc_judge_f32.restype = C.c_long
        L.orc_residual_f64.argtypes = [_dp, _dp, _I, _I, _I, _D, _D, _D]
        L.orc_residual_f64.restype = _D
        L.orc_sweep_polish64.argtypes = [_dp, _dp, _I, _I, _I, _D, _D, _D, _I]
        L.orc_sweep_polish64.restype = _D
        L.orc_polish64.argtypes = [_dp, _dp, _I, _I, _I, _D, _D, _D, _D, _I, C.POINTER(_D)]
        L.orc_polish64.restype = _I
        L.orc_godunov_f32_speed.argtypes = [C.c_float] * 4 + [_D] * 3 + [C.c_float]
        L.orc_godunov_f32_speed.restype = C.c_float
        L.orc_godunov_d64_speed.argtypes = [_D] * 8
        L.orc_godunov_d64_speed.restype = _D
        L.orc_sweep_f32_speed.argtypes = [_fp, _fp, _I, _I, _I, _D, _D, _D, _I]
        L.orc_sweep_f32_speed.restype = None
        L.orc_sweep_d64_speed.argtypes = [_dp, _dp, _I, _I, _I, _D, _D, _D, _I]
        L.orc_sweep_d64_speed.restype = None
        L.orc_solve_f32_speed.argtypes = [_fp, C.c_void_p, _I, _I, _I, _D, _D, _D, _I,
                                          C.POINTER(C.c_long)]
        L.orc_solve_f32_speed.restype = _I
        L.orc_sweep_polish64_speed.argtypes = [_dp, _dp, C.c_void_p, _I, _I, _I, _D, _D, _D, _I]
        L.orc_sweep_polish64_speed.restype = _D
        L.orc_polish64_speed.argtypes = [_dp, _dp, C.c_void_p, _I, _I, _I, _D, _D, _D, _D, _I,
                                         C.POINTER(_D)]
        L.orc_polish64_speed.restype = _I
        L.orc_judge_f32_speed.argtypes = [_fp, _fp, C.c_void_p, _I, _I, _I, _D, _D, _D,
                                          C.c_void_p]
        L.orc_judge_f32_speed.restype = C.c_long
        _up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
        L.orc_fx_bits.argtypes = [_I, _I, _I]
        L.orc_fx_bits.restype = _I
        L.orc_fx_min.argtypes = [_fp, C.c_size_t]
        L.orc_fx_min.restype = C.c_float
        L.orc_fx_encode.argtypes = [_fp, _up, C.c_size_t, _D, _D]
        L.orc_fx_decode.argtypes = [_up, _fp, C.c_size_t, _D, _D]
        L.orc_godunov_fx.argtypes = [C.c_uint32] * 4 + [_I]
        L.orc_godunov_fx.restype = C.c_uint32
        L.orc_sweep_fx.argtypes = [_up, _I, _I, _I, _I, _I, C.c_void_p]
        L.orc_sweep_fx.restype = C.c_long
        L.orc_solve_fx.argtypes = [_up, _I, _I, _I, _I, _I, _I, C.POINTER(C.c_long)]
        L.orc_solve_fx.restype = _I
        L.orc_judge_fx.argtypes = [_up, _fp, _I, _I, _I, _D, _D, _D, _D, C.c_void_p]
        L.orc_judge_fx.restype = C.c_long
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_SO)
        R.ref_run_sn.argtypes = [_dp, _I, _I, _I, _D, _D, _D, _I]
        R.ref_run_sn.restype = _I
        R.ref_solve.argtypes = [_dp, _I, _I, _I, _D, _D, _D, _D, _I, C.POINTER(_D), C.POINTER(_D)]
        R.ref_solve.restype = _I
        R.ref_time_orderings.argtypes = [_dp, _I, _I, _I, _D, _D, _D, _I]
        R.ref_time_orderings.restype = _D
        _ref = R
    return _ref


def dims(a):
    z, y, x = a.shape
    return x - 2, y - 2, z - 2


def padded(X, Y, Z, dtype=np.float64, fill=np.inf):
    return np.full((Z + 2, Y + 2, X + 2), fill, dtype=dtype)


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


# ---- convenience wrappers (all in place, return the array) ----
def sweep_ref64(a, sn, h=(1.0, 1.0, 1.0)):
    lib().orc_sweep_ref64(a, *dims(a), *h, sn)
    return a


def sweep_lex64(a, sn, h=(1.0, 1.0, 1.0)):
    lib().orc_sweep_lex64(a, *dims(a), *h, sn)
    return a


def sweep_f32(a, sn, h=(1.0, 1.0, 1.0)):
    lib().orc_sweep_f32(a, *dims(a), *h, sn)
    return a


def sweep_ref64_speed(a, s, sn, h=(1.0, 1.0, 1.0)):
    lib().orc_sweep_ref64_speed(a, s, *dims(a), *h, sn)
    return a


def solve_ref64_speed(a, s, h=(1.0, 1.0, 1.0), tol=1e-12, max_passes=100):
    md = _D()
    p = lib().orc_solve_ref64_speed(a, s, *dims(a), *h, tol, max_passes, C.byref(md))
    return p, md.value


def solve_ref64(a, h=(1.0, 1.0, 1.0), tol=1e-12, max_passes=100):
    md = _D()
    p = lib().orc_solve_ref64(a, *dims(a), *h, tol, max_passes, C.byref(md))
    return p, md.value


def solve_d64(a, h=(1.0, 1.0, 1.0), tol=0.0, max_passes=100):
    return lib().orc_solve_d64(a, *dims(a), *h, tol, max_passes)


def solve_f32(a, h=(1.0, 1.0, 1.0), max_passes=100):
    ch = C.c_long()
    p = lib().orc_solve_f32(a, *dims(a), *h, max_passes, C.byref(ch))
    return p, ch.value


def sweep_polish64(a, init, sn, h=(1.0, 1.0, 1.0)):
    """One ordering of u = min(init, G64(u)) in place; returns the largest change."""
    return lib().orc_sweep_polish64(a, init, *dims(a), *h, sn)


def polish64(a, init, h=(1.0, 1.0, 1.0), tol=2e-5, max_passes=100):
    """Polish passes until a pass's last ordering changes nothing by more than
    tol (in place); returns (passes, last pass's largest change)."""
    md = _D()
    p = lib().orc_polish64(a, init, *dims(a), *h, tol, max_passes, C.byref(md))
    return p, md.value


def solve_f32_polished(init, h=(1.0, 1.0, 1.0), tol_h=2e-5, max_passes=100):
    """The FP32 solve followed by the FP64 polish, as the GPU's default FP32
    solve runs it: returns (FP32 result (padded), FP32 passes, polish passes).
    init: padded FP64 whose interior values are FP32-representable."""
    u = init.astype(np.float32)
    p32, _ = solve_f32(u, h=h, max_passes=max_passes)
    v = u.astype(np.float64)
    i64 = init.astype(np.float32).astype(np.float64)
    pp, _ = polish64(v, i64, h=h, tol=tol_h * min(h), max_passes=max_passes)
    return v.astype(np.float32), p32, pp


def _fptr(s):
    if s is None:
        return None
    assert s.dtype == np.float32 and s.flags["C_CONTIGUOUS"]
    return s.ctypes.data


def sweep_f32_speed(a, s, sn, h=(1.0, 1.0, 1.0)):
    """One FP32 ordering of |grad u| = s in place (a, s padded float32)."""
    lib().orc_sweep_f32_speed(a, s, *dims(a), *h, sn)
    return a


def sweep_d64_speed(a, s, sn, h=(1.0, 1.0, 1.0)):
    lib().orc_sweep_d64_speed(a, s, *dims(a), *h, sn)
    return a


def solve_f32_speed(a, s, h=(1.0, 1.0, 1.0), max_passes=100):
    """FP32 passes of |grad u| = s until a pass changes nothing (s padded
    float32, or None for 1); returns (passes, points changed in the last)."""
    ch = C.c_long()
    p = lib().orc_solve_f32_speed(a, _fptr(s), *dims(a), *h, max_passes, C.byref(ch))
    return p, ch.value


def solve_d64_speed(a, s, h=(1.0, 1.0, 1.0), tol=0.0, max_passes=100):
    """FP64 difference-form passes of |grad u| = s (s padded float64) until a
    pass changes nothing by more than tol; returns the passes run."""
    for p in range(1, max_passes + 1):
        prev = a.copy()
        for sn in range(1, 9):
            sweep_d64_speed(a, s, sn, h)
        with np.errstate(invalid="ignore"):
            d = np.abs(a - prev)
        d[a == prev] = 0.0
        if not (np.nanmax(d) <= tol):
            continue
        return p
    return max_passes


def polish64_speed(a, init, s, h=(1.0, 1.0, 1.0), tol=2e-5, max_passes=100):
    md = _D()
    p = lib().orc_polish64_speed(a, init, _fptr(s), *dims(a), *h, tol, max_passes, C.byref(md))
    return p, md.value


def solve_f32_speed_polished(init, s, h=(1.0, 1.0, 1.0), tol_h=2e-5, max_passes=100):
    """The FP32 speed solve followed by the FP64 speed polish, as the GPU's
    default speed solve with an init runs it: returns (FP32 result (padded),
    FP32 passes, polish passes). s padded float32."""
    u = init.astype(np.float32)
    p32, _ = solve_f32_speed(u, s, h=h, max_passes=max_passes)
    v = u.astype(np.float64)
    i64 = init.astype(np.float32).astype(np.float64)
    pp, _ = polish64_speed(v, i64, s, h=h, tol=tol_h * min(h), max_passes=max_passes)
    return v.astype(np.float32), p32, pp


def judge_f32_speed(v, init, s, h=(1.0, 1.0, 1.0)):
    return lib().orc_judge_f32_speed(v, init, _fptr(s), *dims(v), *h, None)


def gen_point(X, Y, Z, seed):
    a = padded(X, Y, Z)
    lib().orc_gen_point(a, X, Y, Z, seed)
    return a


def gen_terrain(X, Y, Z, seed, h=1.0, nhills=8, amp=(5.0, 40.0), sig=(10.0, 60.0)):
    a = padded(X, Y, Z)
    lib().orc_gen_terrain(a, X, Y, Z, h, seed, nhills, amp[0], amp[1], sig[0], sig[1])
    return a


def gen_spheres(X, Y, Z, seed, h=1.0, n=64, r=(4.0, 32.0)):
    a = padded(X, Y, Z)
    lib().orc_gen_spheres(a, X, Y, Z, h, seed, n, r[0], r[1])
    return a


def judge_f32(v, init, h=(1.0, 1.0, 1.0), want_flags=False):
    flags = np.zeros(v.shape, np.uint8) if want_flags else None
    n = lib().orc_judge_f32(v, init, *dims(v), *h,
                            flags.ctypes.data if want_flags else None)
    return (n, flags) if want_flags else n


def ref_run_sn(a, sn, h=(1.0, 1.0, 1.0)):
    ref().ref_run_sn(a, *dims(a), *h, sn)
    return a


def ref_solve(a, h=(1.0, 1.0, 1.0), tol=1e-12, max_passes=100):
    sec, md = _D(), _D()
    p = ref().ref_solve(a, *dims(a), *h, tol, max_passes, C.byref(sec), C.byref(md))
    return p, sec.value, md.value


def random_field(X, Y, Z, seed, p_src=0.02, p_far=0.3, hi=20.0):
    """Random padded init: sources (0), far cells (1e6), the rest U(0, hi)."""
    rng = np.random.default_rng(seed)
    a = padded(X, Y, Z)
    v = rng.uniform(0.0, hi, size=(Z, Y, X))
    r = rng.uniform(size=(Z, Y, X))
    v[r < p_src] = 0.0
    v[r > 1.0 - p_far] = 1e6
    a[1:-1, 1:-1, 1:-1] = v
    return a


# ---- fixed-point storage ----
FX_INF = 0xFFFFFFFF


def fx_bits(X, Y, Z):
    return lib().orc_fx_bits(X, Y, Z)


def fx_params(v32, h=1.0):
    """(k, base, d) of a start field (FP32 array, any shape): base = its
    smallest non-NaN value (0 when it has none or only +inf), d = h 2^-k."""
    Z, Y, X = v32.shape if v32.ndim == 3 else (1, 1, v32.size)
    flat = np.ascontiguousarray(v32, dtype=np.float32).ravel()
    b = float(lib().orc_fx_min(flat, flat.size))
    if not np.isfinite(b):
        b = 0.0
    k = fx_bits(X, Y, Z)
    return k, b, float(np.ldexp(h, -k))


def fx_encode(v32, base, d):
    flat = np.ascontiguousarray(v32, dtype=np.float32).ravel()
    q = np.empty(flat.size, np.uint32)
    lib().orc_fx_encode(flat, q, flat.size, base, d)
    return q.reshape(v32.shape)


def fx_decode(q, base, d):
    flat = np.ascontiguousarray(q, dtype=np.uint32).ravel()
    v = np.empty(flat.size, np.float32)
    lib().orc_fx_decode(flat, v, flat.size, base, d)
    return v.reshape(q.shape)


def fx_padded(q_interior):
    """Padded uint32 array (ghosts +inf) around an interior [Z, Y, X]."""
    Z, Y, X = q_interior.shape
    a = np.full((Z + 2, Y + 2, X + 2), FX_INF, np.uint32)
    a[1:-1, 1:-1, 1:-1] = q_interior
    return a


def sweep_fx(q, sn, k, want_drop=False):
    """One fixed-point ordering on a padded uint32 array, in place; returns
    the points changed (and the largest drop in quanta)."""
    md = C.c_uint32()
    n = lib().orc_sweep_fx(q, *dims(q), k, sn, C.addressof(md))
    return (n, md.value) if want_drop else n


def solve_fx(q, k, max_passes=100, stop_quanta=4):
    ch = C.c_long()
    p = lib().orc_solve_fx(q, *dims(q), k, stop_quanta, max_passes, C.byref(ch))
    return p, ch.value


def solve_fixed(start, h=1.0, max_passes=100, stop_quanta=4):
    """The GPU's fixed-point solve of a dense FP32 start field [Z, Y, X]:
    base, quantum, encode, passes until one changes no value by more than
    stop_quanta quanta, decode. Returns (FP32 result, passes, (k, base, d),
    padded uint32 result)."""
    start = np.ascontiguousarray(start, dtype=np.float32)
    k, b, d = fx_params(start, h)
    q = fx_padded(fx_encode(start, b, d))
    p, _ = solve_fx(q, k, max_passes, stop_quanta)
    return fx_decode(interior(q).copy(), b, d), p, (k, b, d), q


def judge_fx(q, init32, h, base, d, tol, want_flags=False):
    """q: padded uint32, init32: padded float32."""
    flags = np.zeros(q.shape, np.uint8) if want_flags else None
    n = lib().orc_judge_fx(q, init32, *dims(q), h, base, d, tol,
                           flags.ctypes.data if want_flags else None)
    return (n, flags) if want_flags else n
