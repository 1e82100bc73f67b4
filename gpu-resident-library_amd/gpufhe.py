"""Python view of libgpufhe_amd.so (ctypes over include/gpufhe_capi.h and
include/gpufhe_kernels.h).  The library is the product; this module only
loads it, declares the signatures and wraps device buffers.  There is no
fallback: if the shared library is missing, import fails loudly."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpufhe_amd.so")


class GpuFheError(RuntimeError):
    pass


if not os.path.exists(LIB_PATH):
    raise ImportError("libgpufhe_amd.so is not built (run __graft_entry__.build() or "
                      "python gpu-resident-library_amd/build.py)")
lib = C.CDLL(LIB_PATH)

vp = C.c_void_p
u32, u64, sz = C.c_uint32, C.c_uint64, C.c_size_t
u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
vpp = C.POINTER(C.c_void_p)

_capi = {
    "gf_last_error": (C.c_char_p, []),
    "gf_malloc": (C.c_int, [vpp, sz]),
    "gf_free": (C.c_int, [vp]),
    "gf_h2d": (C.c_int, [vp, vp, sz]),
    "gf_d2h": (C.c_int, [vp, vp, sz]),
    "gf_d2d": (C.c_int, [vp, vp, sz]),
    "gf_device_sync": (C.c_int, []),
    "gf_stream_create": (C.c_int, [vpp]),
    "gf_stream_destroy": (C.c_int, [vp]),
    "gf_stream_sync": (C.c_int, [vp]),
    "gf_native_primes": (C.c_int, [u64p, u32, u32, u32, u64p, u32]),
    "gf_native_root": (u64, [u32, u64]),
    "gf_native_rescale_consts": (None, [u64p, u64p, u64p, u32]),
    "gf_ctx_create": (vp, [u32, u64p, u32, u64p, u32, u64p, u32]),
    "gf_ctx_destroy": (None, [vp]),
    "gf_ws_bytes": (sz, [vp]),
    "gf_rescale_ws_bytes": (sz, [vp]),
    "gf_key_create": (vp, [vp, u64p, u64p, u32, vp]),
    "gf_key_destroy": (None, [vp]),
    "gf_ctx_create_alpha": (vp, [u32, u64p, u32, u64p, u32, u64p, u32]),
    "gf_ctx_towers": (u32, [vp]),
    "gf_ctx_cached_plans": (u32, [vp]),
    "gf_key_create_rows": (vp, [vp, u64p, u64p, u32, u32, u32, vp]),
    "gf_key_create_uniform": (vp, [vp, u32, u64, vp]),
    "gf_key_generate": (vp, [vp, vp, vp, vp, u64, C.c_double, u64, vp]),
    "gf_key_download": (C.c_int, [vp, u64p, u64p, vp]),
    "gf_key_generate_ctr": (vp, [vp, vp, vp, vp, u64, C.c_double, u64, vp]),
    "gf_keygen": (C.c_int, [vp, u64, C.c_double, u64, vp, vp, vp, vp]),
    "gf_encrypt": (C.c_int, [vp, u32, vp, vp, vp, u64, C.c_double, u64, vp, vp, vp]),
    "gf_decrypt": (C.c_int, [vp, vp, vp, vp, vp, vp]),
    "gf_key_shape": (None, [vp, u32p, u32p, u32p]),
    "gf_ct_packed_bytes": (sz, [u32, u32]),
    "gf_ct_pack": (C.c_int, [u64p, u64p, u32, u32, u64p, C.c_double, vp, sz]),
    "gf_ct_unpack_header": (C.c_int, [vp, sz, u32p, u32p, C.POINTER(C.c_double)]),
    "gf_ct_unpack": (C.c_int, [vp, sz, u64p, u64p, u64p]),
    "gf_key_packed_bytes": (sz, [vp]),
    "gf_key_pack": (C.c_int, [vp, vp, sz, vp]),
    "gf_key_unpack": (vp, [vp, vp, sz, vp]),
    "gf_root_tables": (C.c_int, [u32, u64, u64, vpp, vpp, vpp, vpp, u64p, u64p]),
    "gf_ntt_towers": (C.c_int, [u32, vp, u32, u64p, u64p, C.c_int, vp]),
    "gf_automorph": (C.c_int, [u32, vp, vp, u32, u64, vp]),
    "gf_rescale_comp": (C.c_int, [vp, u32, vp, vp, vp, vp]),
    "gf_evalmult_rescale": (C.c_int, [vp, vp, u32] + [vp] * 7 + [vp]),
    "gf_evalmult": (C.c_int, [vp, vp, u32] + [vp] * 7 + [vp]),
    "gf_evalmult_rescale_noalloc": (C.c_int, [vp, vp, u32] + [vp] * 6 + [vp]),
    "gf_noalloc_scratch_bytes": (sz, []),
    "gf_keyswitch": (C.c_int, [vp, vp, u32, vp, vp, vp, vp, vp]),
    "gf_rescale": (C.c_int, [vp, u32, vp, vp, vp, vp, vp, vp]),
    "gf_rotate": (C.c_int, [vp, vp, u32, u64, vp, vp, vp, vp, vp, vp]),
    "gf_add": (C.c_int, [vp, u32, vp, vp, vp, vp, vp, vp, vp]),
    "gf_mul_plain": (C.c_int, [vp, u32, vp, vp, vp, vp, vp, vp]),
    "gf_add_plain": (C.c_int, [vp, u32, vp, vp, vp, vp, vp, vp]),
    "gf_rotate_hoisted": (C.c_int, [vp, u32, vp, vp, u32, vpp, u64p, vpp, vpp, vp, vpp, vp]),
    "gf_linear_transform": (C.c_int, [vp, u32, vp, vp, u32, vpp, u64p, vpp, vp, vp, vp, vpp, vp]),
    "gf_plan_mul_plain": (vp, [vp, u32, u32p] + [vpp] * 5),
    "gf_plan_add_plain": (vp, [vp, u32, u32p] + [vpp] * 5),
    "gf_plan_evalmult_rescale": (vp, [vp, vp, u32, u32p] + [vpp] * 7),
    "gf_plan_rotate": (vp, [vp, vpp, u32, u32p, u64p] + [vpp] * 5),
    "gf_plan_add": (vp, [vp, u32, u32p] + [vpp] * 6),
    "gf_plan_rotate_hoisted": (vp, [vp, u32, u32p, vpp, vpp, u32, vpp, u64p, vpp, vpp, vpp, vpp]),
    "gf_plan_linear_transform": (vp, [vp, u32, u32p, vpp, vpp, u32, vpp, u64p, vpp, vpp, vpp, vpp, vpp]),
    "gf_plan_run": (C.c_int, [vp, vp]),
    "gf_plan_run_per_job": (C.c_int, [vp, vp]),
    "gf_plan_run_graph": (C.c_int, [vp, vp]),
    "gf_plan_launches": (u32, [vp]),
    "gf_plan_destroy": (None, [vp]),
    "gf_plan_num_steps": (u32, [vp]),
    "gf_plan_step_info": (C.c_int, [vp, u32, u32p, u32p]),
    "gf_plan_run_step": (C.c_int, [vp, u32, vp]),
    "gf_plan_step_kernels": (u32, [vp, u32]),
    "gf_plan_run_step_kernel": (C.c_int, [vp, u32, u32, vp]),
    "gf_fill_uniform": (C.c_int, [vp, vp, u32, u64, vp]),
    "gf_ctx_set_ntt_radix": (C.c_int, [vp, u32]),
    "gf_ctx_get_ntt_radix": (u32, [vp]),
    "gf_ctx_autotune": (C.c_int, [vp, vp, u32p, C.POINTER(C.c_float)]),
    "gf_ctx_set_fusion": (C.c_int, [vp, C.c_int]),
    "gf_ctx_autotune_plan": (C.c_int, [vp, u32, vp]),
    "gf_ctx_tuning_report": (C.c_char_p, [vp]),
    "gf_ctx_set_knob": (C.c_int, [vp, C.c_char_p, u32]),
    "gf_ctx_get_knob": (C.c_int, [vp, C.c_char_p, u32p]),
    "gf_event_create": (C.c_int, [vpp]),
    "gf_event_destroy": (C.c_int, [vp]),
    "gf_event_record": (C.c_int, [vp, vp]),
    "gf_event_elapsed_ms": (C.c_int, [C.POINTER(C.c_float), vp, vp]),
    "gf_capture_begin": (C.c_int, [vp]),
    "gf_capture_end": (C.c_int, [vp, vpp]),
    "gf_graph_launch": (C.c_int, [vp, vp]),
    "gf_graph_destroy": (C.c_int, [vp]),
}
_kern = {
    "LaunchNTT_CT": [vp, vp, vp, u32, u64, vp],
    "LaunchINTT_GS": [vp, vp, vp, u32, u64, u64, u64, vp],
    "LaunchRNSMultTower": [vp, vp, vp, u64, u32, vp],
    "LaunchSwitchModulus": [vp, u64, u64, u32, vp],
    "LaunchRescaleFuse": [vp, vp, u64, u64, u64, u32, vp],
    "LaunchApproxSwitchCRTBasis": [vp] * 9 + [u32, u32, u32, vp],
    "LaunchKSMultAcc": [vp, vp, vp, vp, vp, u64, u32, vp],
    "LaunchKSMultAccBatched": [vp] * 7 + [u32, u32, vp],
    "LaunchModDownCombine": [vp, vp, vp, u64, u64, u32, vp],
    "LaunchAutomorphPermute": [vp, vp, u32, u64, u64, vp],
    "LaunchAddInto": [vp, vp, u64, u32, vp],
    "LaunchPtMulAcc": [vp, vp, vp, vp, vp, u64, u32, vp],
    "LaunchTensor": [vp] * 8 + [u32, u32, vp],
    "LaunchCombine": [vp] * 7 + [u32, u32, vp],
}
for _n, (_r, _a) in _capi.items():
    _f = getattr(lib, _n)
    _f.restype = _r
    _f.argtypes = _a
for _n, _a in _kern.items():
    _f = getattr(lib, _n)
    _f.restype = None
    _f.argtypes = _a


def check(rc):
    if rc != 0:
        raise GpuFheError(lib.gf_last_error().decode())


def P64(a):
    return np.ascontiguousarray(a, dtype=np.uint64).ctypes.data_as(u64p)


def ptr_array(ptrs):
    return (C.c_void_p * len(ptrs))(*[int(p) for p in ptrs])


class Dev:
    """A device buffer of uint64 words."""

    def __init__(self, words):
        self.words = int(words)
        p = C.c_void_p()
        check(lib.gf_malloc(C.byref(p), max(self.words, 1) * 8))
        self.ptr = p.value

    @classmethod
    def of(cls, arr):
        a = np.ascontiguousarray(arr, dtype=np.uint64)
        d = cls(a.size)
        d.upload(a)
        return d

    def upload(self, arr, word_off=0):
        a = np.ascontiguousarray(arr, dtype=np.uint64)
        check(lib.gf_h2d(self.ptr + word_off * 8, a.ctypes.data, a.nbytes))

    def download(self, words=None, word_off=0):
        w = self.words - word_off if words is None else int(words)
        out = np.empty(w, np.uint64)
        check(lib.gf_d2h(out.ctypes.data, self.ptr + word_off * 8, w * 8))
        return out

    def at(self, word_off):
        return self.ptr + int(word_off) * 8

    def free(self):
        if getattr(self, "ptr", None):
            lib.gf_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(lib.gf_stream_create(C.byref(s)))
        self.s = s.value

    def sync(self):
        check(lib.gf_stream_sync(self.s))

    def __del__(self):
        if getattr(self, "s", None):
            lib.gf_stream_destroy(self.s)
            self.s = None


def device_sync():
    check(lib.gf_device_sync())


class Context:
    """Resident parameter set (Q moduli, P moduli, roots, dnum)."""

    def __init__(self, n, q, p, roots, dnum):
        self.n, self.L, self.K, self.dnum = n, len(q), len(p), dnum
        self.q = np.asarray(q, np.uint64)
        self.p = np.asarray(p, np.uint64)
        self.h = lib.gf_ctx_create(n, P64(q), len(q), P64(p), len(p), P64(roots), dnum)
        if not self.h:
            raise GpuFheError(lib.gf_last_error().decode())
        self.ws_bytes = lib.gf_ws_bytes(self.h)

    def workspace(self):
        return Dev(self.ws_bytes // 8)

    def set_ntt_radix(self, r):
        check(lib.gf_ctx_set_ntt_radix(self.h, r))

    def set_fusion(self, on):
        check(lib.gf_ctx_set_fusion(self.h, 1 if on else 0))

    def ntt_radix(self):
        return lib.gf_ctx_get_ntt_radix(self.h)

    def cached_plans(self):
        """single-op plans held in the context's cache"""
        return lib.gf_ctx_cached_plans(self.h)

    def autotune(self, stream, batch=32):
        """NTT register radix (synthetic tower batch), then the plan knobs on
        an EvalMult+Rescale batch of `batch` ops at the top level (batch=0
        skips them).  Returns (radix, ms per 96 forward + inverse transforms)."""
        r, ms = C.c_uint32(), C.c_float()
        check(lib.gf_ctx_autotune(self.h, stream.s, C.byref(r), C.byref(ms)))
        if batch:
            check(lib.gf_ctx_autotune_plan(self.h, batch, stream.s))
        return r.value, ms.value

    def tuning_report(self):
        import json
        rep = lib.gf_ctx_tuning_report(self.h).decode()
        return json.loads(rep) if rep else None

    def set_knob(self, name, value):
        check(lib.gf_ctx_set_knob(self.h, name.encode(), int(value)))

    def knob(self, name):
        v = C.c_uint32()
        check(lib.gf_ctx_get_knob(self.h, name.encode(), C.byref(v)))
        return v.value

    def __del__(self):
        if getattr(self, "h", None):
            lib.gf_ctx_destroy(self.h)
            self.h = None


class Key:
    def __init__(self, ctx, av, bv, nparts, stream=None):
        self.ctx = ctx
        self.h = lib.gf_key_create(ctx.h, P64(av), P64(bv), nparts, stream.s if stream else None)
        if not self.h:
            raise GpuFheError(lib.gf_last_error().decode())

    @classmethod
    def generate(cls, ctx, d_sold, d_pka, d_pkb, seed, ns=1, sigma=3.19, stream=None):
        """Switching key generated on the device from device arrays [L+K][n]."""
        k = cls.__new__(cls)
        k.ctx = ctx
        k.h = lib.gf_key_generate(ctx.h, d_sold.ptr, d_pka.ptr, d_pkb.ptr, ns, sigma, seed,
                                  stream.s if stream else None)
        if not k.h:
            raise GpuFheError(lib.gf_last_error().decode())
        return k

    @classmethod
    def generate_ctr(cls, ctx, d_sold, d_pka, d_pkb, seed, ns=1, sigma=3.19, stream=None):
        """gf_key_generate with u, e0, e1 drawn on the device (counter sampler)."""
        k = cls.__new__(cls)
        k.ctx = ctx
        k.h = lib.gf_key_generate_ctr(ctx.h, d_sold.ptr, d_pka.ptr, d_pkb.ptr, ns, sigma, seed,
                                      stream.s if stream else None)
        if not k.h:
            raise GpuFheError(lib.gf_last_error().decode())
        return k

    def download(self, stream=None):
        """(av, bv) in plain form, [nparts][rows][n] flattened."""
        p, r, f = C.c_uint32(), C.c_uint32(), C.c_uint32()
        lib.gf_key_shape(self.h, C.byref(p), C.byref(r), C.byref(f))
        words = p.value * r.value * self.ctx.n
        av, bv = np.empty(words, np.uint64), np.empty(words, np.uint64)
        check(lib.gf_key_download(self.h, av.ctypes.data_as(u64p), bv.ctypes.data_as(u64p),
                                  stream.s if stream else None))
        return av, bv

    @classmethod
    def uniform(cls, ctx, nparts, seed, stream=None):
        """Synthetic key of uniform residues generated on the device."""
        k = cls.__new__(cls)
        k.ctx = ctx
        k.h = lib.gf_key_create_uniform(ctx.h, nparts, seed, stream.s if stream else None)
        if not k.h:
            raise GpuFheError(lib.gf_last_error().decode())
        return k

    def __del__(self):
        if getattr(self, "h", None):
            lib.gf_key_destroy(self.h)
            self.h = None


class Plan:
    def __init__(self, h, keep=()):
        if not h:
            raise GpuFheError(lib.gf_last_error().decode())
        self.h = h
        self.keep = keep
        self.launches = lib.gf_plan_launches(h)

    def run(self, stream):
        check(lib.gf_plan_run(self.h, stream.s if stream else None))

    def run_per_job(self, stream):
        check(lib.gf_plan_run_per_job(self.h, stream.s if stream else None))

    def run_graph(self, stream):
        check(lib.gf_plan_run_graph(self.h, stream.s))

    def steps(self):
        out = []
        k, j = C.c_uint32(), C.c_uint32()
        for i in range(lib.gf_plan_num_steps(self.h)):
            check(lib.gf_plan_step_info(self.h, i, C.byref(k), C.byref(j)))
            out.append((STEP_KINDS[k.value], j.value))
        return out

    def run_step(self, i, stream):
        check(lib.gf_plan_run_step(self.h, i, stream.s))

    def step_kernels(self, i):
        return int(lib.gf_plan_step_kernels(self.h, i))

    def run_step_kernel(self, i, which, stream):
        check(lib.gf_plan_run_step_kernel(self.h, i, which, stream.s))

    def __del__(self):
        if getattr(self, "h", None):
            lib.gf_plan_destroy(self.h)
            self.h = None


STEP_KINDS = ["ntt_fwd", "ntt_inv", "tensor", "conv", "inner", "combine", "switch", "fuse", "add",
              "perm", "copy", "rsv", "ptacc", "ntt_fwd_a", "ksi", "ntt_inv_b", "modup", "moddown", "ntt_fwd_b",
              "tensor_ib"]


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(lib.gf_event_create(C.byref(e)))
        self.e = e.value

    def record(self, stream):
        check(lib.gf_event_record(self.e, stream.s))

    def elapsed_ms(self, later):
        ms = C.c_float()
        check(lib.gf_event_elapsed_ms(C.byref(ms), self.e, later.e))
        return ms.value

    def __del__(self):
        if getattr(self, "e", None):
            lib.gf_event_destroy(self.e)
            self.e = None


class Graph:
    """What fn() launches on `stream`, captured once and replayed."""

    def __init__(self, stream, fn):
        check(lib.gf_capture_begin(stream.s))
        try:
            fn()
        finally:
            g = C.c_void_p()
            rc = lib.gf_capture_end(stream.s, C.byref(g))
        check(rc)
        self.g = g.value

    def launch(self, stream):
        check(lib.gf_graph_launch(self.g, stream.s))

    def __del__(self):
        if getattr(self, "g", None):
            lib.gf_graph_destroy(self.g)
            self.g = None


class RefSequence:
    """One chain of the reference's EvalMult+Rescale launch sequence
    (bench/bench_evalmult_resident.cpp:96-100) over the drop-in boundary:
    LaunchTensor -> keyswitch_resident (gf_keyswitch) -> LaunchCombine ->
    rescale_resident_raw (gf_rescale_comp) on each component, on the chain's
    own stream and buffers.  Outputs: r0, r1 (level - 1 towers, in place)."""

    def __init__(self, ctx, key, level, d_mods, seed):
        n = ctx.n
        self.ctx, self.key, self.level, self.d_mods = ctx, key, level, d_mods
        self.stream = Stream()
        self.ins = [Dev(level * n) for _ in range(4)]
        for j, d in enumerate(self.ins):
            check(lib.gf_fill_uniform(ctx.h, d.ptr, level, seed + j, self.stream.s))
        self.t = [Dev(level * n) for _ in range(3)]
        self.ba = [Dev(level * n) for _ in range(2)]
        self.r = [Dev(level * n) for _ in range(2)]
        self.ws = ctx.workspace()
        self.rs = Dev(lib.gf_rescale_ws_bytes(ctx.h) // 8)
        self.graph = None
        self.stream.sync()

    def launch(self):
        n, L, s = self.ctx.n, self.level, self.stream.s
        a0, a1, b0, b1 = (d.ptr for d in self.ins)
        t0, t1, t2 = (d.ptr for d in self.t)
        lib.LaunchTensor(a0, a1, b0, b1, t0, t1, t2, self.d_mods.ptr, L, n, s)
        check(lib.gf_keyswitch(self.ctx.h, self.key.h, L, t2, self.ba[0].ptr, self.ba[1].ptr, self.ws.ptr, s))
        lib.LaunchCombine(self.r[0].ptr, self.r[1].ptr, t0, t1, self.ba[0].ptr, self.ba[1].ptr, self.d_mods.ptr,
                          L, n, s)
        for r in self.r:
            check(lib.gf_rescale_comp(self.ctx.h, L, r.ptr, r.ptr, self.rs.ptr, s))

    def capture(self):
        """Capture the sequence as a graph (call launch() once first, so the
        single-op plans it uses are built outside the capture)."""
        self.graph = Graph(self.stream, self.launch)

    def outputs(self):
        w = (self.level - 1) * self.ctx.n
        return self.r[0].download(w), self.r[1].download(w)


def keygen(ctx, seed, sigma=3.19, ns=1, stream=None):
    """Public key generated on the device: (s, pkA, pkB) as device buffers of
    [L+K][n] words, evaluation form."""
    w = (ctx.L + ctx.K) * ctx.n
    s, a, b = Dev(w), Dev(w), Dev(w)
    check(lib.gf_keygen(ctx.h, seed, sigma, ns, s.ptr, a.ptr, b.ptr, stream.s if stream else None))
    return s, a, b


def plan_evalmult_rescale(ctx, key, levels, c0a, c1a, c0b, c1b, o0, o1, ws):
    cnt = len(levels)
    lv = np.asarray(levels, np.uint32)
    arrs = [ptr_array(x) for x in (c0a, c1a, c0b, c1b, o0, o1, ws)]
    return Plan(lib.gf_plan_evalmult_rescale(ctx.h, key.h, cnt, lv.ctypes.data_as(u32p), *arrs),
                keep=(lv, arrs))


def plan_rotate(ctx, keys, levels, ks, c0, c1, o0, o1, ws):
    cnt = len(levels)
    lv = np.asarray(levels, np.uint32)
    kk = np.asarray(ks, np.uint64)
    kh = ptr_array([k.h for k in keys])
    arrs = [ptr_array(x) for x in (c0, c1, o0, o1, ws)]
    return Plan(lib.gf_plan_rotate(ctx.h, kh, cnt, lv.ctypes.data_as(u32p), kk.ctypes.data_as(u64p), *arrs),
                keep=(lv, kk, kh, arrs, keys))


def plan_rotate_hoisted(ctx, levels, c0, c1, keys, ks, o0, o1, ws_base, ws_rot):
    """Each ciphertext i rotated by every ks[j]; o0/o1/ws_rot indexed [i*m + j]."""
    m = len(ks)
    lv = np.asarray(levels, np.uint32)
    kk = np.asarray(ks, np.uint64)
    kh = ptr_array([k.h for k in keys])
    arrs = [ptr_array(x) for x in (c0, c1)]
    outs = [ptr_array(x) for x in (o0, o1, ws_base, ws_rot)]
    return Plan(lib.gf_plan_rotate_hoisted(ctx.h, len(levels), lv.ctypes.data_as(u32p), arrs[0], arrs[1], m, kh,
                                           kk.ctypes.data_as(u64p), *outs),
                keep=(lv, kk, kh, arrs, outs, keys))


def plan_linear_transform(ctx, levels, c0, c1, keys, ks, pts, o0, o1, ws_base, ws_rot):
    """o_i = sum_j pts[j] * rot_ks[j](c_i); keys[j] None with ks[j] = 1 for the
    unrotated term; ws_rot indexed [i*m + j]."""
    m = len(ks)
    lv = np.asarray(levels, np.uint32)
    kk = np.asarray(ks, np.uint64)
    kh = ptr_array([k.h if k is not None else 0 for k in keys])
    arrs = [ptr_array(x) for x in (c0, c1, pts, o0, o1, ws_base, ws_rot)]
    return Plan(lib.gf_plan_linear_transform(ctx.h, len(levels), lv.ctypes.data_as(u32p), arrs[0], arrs[1], m, kh,
                                             kk.ctypes.data_as(u64p), *arrs[2:]),
                keep=(lv, kk, kh, arrs, keys))


def plan_mul_plain(ctx, levels, c0, c1, pt, o0, o1):
    lv = np.asarray(levels, np.uint32)
    arrs = [ptr_array(x) for x in (c0, c1, pt, o0, o1)]
    return Plan(lib.gf_plan_mul_plain(ctx.h, len(levels), lv.ctypes.data_as(u32p), *arrs), keep=(lv, arrs))


def plan_add_plain(ctx, levels, c0, c1, pt, o0, o1):
    lv = np.asarray(levels, np.uint32)
    arrs = [ptr_array(x) for x in (c0, c1, pt, o0, o1)]
    return Plan(lib.gf_plan_add_plain(ctx.h, len(levels), lv.ctypes.data_as(u32p), *arrs), keep=(lv, arrs))


def mul_plain(ctx, level, c0, c1, pt, o0, o1, stream):
    check(lib.gf_mul_plain(ctx.h, level, c0, c1, pt, o0, o1, stream.s))


def add_plain(ctx, level, c0, c1, pt, o0, o1, stream):
    check(lib.gf_add_plain(ctx.h, level, c0, c1, pt, o0, o1, stream.s))


def rotate_hoisted(ctx, level, c0, c1, keys, ks, o0, o1, ws_base, ws_rot, stream):
    kk = np.asarray(ks, np.uint64)
    check(lib.gf_rotate_hoisted(ctx.h, level, c0, c1, len(ks), ptr_array([k.h for k in keys]),
                                kk.ctypes.data_as(u64p), ptr_array(o0), ptr_array(o1), ws_base, ptr_array(ws_rot),
                                stream.s))


def linear_transform(ctx, level, c0, c1, keys, ks, pts, o0, o1, ws_base, ws_rot, stream):
    kk = np.asarray(ks, np.uint64)
    check(lib.gf_linear_transform(ctx.h, level, c0, c1, len(ks),
                                  ptr_array([k.h if k is not None else 0 for k in keys]), kk.ctypes.data_as(u64p),
                                  ptr_array(pts), o0, o1, ws_base, ptr_array(ws_rot), stream.s))


def plan_add(ctx, levels, a0, a1, b0, b1, o0, o1):
    lv = np.asarray(levels, np.uint32)
    arrs = [ptr_array(x) for x in (a0, a1, b0, b1, o0, o1)]
    return Plan(lib.gf_plan_add(ctx.h, len(levels), lv.ctypes.data_as(u32p), *arrs), keep=(lv, arrs))


# ---- serving ----
OP_EVALMULT_RESCALE, OP_ROTATE, OP_ADD = 0, 1, 2


class Request(C.Structure):
    _fields_ = [("ring", C.c_uint32), ("op", C.c_uint32), ("level", C.c_uint32), ("rot", C.c_uint32),
                ("a0", C.c_void_p), ("a1", C.c_void_p), ("b0", C.c_void_p), ("b1", C.c_void_p),
                ("o0", C.c_void_p), ("o1", C.c_void_p), ("arrival_ms", C.c_double)]


class Completion(C.Structure):
    _fields_ = [("id", C.c_uint64), ("arrival_ms", C.c_double), ("launch_ms", C.c_double),
                ("done_ms", C.c_double), ("batch_size", C.c_uint32), ("ring", C.c_uint32),
                ("op", C.c_uint32), ("level", C.c_uint32)]


_serve = {
    "gf_server_create": (vp, [sz, u32, u32]),
    "gf_server_destroy": (None, [vp]),
    "gf_server_add_ring": (C.c_int, [vp, vp, vp, vpp, u64p, u32, u32p]),
    "gf_server_slot_bytes": (sz, [vp, u32]),
    "gf_server_now_ms": (C.c_double, [vp]),
    "gf_server_submit": (C.c_int, [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "gf_server_step": (C.c_int, [vp, u32p]),
    "gf_server_poll": (u32, [vp, C.POINTER(Completion), u32]),
    "gf_server_pending": (u32, [vp]),
    "gf_server_drain": (C.c_int, [vp]),
    "gf_server_run_trace": (C.c_int, [vp, u32, C.POINTER(Request), C.POINTER(Completion)]),
    "gf_server_stats": (C.c_int, [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_size_t)]),
}
for _n, (_r, _a) in _serve.items():
    _f = getattr(lib, _n)
    _f.restype = _r
    _f.argtypes = _a


class Server:
    """Continuous-batching server over resident ciphertexts (gf_server_*)."""

    def __init__(self, ws_budget_bytes, lanes=4, max_batch=64):
        self.h = lib.gf_server_create(ws_budget_bytes, lanes, max_batch)
        if not self.h:
            raise GpuFheError(lib.gf_last_error().decode())
        self._keep = []

    def add_ring(self, ctx, rlk=None, rot_keys=(), rot_k=()):
        arr = (C.c_void_p * max(1, len(rot_keys)))(*[k.h for k in rot_keys])
        ks = np.asarray(list(rot_k) or [1], np.uint64)
        out = C.c_uint32()
        check(lib.gf_server_add_ring(self.h, ctx.h, rlk.h if rlk else None, arr, ks.ctypes.data_as(u64p),
                                     len(rot_keys), C.byref(out)))
        self._keep.append((ctx, rlk, tuple(rot_keys)))
        return out.value

    def slot_bytes(self, ring):
        return int(lib.gf_server_slot_bytes(self.h, ring))

    def run_trace(self, reqs):
        n = len(reqs)
        arr = (Request * n)(*reqs)
        out = (Completion * n)()
        check(lib.gf_server_run_trace(self.h, n, arr, out))
        return list(out)

    def submit(self, req):
        rid = C.c_uint64()
        check(lib.gf_server_submit(self.h, C.byref(req), C.byref(rid)))
        return rid.value

    def step(self):
        n = C.c_uint32()
        check(lib.gf_server_step(self.h, C.byref(n)))
        return n.value

    def poll(self, max_n=4096):
        out = (Completion * max_n)()
        k = lib.gf_server_poll(self.h, out, max_n)
        return list(out[:k])

    def drain(self):
        check(lib.gf_server_drain(self.h))

    def stats(self):
        b, a, s = C.c_uint64(), C.c_uint64(), C.c_size_t()
        lib.gf_server_stats(self.h, C.byref(b), C.byref(a), C.byref(s))
        return dict(batches=b.value, admitted=a.value, slot_bytes=s.value)

    def __del__(self):
        if getattr(self, "h", None):
            lib.gf_server_destroy(self.h)
            self.h = None


# ---- wire format ----
def pack_ciphertext(c0, c1, n, moduli, scale=0.0):
    """Host arrays (towers x n each) -> bytes with header and checksum."""
    c0 = np.ascontiguousarray(c0, np.uint64)
    c1 = np.ascontiguousarray(c1, np.uint64)
    mods = np.ascontiguousarray(moduli, np.uint64)
    towers = len(mods)
    out = C.create_string_buffer(lib.gf_ct_packed_bytes(n, towers))
    check(lib.gf_ct_pack(c0.ctypes.data_as(u64p), c1.ctypes.data_as(u64p), n, towers, mods.ctypes.data_as(u64p),
                         scale, out, len(out)))
    return out.raw


def unpack_ciphertext(buf):
    """bytes -> (c0, c1, n, moduli, scale); raises on any mismatch."""
    n, t, sc = C.c_uint32(), C.c_uint32(), C.c_double()
    check(lib.gf_ct_unpack_header(buf, len(buf), C.byref(n), C.byref(t), C.byref(sc)))
    mods = np.empty(t.value, np.uint64)
    c0 = np.empty(t.value * n.value, np.uint64)
    c1 = np.empty(t.value * n.value, np.uint64)
    check(lib.gf_ct_unpack(buf, len(buf), mods.ctypes.data_as(u64p), c0.ctypes.data_as(u64p),
                           c1.ctypes.data_as(u64p)))
    return c0, c1, n.value, mods, sc.value


def pack_key(key, stream=None):
    out = C.create_string_buffer(lib.gf_key_packed_bytes(key.h))
    check(lib.gf_key_pack(key.h, out, len(out), stream.s if stream else None))
    return out.raw


def unpack_key(ctx, buf, stream=None):
    k = Key.__new__(Key)
    k.ctx = ctx
    k.h = lib.gf_key_unpack(ctx.h, buf, len(buf), stream.s if stream else None)
    if not k.h:
        raise GpuFheError(lib.gf_last_error().decode())
    return k
