"""MI355X (gfx950) Barnes-Hut gravity: Python host bindings over the C ABI (include/ryoanji_amd.h).

The compute path is the HIP library ``libryoanji_amd.so`` built in this directory by ``build()`` in
``__graft_entry__.py`` (``make -C ryoanji_amd/csrc``). Nothing here falls back to a CPU path: if the
library is missing, importing ``lib()`` raises.

Device memory is passed as raw pointers; helpers accept torch tensors on the GPU (``.data_ptr()``)
and use torch only as an allocator and for host<->device copies.
"""

from __future__ import annotations

import ctypes as C
import os
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libryoanji_amd.so")

RYO_OK, RYO_ERR_ARG, RYO_ERR_HIP, RYO_ERR_UNSUPPORTED, RYO_ERR_CAPACITY, RYO_ERR_TRAVERSAL = range(6)

_P = C.c_void_p
_I = C.c_int
_U = C.c_uint
_SZ = C.c_size_t
_F = C.c_float
_D = C.c_double


class RyoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ryoanji_amd error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    """ryo_params: physics and tuning parameters of a solve."""

    _fields_ = [
        ("bucket_size", C.c_uint),
        ("theta", C.c_float),
        ("G", C.c_float),
        ("group_size", C.c_int),
        ("num_shells", C.c_int),
        ("ewald", C.c_int),
        ("ewald_settings", C.c_double * 5),
        ("init_node", C.c_int),
        ("soft_mac", C.c_int),
        ("group_split_tol", C.c_float),
        ("multipole", C.c_int),
    ]


class Request(C.Structure):
    """ryo_request: one served particle system."""

    _fields_ = [
        ("n", C.c_size_t),
        ("x", _P), ("y", _P), ("z", _P), ("m", _P), ("h", _P),
        ("limits", C.c_float * 6),
        ("periodic", C.c_int),
        ("theta", C.c_float), ("G", C.c_float), ("dt", C.c_float),
        ("steps", C.c_int),
        ("bucket_size", C.c_uint),
        ("group_size", C.c_int),
        ("arrival_ms", C.c_double),
        ("accuracy_targets", C.c_int),
        ("num_shells", C.c_int),
        ("ewald", C.c_int),
    ]


class Result(C.Structure):
    """ryo_result: timing and outcome of one served request."""

    _fields_ = [("arrival_ms", C.c_double), ("admit_ms", C.c_double), ("finish_ms", C.c_double),
                ("total_pot", C.c_double), ("status", C.c_int), ("slot", C.c_int),
                ("err_median", C.c_double), ("err_p99", C.c_double), ("err_max", C.c_double),
                ("limits", C.c_float * 6)]


# exported symbol -> (restype, argtypes)
SIGNATURES = {
    "ryo_system_leapfrog": (_I, [_P, C.POINTER(Params), _F, _I, C.POINTER(_D)]),
    "ryo_system_limits": (_I, [_P, _P]),
    "ryo_server_create": (_I, [_I, _SZ, _SZ, C.POINTER(_P)]),
    "ryo_server_destroy": (_I, [_P]),
    "ryo_server_set_batching": (_I, [_P, _SZ, _SZ, _I]),
    "ryo_server_set_order": (_I, [_P, _I, C.c_double]),
    "ryo_server_footprint": (_SZ, [_SZ]),
    "ryo_server_run": (_I, [_P, _I, C.POINTER(Request), C.POINTER(Result), _P]),
    "ryo_ctx_create": (_I, [C.POINTER(_P), _P]),
    "ryo_ctx_destroy": (_I, [_P]),
    "ryo_last_error": (C.c_char_p, [_P]),
    "ryo_ctx_sync": (_I, [_P]),
    "ryo_ctx_set_timing": (_I, [_P, _I]),
    "ryo_ctx_set_traversal_shape": (_I, [_P, _I, _I, _I]),
    "ryo_last_traverse_ms": (_D, [_P]),
    "ryo_sfc_keys_f32": (_I, [_P, _P, _P, _P, _P, _SZ, _P, _I, _I]),
    "ryo_sfc_keys_u32_f32": (_I, [_P, _P, _P, _P, _P, _SZ, _P, _I, _I]),
    "ryo_sfc_keys_u32_f64": (_I, [_P, _P, _P, _P, _P, _SZ, _P, _I, _I]),
    "ryo_sfc_keys_f64": (_I, [_P, _P, _P, _P, _P, _SZ, _P, _I, _I]),
    "ryo_sort_temp_bytes": (_SZ, [_SZ]),
    "ryo_sort_pairs_u64_i32": (_I, [_P, _P, _P, _SZ, _P, _P, _P, _SZ]),
    "ryo_sort_check": (_I, [_P, _P, _SZ]),
    "ryo_sequence_i32": (_I, [_P, _P, _SZ, C.c_int32]),
    "ryo_gather_f32": (_I, [_P, _P, _SZ, _I, _P, _P]),
    "ryo_gather_f64": (_I, [_P, _P, _SZ, _I, _P, _P]),
    "ryo_build_leaves_u64": (_I, [_P, _P, _SZ, _U, _P, _P, _P, _SZ, C.POINTER(_I)]),
    "ryo_link_octree_u64": (_I, [_P, _P, _I, _P, _P, _P, _P, _P, _P, _P]),
    "ryo_upsweep_f32": (_I, [_P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _F, _P, _P, _P, _P]),
    "ryo_upsweep_mdq_f32": (_I, [_P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _F, _P, _P, _P, _P]),
    "ryo_upsweep_f64": (_I, [_P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _F, _P, _P, _P]),
    "ryo_upsweep_mdq_f64": (_I, [_P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _F, _P, _P, _P]),
    "ryo_fixed_groups": (_I, [_P, C.c_uint32, C.c_uint32, _U, _P, C.POINTER(_I)]),
    "ryo_group_splits_f32": (_I, [_P, C.c_uint32, C.c_uint32, _P, _P, _P, _P, _P, _I, _P, _P, _U, _F, _F, _P,
                                  C.c_size_t, C.POINTER(_I)]),
    "ryo_traverse_f32": (_I, [_P, _I, _P, _P, _I] + [_P] * 10 + [_P, _P, _P, _P, _P, _F, _I, _P, _P, _P, _P, _P,
                                                                   _P]),
    "ryo_traverse_f32_noalloc": (_I, [_P, _I, _P, _P, _I] + [_P] * 10 + [_P, _P, _P, _P, _P, _F, _I, _P, _P, _P, _P,
                                                                       _P, _P]),
    "ryo_traverse_mdq_f32": (_I, [_P, _I, _P, _P, _I] + [_P] * 10 + [_P, _P, _P, _P, _P, _F, _I, _P, _P, _P, _P, _P,
                                                                   _P]),
    "ryo_traverse_f64": (_I, [_P, _I, _P, _P, _I] + [_P] * 10 + [_P, _P, _P, _P, _P, _D, _I, _P, _P, _P, _P, _P,
                                                                   _P]),
    "ryo_traverse_mdq_f64": (_I, [_P, _I, _P, _P, _I] + [_P] * 10 + [_P, _P, _P, _P, _P, _D, _I, _P, _P, _P, _P, _P,
                                                                   _P]),
    "ryo_traverse_f64_f32": (_I, [_P, _I, _P, _P, _I] + [_P] * 10 + [_P, _P, _P, _P, _P, _D, _I, _P, _P, _P, _P,
                                                                       _P, _P]),
    "ryo_traverse_mdq_f64_f32": (_I, [_P, _I, _P, _P, _I] + [_P] * 10 + [_P, _P, _P, _P, _P, _D, _I, _P, _P, _P, _P,
                                                                       _P, _P]),
    "ryo_node_counts_u64": (_I, [_P, _P, _I, _P, _SZ, _U, _P]),
    "ryo_node_ops_u64": (_I, [_P, _P, _I, _P, _U, _P, C.POINTER(_I)]),
    "ryo_rebalance_tree_u64": (_I, [_P, _P, _I, _I, _P, _P, C.POINTER(_I)]),
    "ryo_leaf_centers": (_I, [_P, _I, _P, _P, _P, _P, _P, _I, _P, _P]),
    "ryo_upsweep_centers": (_I, [_P, _I, _I, _P, _P, _P]),
    "ryo_set_mac": (_I, [_P, _I, _P, _I, _P, _F, _P]),
    "ryo_leaf_multipoles": (_I, [_P, _I, _I, _P, _P, _P, _P, _P, _I, _P, _P, _P]),
    "ryo_upsweep_multipoles": (_I, [_P, _I, _I, _I, _I, _P, _P, _P]),
    "ryo_direct_sum_range": (_I, [_P, _I, _SZ, _SZ, _SZ, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ryo_malloc": (_I, [_P, C.POINTER(_P), _SZ]),
    "ryo_free": (_I, [_P, _P]),
    "ryo_memcpy": (_I, [_P, _P, _P, _SZ]),
    "ryo_read_stats": (_I, [_P, _P]),
    "ryo_traversal_retries": (C.c_uint64, [_P]),
    "ryo_ewald_f32": (_I, [_P, _P, _P, C.c_uint32, C.c_uint32, _P, _P, _P, _P, _P, _F, _P, _P, _P, _P, _P, _P]),
    "ryo_direct_sum_f64": (_I, [_P, _P, _P, _P, _P, _P, _SZ, _P, _SZ, _D, _P, _I, _P]),
    "ryo_ewald_pair_sum_f64": (_I, [_P, _P, _P, _P, _P, _P, _SZ, _P, _SZ, _D, _D, _D, _P]),
    "ryo_default_params": (None, [C.POINTER(Params)]),
    "ryo_system_create": (_I, [_P, _SZ, C.POINTER(_P)]),
    "ryo_system_destroy": (_I, [_P]),
    "ryo_system_arrays": (_I, [_P, _P]),
    "ryo_system_load": (_I, [_P, _SZ, _P, _P, _P, _P, _P, _P, _I]),
    "ryo_system_store": (_I, [_P, _P, _P, _P, _P, _P]),
    "ryo_system_build": (_I, [_P, C.POINTER(Params)]),
    "ryo_system_gravity": (_I, [_P, C.POINTER(Params), C.POINTER(_D)]),
    "ryo_system_results": (_I, [_P, _P, _P]),
    "ryo_system_info": (_I, [_P, _P]),
    "ryo_system_tree": (_I, [_P, _P]),
    "ryo_system_phase_ms": (_I, [_P, _P]),
    "ryo_system_accuracy": (_I, [_P, _I, C.c_uint64, _P]),
    "ryo_batch_create": (_I, [_P, _SZ, _I, C.POINTER(_P)]),
    "ryo_batch_destroy": (_I, [_P]),
    "ryo_batch_load": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P]),
    "ryo_batch_set_periodic": (_I, [_P, _P]),
    "ryo_batch_solve": (_I, [_P, C.POINTER(Params), _P, _P, _P]),
    "ryo_batch_solve_graph": (_I, [_P, C.POINTER(Params), _P, _P, _P]),
    "ryo_ctx_set_traversal_waves": (_I, [_P, _I]),
    "ryo_batch_graph_stats": (_I, [_P, _P]),
    "ryo_batch_results": (_I, [_P, _P, _P]),
    "ryo_batch_offsets": (_I, [_P, _P, _P, _P, _P, _P]),
    "ryo_batch_tree": (_I, [_P, _P]),
}
# entry points that are not part of the public header (test hooks)
EXTRA_SIGNATURES = {
    "ryo_traverse_mdq_f32_stackcap": (_I, [_P, _I, _I, _P, _P, _I] + [_P] * 10 + [_P, _P, _P, _P, _P, _F, _I, _P, _P,
                                                                                _P, _P, _P, _P]),
    "ryo_traverse_f32_stackcap": (_I, [_P, _I, _I, _P, _P, _I] + [_P] * 10 + [_P, _P, _P, _P, _P, _F, _I, _P, _P,
                                                                            _P, _P, _P, _P]),
}

_lib = None


def lib() -> C.CDLL:
    """Load the gfx950 library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C ryoanji_amd/csrc` "
                           "(or __graft_entry__.build())")
    # torch ships its own HIP runtime under the soname libamdhip64.so.7, the soname the library links against
    # (/opt/rocm's). Loaded after torch, the library binds to torch's runtime; loaded first, it would bring in
    # /opt/rocm's copy, torch would then load its own beside it, and that second runtime finds no GPU. Importing
    # torch here keeps one runtime per process whatever the caller imports first.
    import torch  # noqa: F401

    L = C.CDLL(LIB_PATH)
    for name, (res, args) in {**SIGNATURES, **EXTRA_SIGNATURES}.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def default_params(**kw) -> Params:
    p = Params()
    lib().ryo_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "ewald_settings":
            for i, s in enumerate(v):
                p.ewald_settings[i] = s
        else:
            setattr(p, k, v)
    return p


def ptr(t) -> int | None:
    """Raw pointer of a torch tensor / numpy array (None passes NULL)."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def host_f32(a):
    return (C.c_float * len(a))(*[float(v) for v in a])


def host_f64(a):
    return (C.c_double * len(a))(*[float(v) for v in a])


class Context:
    """A ryo_ctx bound to a HIP stream (default: torch's current stream on the current device)."""

    def __init__(self, stream=None):
        import torch

        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self.h = _P()
        self._systems = weakref.WeakSet()  # destroyed before the context (they use its stream and scratch)
        self._check(lib().ryo_ctx_create(C.byref(self.h), _P(stream)))

    def _check(self, rc):
        if rc != RYO_OK:
            msg = lib().ryo_last_error(self.h).decode() if self.h else ""
            raise RyoError(rc, msg)
        return rc

    def call(self, name, *args):
        return self._check(getattr(lib(), name)(self.h, *args))

    def close(self):
        if self.h:
            for s in list(self._systems):
                s.close()
            lib().ryo_ctx_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        s = (C.c_uint64 * 5)()
        self.call("ryo_read_stats", s)
        return dict(zip(("sumP2P", "maxP2P", "sumM2P", "maxM2P", "maxStack"), list(s)))


class System:
    """A resident particle system (ryo_system): keys, tree and multipoles stay on the device."""

    def __init__(self, ctx: Context, capacity: int):
        self.ctx = ctx
        self.h = _P()
        if not ctx.h:
            raise RyoError(RYO_ERR_ARG, "system: the context is closed")
        ctx._check(lib().ryo_system_create(ctx.h, capacity, C.byref(self.h)))
        ctx._systems.add(self)
        self.capacity = capacity

    def _check(self, rc):
        return self.ctx._check(rc)

    def load(self, x, y, z, m, h, limits, periodic=False):
        n = int(x.numel() if hasattr(x, "numel") else x.size)
        self._check(lib().ryo_system_load(self.h, n, ptr(x), ptr(y), ptr(z), ptr(m), ptr(h), host_f32(limits),
                                          int(periodic)))
        self.n = n

    def build(self, params: Params):
        self._check(lib().ryo_system_build(self.h, C.byref(params)))
        self._mdq = params.multipole == 1

    def gravity(self, params: Params) -> float:
        pot = _D()
        self._check(lib().ryo_system_gravity(self.h, C.byref(params), C.byref(pot)))
        return pot.value

    def info(self):
        v = (C.c_int64 * 4)()
        self._check(lib().ryo_system_info(self.h, v))
        return dict(zip(("n", "numLeaves", "numNodes", "numGroups"), list(v)))

    def accuracy(self, num_targets=4096, seed=1):
        """errors of the last gravity call vs the FP64 direct sum on sampled targets"""
        v = (C.c_double * 4)()
        self._check(lib().ryo_system_accuracy(self.h, int(num_targets), int(seed), v))
        return dict(zip(("err_median", "err_p99", "err_max", "pot_rel_err"), list(v)))

    def phase_ms(self):
        v = (C.c_double * 8)()
        self._check(lib().ryo_system_phase_ms(self.h, v))
        return dict(zip(("keys", "sort", "permute", "leaves", "link", "upsweep", "traverse", "ewald"), list(v)))

    def _tensor(self, p, count, dtype):
        """Copy `count` elements at device pointer p into a new torch tensor."""
        import torch

        out = torch.empty(count, dtype=dtype, device="cuda")
        nbytes = out.element_size() * count
        if count:
            torch.cuda.current_stream().synchronize()
            _memcpy_d2d(out.data_ptr(), p, nbytes)
        return out

    def arrays(self):
        import torch

        ptrs = (_P * 8)()
        self._check(lib().ryo_system_arrays(self.h, ptrs))
        return [self._tensor(ptrs[k], self.n, torch.float32) for k in range(8)]

    def results(self):
        """(p, ax, ay, az, order) as torch tensors, in the system's SFC order."""
        import torch

        outs = (_P * 4)()
        order = _P()
        self._check(lib().ryo_system_results(self.h, outs, C.byref(order)))
        res = [self._tensor(outs[k], self.n, torch.float32) for k in range(4)]
        res.append(self._tensor(order.value, self.n, torch.int32))
        return res

    def tree(self):
        """Device tree arrays of the last build, copied into torch tensors."""
        import torch

        info = self.info()
        L, nn = info["numLeaves"], info["numNodes"]
        w = 12 if getattr(self, "_mdq", False) else 8  # values per node of the last build's multipoles
        p = (_P * 14)()
        self._check(lib().ryo_system_tree(self.h, p))
        spec = [("keys", self.n, torch.int64), ("leaves", L + 1, torch.int64), ("counts", L, torch.int32),
                ("layout", L + 1, torch.int32), ("prefixes", nn, torch.int64), ("childOffsets", nn, torch.int32),
                ("parents", max(1, (nn - 1) // 8), torch.int32), ("levelRange", 23, torch.int32),
                ("internalToLeaf", nn, torch.int32), ("leafToInternal", nn, torch.int32),
                ("centers", 4 * nn, torch.float32), ("multipoles", w * nn, torch.float32),
                ("centers64", 4 * nn, torch.float64), ("multipoles64", w * nn, torch.float64)]
        return {name: self._tensor(p[k], cnt, dt) for k, (name, cnt, dt) in enumerate(spec)}

    def close(self):
        if self.h:
            lib().ryo_system_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many particle systems solved together (ryo_batch): one pass over all of them per solve."""

    def __init__(self, ctx: Context, max_particles: int, max_systems: int):
        self.ctx = ctx
        self.h = _P()
        if not ctx.h:
            raise RyoError(RYO_ERR_ARG, "batch: the context is closed")
        ctx._check(lib().ryo_batch_create(ctx.h, int(max_particles), int(max_systems), C.byref(self.h)))
        ctx._systems.add(self)
        self.S = 0

    def _check(self, rc):
        return self.ctx._check(rc)

    def load(self, systems, periodic=None):
        """systems: objects with x, y, z, m, h arrays (numpy float32) and limits[6]; periodic: per-system flags
        (default: each system's own `periodic` attribute, else open)"""
        import numpy as np

        counts = np.array([s.n for s in systems], np.uint32)
        cat = [np.ascontiguousarray(np.concatenate([getattr(s, k) for s in systems]), np.float32)
               for k in ("x", "y", "z", "m", "h")]
        lim = np.ascontiguousarray(np.concatenate([np.asarray(s.limits, np.float32) for s in systems]), np.float32)
        self._keep = (counts, cat, lim)
        self._check(lib().ryo_batch_load(self.h, len(systems), counts.ctypes.data, *(a.ctypes.data for a in cat),
                                         lim.ctypes.data))
        self.S = len(systems)
        if periodic is None:
            periodic = [bool(getattr(s, "periodic", False)) for s in systems]
        flags = np.ascontiguousarray(periodic, np.int32)
        if flags.any():
            self._check(lib().ryo_batch_set_periodic(self.h, flags.ctypes.data))

    def solve(self, params: Params, theta=None, G=None, graph: bool = False):
        """one pass over every system; graph=True replays the pipeline as one captured HIP graph (the same
        results, bit for bit)"""
        import numpy as np

        th = None if theta is None else np.ascontiguousarray(theta, np.float32)
        g = None if G is None else np.ascontiguousarray(G, np.float32)
        pot = np.zeros(self.S, np.float64)
        fn = lib().ryo_batch_solve_graph if graph else lib().ryo_batch_solve
        self._check(fn(self.h, C.byref(params), None if th is None else th.ctypes.data,
                       None if g is None else g.ctypes.data, pot.ctypes.data))
        return pot

    def graph_stats(self):
        """{captures, replays, fallbacks, capacity}: graphs captured, replays, replays solved again exactly
        because the trees outgrew the captured arrays, leaf-slot capacity of the node arrays (the current graph's,
        which the eager solve after its first call uses too)"""
        import numpy as np

        out = np.zeros(4, np.uint64)
        self._check(lib().ryo_batch_graph_stats(self.h, out.ctypes.data))
        return dict(zip(("captures", "replays", "fallbacks", "capacity"), (int(v) for v in out)))

    def offsets(self):
        import numpy as np

        t = [np.zeros(self.S + 1, np.uint32) for _ in range(5)]
        self._check(lib().ryo_batch_offsets(self.h, *(a.ctypes.data for a in t)))
        return dict(zip(("P", "LO", "NO", "PO", "GO"), t))

    def results(self):
        """device copies (torch): x, y, z, m, h, p, ax, ay, az (packed, SFC order per system) and the order"""
        import torch

        ptrs = (_P * 9)()
        order = _P()
        self._check(lib().ryo_batch_results(self.h, ptrs, C.byref(order)))
        n = int(self.offsets()["P"][-1])
        out = [_tensor_from(ptrs[k], n, torch.float32) for k in range(9)]
        return out, _tensor_from(order.value, n, torch.int32)

    def tree(self):
        """device copies (torch) of the packed tree arrays (local indices per system)"""
        import torch

        o = self.offsets()
        slots, nodes, par = int(o["LO"][-1]), int(o["NO"][-1]), int(o["PO"][-1])
        ptrs = (_P * 14)()
        self._check(lib().ryo_batch_tree(self.h, ptrs))
        n = int(o["P"][-1])
        spec = [("keys", n, torch.int64), ("leaves", slots, torch.int64), ("layout", slots, torch.int32),
                ("counts", slots, torch.int32), ("prefixes", nodes, torch.int64), ("childOffsets", nodes, torch.int32),
                ("parents", par, torch.int32), ("internalToLeaf", nodes, torch.int32),
                ("leafToInternal", nodes, torch.int32), ("levelRange", 23 * self.S, torch.int32),
                ("centers", 4 * nodes, torch.float32), ("multipoles", 8 * nodes, torch.float32),
                ("centers64", 4 * nodes, torch.float64), ("multipoles64", 8 * nodes, torch.float64)]
        return {name: _tensor_from(ptrs[k], cnt, dt) for k, (name, cnt, dt) in enumerate(spec)}

    def close(self):
        if self.h:
            lib().ryo_batch_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _tensor_from(p, count, dtype):
    """Copy `count` elements at device pointer p into a new torch tensor."""
    import torch

    out = torch.empty(count, dtype=dtype, device="cuda")
    if count:
        _memcpy_d2d(out.data_ptr(), p, out.element_size() * count)
    return out


def _memcpy_d2d(dst, src, nbytes):
    import torch

    # a typed view over the raw pointer is not available in torch; use hipMemcpy via the HIP runtime
    _hip().hipMemcpy(_P(dst), _P(src), _SZ(nbytes), 3)  # hipMemcpyDeviceToDevice
    del torch


_hiprt = None


def _hip():
    global _hiprt
    if _hiprt is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7", "libamdhip64.so.6"):
            try:
                _hiprt = C.CDLL(name)
                break
            except OSError:
                continue
        if _hiprt is None:
            raise RuntimeError("HIP runtime library not found")
        _hiprt.hipMemcpy.restype = _I
        _hiprt.hipMemcpy.argtypes = [_P, _P, _SZ, _I]
    return _hiprt
