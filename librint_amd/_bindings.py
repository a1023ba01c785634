"""ctypes bindings of librint_amd/lib/librint.so (C ABI in include/librint.h).

Mirrors the reference's loader (reference python/librint/_bindings.py:9-179):
the LIBRINT_SO environment variable overrides the bundled library path.
Loading fails loudly: there is no fallback implementation.
"""
import ctypes
import os

current_dir = os.path.dirname(os.path.abspath(__file__))
so_path = os.environ.get("LIBRINT_SO", os.path.join(current_dir, "lib", "librint.so"))

if not os.path.exists(so_path):
    raise ImportError(
        "librint_amd: native library %s is missing; build it with "
        "`python -c 'import __graft_entry__ as g; g.build()'` or `python librint_amd/build.py`" % so_path)

library = ctypes.CDLL(so_path)

_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_sz = ctypes.c_size_t
_mol = (_ip, _sz, _ip, _sz, _dp, _sz)

library.free_c.argtypes = (_dp, _sz)
library.free_c.restype = None
library.int1e_c.argtypes = _mol + (ctypes.c_int, ctypes.c_int)
library.int1e_c.restype = _dp
library.int2e_c.argtypes = _mol + (ctypes.c_int,)
library.int2e_c.restype = _dp
library.density_c.argtypes = _mol + (_sz, ctypes.c_int, ctypes.c_double)
library.density_c.restype = _dp
library.energy_c.argtypes = _mol + (_dp, _sz)
library.energy_c.restype = ctypes.c_double
library.scf_c.argtypes = _mol + (_sz, ctypes.c_int, ctypes.c_double)
library.scf_c.restype = ctypes.c_double
library.dS_u.argtypes = _mol
library.dS_u.restype = _dp
for _name in ("grad_c", "dS_c", "dHcore_c", "dR_c", "danalytical_c", "denergy_c"):
    getattr(library, _name).argtypes = _mol + (_dp, _sz)
    getattr(library, _name).restype = _dp

library.fock_c.argtypes = _mol + (_dp, _sz, ctypes.c_int, ctypes.c_int, ctypes.c_double)
library.fock_c.restype = _dp
library.librint_ctx_create.argtypes = _mol
library.librint_ctx_create.restype = ctypes.c_void_p
library.librint_ctx_destroy.argtypes = (ctypes.c_void_p,)
library.librint_ctx_destroy.restype = None
library.librint_ctx_nao.argtypes = (ctypes.c_void_p,)
library.librint_ctx_nao.restype = ctypes.c_int
library.librint_ctx_set_schwarz.argtypes = (ctypes.c_void_p, ctypes.c_double)
library.librint_ctx_set_schwarz.restype = ctypes.c_int
library.librint_ctx_set_shard.argtypes = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int)
library.librint_ctx_set_shard.restype = ctypes.c_int
library.librint_fock.argtypes = (ctypes.c_void_p, _dp, _dp)
library.librint_fock.restype = ctypes.c_int
library.librint_fock_device.argtypes = (ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)
library.librint_fock_device.restype = ctypes.c_int
library.librint_fock_device_noalloc.argtypes = (ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p)
library.librint_fock_device_noalloc.restype = ctypes.c_int
library.librint_int2e_shells.argtypes = (ctypes.c_void_p, _ip, ctypes.c_int, _dp)
library.librint_int2e_shells.restype = _sz
library.librint_schwarz.argtypes = (ctypes.c_void_p, _dp)
library.librint_schwarz.restype = ctypes.c_int
library.librint_ctx_stats.argtypes = (ctypes.c_void_p, _dp, ctypes.c_int)
library.librint_ctx_stats.restype = ctypes.c_int
library.librint_ctx_class_stats.argtypes = (ctypes.c_void_p, _dp, ctypes.c_int)
library.librint_ctx_class_stats.restype = ctypes.c_int
library.librint_ctx_set_precision.argtypes = (ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double)
library.librint_ctx_set_precision.restype = ctypes.c_int
library.librint_ctx_flag_counts.argtypes = (ctypes.c_void_p, _dp)
library.librint_ctx_flag_counts.restype = ctypes.c_int
library.librint_scf.argtypes = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                ctypes.c_int, _dp, _ip)
library.librint_scf.restype = ctypes.c_double
library.librint_ctx_gradient.argtypes = (ctypes.c_void_p, _dp, _dp, _dp, _dp)
library.librint_ctx_gradient.restype = ctypes.c_int
library.grad_nuc_c.argtypes = _mol + (_dp, _sz)
library.grad_nuc_c.restype = _dp
library.librint_scf_ex.argtypes = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, _dp, _ip)
library.librint_scf_ex.restype = ctypes.c_double
_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
library.librint_mlp_gemm.argtypes = (_i, _i, _i, _vp, _i, _i, _vp, _i, _i, _vp, _i, _vp, _i, _i, _vp, _i, _vp)
library.librint_mlp_relu_bwd.argtypes = (_sz, _vp, _vp, _vp, _vp, _vp)
library.librint_mlp_colsum.argtypes = (_i, _i, _vp, _i, _vp, _i, _vp)
library.librint_mlp_mse.argtypes = (_i, _i, _vp, _i, _vp, _vp, _i, _f, _vp, _vp, _vp, _vp)
library.librint_mlp_adam.argtypes = (_sz, _vp, _vp, _vp, _vp, _f, _f, _f, _f, _i, _vp, _vp)
library.librint_mlp_to_bf16.argtypes = (_sz, _vp, _vp, _vp)
for _name in ("librint_mlp_gemm", "librint_mlp_relu_bwd", "librint_mlp_colsum", "librint_mlp_mse",
              "librint_mlp_adam", "librint_mlp_to_bf16"):
    getattr(library, _name).restype = ctypes.c_int
library.librint_last_error.argtypes = ()
library.librint_last_error.restype = ctypes.c_char_p

# every symbol include/librint.h declares
EXPORTED = (
    "free_c", "int1e_c", "int2e_c", "density_c", "energy_c", "scf_c", "dS_u", "grad_c", "dS_c",
    "dHcore_c", "dR_c", "danalytical_c", "denergy_c", "fock_c", "librint_ctx_create",
    "librint_ctx_destroy", "librint_ctx_nao", "librint_ctx_set_schwarz", "librint_fock",
    "librint_fock_device", "librint_int2e_shells", "librint_schwarz", "librint_ctx_stats",
    "librint_ctx_class_stats", "librint_ctx_set_precision", "librint_ctx_flag_counts",
    "librint_scf", "librint_last_error", "librint_ctx_gradient", "grad_nuc_c", "librint_scf_ex",
    "librint_mlp_gemm", "librint_mlp_relu_bwd", "librint_mlp_colsum", "librint_mlp_mse", "librint_mlp_adam",
    "librint_mlp_to_bf16", "librint_ctx_set_shard", "librint_fock_device_noalloc",
)
