"""ctypes binding of libbrax_gen.so (declared in include/brax_gen.h).

The library is loaded after torch so that both use one HIP runtime: torch
ships libamdhip64.so with SONAME libamdhip64.so.7, which satisfies the
library's dependency of the same SONAME.  ``load()`` verifies that only one
HIP runtime is mapped and raises if the library is missing: there is no
fallback path.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional

import torch  # noqa: F401  (must be loaded before the library)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        "libbrax_gen.so")

_lib: Optional[ctypes.CDLL] = None

c_int, c_i32, c_i64 = ctypes.c_int, ctypes.c_int32, ctypes.c_int64
c_fp, c_vp, c_ip = ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p

# name -> (restype, argtypes)
SIGNATURES = {
    "bgx_version": (c_int, []),
    "bgx_last_error": (ctypes.c_char_p, []),
    "bgx_num_topologies": (c_int, []),
    "bgx_topology_name": (ctypes.c_char_p, [c_int]),
    "bgx_topology_signature": (ctypes.c_char_p, [c_int]),
    "bgx_topology_dims": (c_int, [c_int, ctypes.POINTER(c_i32)]),
    "bgx_topology_per_env": (c_int, [c_int]),
    "bgx_topology_tiles": (c_int, [c_int, ctypes.POINTER(c_i32)]),
    "bgx_model_create": (c_int, [c_int, c_vp, c_i64, ctypes.POINTER(c_vp)]),
    "bgx_model_destroy": (c_int, [c_vp]),
    "bgx_model_topology": (c_int, [c_vp]),
    "bgx_model_set_launch": (c_int, [c_vp, c_int, c_i32]),
    "bgx_model_get_launch": (c_int, [c_vp, c_int]),
    "bgx_init": (c_int, [c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "bgx_step": (c_int, [c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp,
                         c_vp, c_vp, c_vp, c_vp]),
    "bgx_rollout": (c_int, [c_vp, c_i32, c_i32, c_i32, c_vp, c_vp, c_vp, c_i64,
                            c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "bgx_contacts": (c_int, [c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "bgx_step_probe": (c_int, [c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, c_vp,
                               c_vp]),
    "bgx_frames_vjp": (c_int, [c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                               c_vp]),
    "bgx_adjoint_workspace_bytes": (c_i64, [c_vp, c_i32, c_i32, c_i32]),
    "bgx_adjoint_workspace_min_bytes": (c_i64, [c_vp, c_i32, c_i32, c_i32]),
    "bgx_rollout_adjoint": (c_int, [c_vp, c_i32, c_i32, c_i32, c_i32, c_vp,
                                    c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp,
                                    c_vp, c_vp, c_i64, c_vp, c_vp]),
    "bgx_batch_create": (c_int, [c_i32, c_vp, c_vp, ctypes.POINTER(c_vp)]),
    "bgx_batch_destroy": (c_int, [c_vp]),
    "bgx_batch_sizes": (c_int, [c_vp, c_vp]),
    "bgx_batch_rollout": (c_int, [c_vp, c_i32, c_i32, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp,
                                  c_vp, c_vp]),
    "bgx_adjoint_scratch_bytes": (c_i64, []),
    "bgx_rollout_adjoint_noalloc": (c_int, [c_vp, c_i32, c_i32, c_i32, c_i32, c_vp,
                                            c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp,
                                            c_vp, c_vp, c_vp]),
}


class BgxError(RuntimeError):
  pass


def _hip_runtimes_mapped():
  paths = set()
  try:
    with open("/proc/self/maps") as f:
      for line in f:
        if "libamdhip64" in line:
          paths.add(os.path.realpath(line.split()[-1]))
  except OSError:
    pass
  return paths


def load() -> ctypes.CDLL:
  global _lib
  if _lib is not None:
    return _lib
  if not os.path.exists(LIB_PATH):
    raise BgxError(
        f"{LIB_PATH} is not built; run __graft_entry__.build() (or "
        "python brax_amd/build.py) first")
  lib = ctypes.CDLL(LIB_PATH)
  for name, (res, args) in SIGNATURES.items():
    fn = getattr(lib, name)
    fn.restype, fn.argtypes = res, args
  runtimes = _hip_runtimes_mapped()
  if len(runtimes) > 1:
    raise BgxError(f"two HIP runtimes are mapped: {sorted(runtimes)}")
  _lib = lib
  return lib


def check(rc: int) -> None:
  if rc != 0:
    msg = load().bgx_last_error().decode()
    raise BgxError(f"libbrax_gen error {rc}: {msg}")


def topologies():
  """{signature: (topo_id, name, dims)} of the compiled topologies."""
  lib = load()
  out = {}
  for i in range(lib.bgx_num_topologies()):
    dims = (c_i32 * 8)()
    check(lib.bgx_topology_dims(i, dims))
    out[lib.bgx_topology_signature(i).decode()] = (
        i, lib.bgx_topology_name(i).decode(), list(dims))
  return out


def ptr(t: Optional[torch.Tensor]):
  return None if t is None else ctypes.c_void_p(t.data_ptr())
