"""C ABI and model-table layout checks (CPU only: no compute calls)."""

import ctypes
import os
import re

import numpy as np
import pytest

from brax_amd import codegen
from brax_amd import systems as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "brax_gen.h")
LAYOUT = os.path.join(ROOT, "brax_amd", "csrc", "bgx_layout.h")


def _declared_functions():
  text = open(HEADER).read()
  return sorted(set(re.findall(r"\b(bgx_[a-z_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
  from brax_amd import _lib
  return _lib.load()


def test_library_exports_every_declared_symbol(lib):
  names = _declared_functions()
  assert len(names) >= 12
  for n in names:
    assert hasattr(lib, n), n


def test_python_binding_covers_header(lib):
  from brax_amd import _lib
  assert set(_declared_functions()) == set(_lib.SIGNATURES)


def test_registered_topologies_compiled(lib):
  from brax_amd import _lib
  compiled = _lib.topologies()
  for sys, per_env in codegen.registered_systems():
    t = codegen.Topology(sys, per_env=per_env)
    assert t.signature() in compiled
    tid, _, dims = compiled[t.signature()]
    assert dims[:7] == [t.L, t.nq, t.nv, t.nu, len(t.lim_d), t.ncand, t.slots]
    assert dims[7] == t.size == codegen.pack_model(sys, t).size
    assert lib.bgx_topology_per_env(tid) == int(per_env)


def test_error_paths_without_gpu(lib):
  dims = (ctypes.c_int32 * 8)()
  assert lib.bgx_topology_dims(99, dims) == 2
  assert b"unknown topology" in lib.bgx_last_error()
  out = ctypes.c_void_p()
  table = np.zeros(3, np.float32)
  assert lib.bgx_model_create(0, table.ctypes.data, 3, ctypes.byref(out)) == 1
  assert lib.bgx_step(None, 1, 1, None, None, None, None, None, None, None,
                      None, None) == 1
  assert lib.bgx_model_set_launch(None, 0, 16) == 1
  assert lib.bgx_model_get_launch(None, 0) == -1


def test_layout_constants_match_header():
  text = open(LAYOUT).read()
  consts = dict((k, int(v)) for k, v in re.findall(r"\b([A-Z_]+) = (\d+)", text))
  assert consts["LINK_STRIDE"] == codegen.LINK_STRIDE
  assert consts["DOF_STRIDE"] == codegen.DOF_STRIDE
  assert consts["ACT_STRIDE"] == codegen.ACT_STRIDE
  assert consts["CAND_STRIDE"] == codegen.CAND_STRIDE
  sys = S.quadruped(0)
  t = codegen.Topology(sys)
  m = codegen.pack_model(sys, t)
  # spot-check records against the System arrays
  o = t.off_link + 3 * codegen.LINK_STRIDE
  assert m[o + consts["LK_MASS"]] == np.float32(sys.link.inertia.mass[3])
  o = t.off_dof + 7 * codegen.DOF_STRIDE
  assert m[o + consts["DF_HI"]] == np.float32(sys.dof.limit[1][7])
  o = t.off_act + 2 * codegen.ACT_STRIDE
  assert m[o + consts["AC_GEAR"]] == 150.0
  o = t.off_cand + 1 * codegen.CAND_STRIDE
  assert m[o + consts["CD_SIGN"]] == 1.0 and m[o + codegen.CAND_STRIDE + consts["CD_SIGN"]] == -1.0


def test_topology_tables():
  t = codegen.Topology(S.quadruped(0))
  assert t.ltype == [0] + [1] * 8
  assert t.dof_parent[:7] == [-1, 0, 1, 2, 3, 4, 5]
  assert t.dof_parent[7] == 6 and t.dof_parent[8] == 5
  assert t.dof_anc[7] == 0b1111111 and t.dof_anc[8] == 0b0111111
  assert t.ncand == 17 and t.cand_link[0] == 0


def test_unsupported_systems_raise():
  sys = S.quadruped(0)
  sys2 = sys.replace(enable_fluid=True)
  with pytest.raises(NotImplementedError):
    codegen.Topology(sys2)
