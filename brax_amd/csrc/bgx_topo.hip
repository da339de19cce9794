// The kernels and launch table of one topology.  build.py compiles this file
// once per topology in topologies.inc, with -DBGX_TOPO=<topology struct>;
// bgx.hip reaches each object's table through the host function
// bgx_ops_<struct>().
#include "bgx_kernels.h"
#include "topologies.inc"

#ifndef BGX_TOPO
#error "compile with -DBGX_TOPO=<topology struct from topologies.inc>"
#endif

#define BGX_OPS_SYM2(T) bgx_ops_##T
#define BGX_OPS_SYM(T) BGX_OPS_SYM2(T)

const bgx::TopoOps* BGX_OPS_SYM(BGX_TOPO)() {
  static const bgx::TopoOps ops = bgx::Ops<BGX_TOPO>::table();
  return &ops;
}
