// Device-side face exchange between blocks: the on-GPU form of the
// reference's MPI ghost-cell exchange
//   communication3d / commun3d      src/Parallel/mod_comm.f90:249-290
//   communication_3dv (2 layers)    src/Parallel/mod_comm.f90:376-402
//   face types with swapped/reversed neighbour axes
//       src/Parallel/mod_mpi_types_two_sided.f90:47-197, 331-512
//       src/Grid/mod_grid_directions.f90:16-40
// A periodic face is an interface of a block with itself.
//
// All faces of all blocks go in one launch.  The host resolves every face
// (orientation included) into a FaceTask: a 3-D index box with a signed
// element stride per box axis on the destination (ghost layers) and on the
// source (the neighbour's interior layers), box axis 0 being the
// destination's contiguous one.  Workgroups map to (task, first point)
// through a chunk table; array v of the destination takes array vmap[v] of
// the source (the same quantity in the source block's stored axis order).
// Only face ghosts are written (tangential extent of
// the face), as the reference's face types do; edges and corners are never
// read by the kernels.  Sources are interior points only, so a launch never
// reads what it writes.
#include "mgpu_internal.h"

namespace mgpu {

namespace {

template <class T>
__global__ __launch_bounds__(kIfaceThreads) void k_iface(IfaceArgsT<T> a) {
  const IfaceChunk ch = a.chunks[blockIdx.x];
  const FaceTask &t = a.tasks[ch.task];
  const int e = ch.start + (int)threadIdx.x;
  const int cnt = t.n[0] * t.n[1] * t.n[2];
  if (e >= cnt) return;
  const int a0 = e % t.n[0];
  const int r = e / t.n[0];
  const int a1 = r % t.n[1];
  const int a2 = r / t.n[1];
  const idx_t pd = (idx_t)(t.dst_base + a0 * t.dst_s[0] + a1 * t.dst_s[1] + a2 * t.dst_s[2]);
  const idx_t ps = (idx_t)(t.src_base + a0 * t.src_s[0] + a1 * t.src_s[1] + a2 * t.src_s[2]);
  T *const *dst = a.ptrs + (size_t)t.dst * a.nvar;
  T *const *src = a.ptrs + (size_t)t.src * a.nvar;
  for (int v = 0; v < a.nvar; v++) stg(dst[v], pd, ldg(src[t.vmap[v]], ps));
}

}  // namespace

hipError_t launch_iface(const IfaceArgs &a, int nchunks, hipStream_t s) {
  if (nchunks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iface<double>, dim3(nchunks), dim3(kIfaceThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_iface(const IfaceArgs32 &a, int nchunks, hipStream_t s) {
  if (nchunks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_iface<float>, dim3(nchunks), dim3(kIfaceThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace mgpu
