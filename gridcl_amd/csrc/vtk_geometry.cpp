// VTU output of a device geometry (tetraloma/src/vtk.rs:108-115, 869-927: VtkWriter::write of one
// rank): the nodes as equispaced nodes (converted on the device when a block has GLL nodes) with
// the origin added in f64, the dofmaps and the attached point and cell data copied to the host,
// then the host writer. Float32 points for an FP32 geometry, Float64 for FP64.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/tetraloma_geom.h"
#include "handles.hpp"
#include "vtk_internal.hpp"

namespace tg::vtk {
int write_file(const char* path, int gdim, const double* nodes, uint64_t num_nodes,
               const std::vector<Block>& blocks, uint32_t options, const tg_vtu_array* point_data,
               int32_t num_point_data, const tg_vtu_array* cell_data, int32_t num_cell_data,
               uint64_t* pyramids);
}

namespace {

const size_t kTypeSize[] = {4, 8, 4, 4, 8, 1};

// Copies device data arrays to host buffers and points the host descriptors at them.
int download(const tg_vtu_array* d, int32_t n, std::vector<std::vector<char>>& bufs,
             std::vector<tg_vtu_array>& host, hipStream_t s) {
  for (int32_t i = 0; i < n; ++i) {
    if (d[i].type < 0 || d[i].type > 5 || !d[i].name)
      return tg::fail(TG_ERR_INVALID_ARGUMENT, "bad data array type or name");
    bufs.emplace_back(d[i].count * kTypeSize[d[i].type]);
    if (d[i].count) {
      if (!d[i].values) return tg::fail(TG_ERR_INVALID_ARGUMENT, "null data values");
      const hipError_t e = hipMemcpyAsync(bufs.back().data(), d[i].values, bufs.back().size(),
                                          hipMemcpyDeviceToHost, s);
      if (e != hipSuccess) return tg::hip_fail(e, "VTU data download");
    }
    tg_vtu_array h = d[i];
    h.values = bufs.back().data();
    host.push_back(h);
  }
  return TG_OK;
}

}  // namespace

extern "C" {

int tg_vtu_write_geometry(const tg_geometry* g, const char* path, uint32_t options,
                          const tg_vtu_array* point_data, int32_t num_point_data,
                          const tg_vtu_array* cell_data, int32_t num_cell_data,
                          uint64_t* pyramids, void* stream) {
  if (!g || !path || num_point_data < 0 || num_cell_data < 0 ||
      (num_point_data && !point_data) || (num_cell_data && !cell_data))
    return tg::fail(TG_ERR_INVALID_ARGUMENT, "null or negative argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool f32 = g->dtype == TG_F32;
  const size_t elem = f32 ? sizeof(float) : sizeof(double);
  const size_t n = g->num_nodes * g->gdim;
  bool gll = false;
  for (const tg::Block& b : g->blocks) gll = gll || b.family == tg::kGll;
  // The equispaced nodes (mesh.rs:227-241): converted on the device if any block is GLL.
  const void* src = g->nodes;
  void* conv = nullptr;
  int rc = TG_OK;
  if (gll && n) {
    hipError_t e = tg::work_alloc(&conv, n * elem, s);
    if (e != hipSuccess) return tg::hip_fail(e, "VTU equispaced nodes");
    rc = tg_equispaced_nodes(g, conv, s);
    src = conv;
  }
  std::vector<char> raw(n * elem);
  std::vector<std::vector<uint32_t>> dofmaps(g->blocks.size());
  std::vector<std::vector<char>> bufs;
  std::vector<tg_vtu_array> pd, cd;
  bufs.reserve(num_point_data + num_cell_data);
  if (rc == TG_OK && n) {
    const hipError_t e = hipMemcpyAsync(raw.data(), src, raw.size(), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) rc = tg::hip_fail(e, "VTU node download");
  }
  for (size_t b = 0; rc == TG_OK && b < g->blocks.size(); ++b) {
    const tg::Block& blk = g->blocks[b];
    dofmaps[b].resize(blk.num_cells * blk.k);
    if (!dofmaps[b].empty()) {
      const hipError_t e = hipMemcpyAsync(dofmaps[b].data(), blk.dofmap,
                                          dofmaps[b].size() * sizeof(uint32_t),
                                          hipMemcpyDeviceToHost, s);
      if (e != hipSuccess) rc = tg::hip_fail(e, "VTU dofmap download");
    }
  }
  if (rc == TG_OK) rc = download(point_data, num_point_data, bufs, pd, s);
  if (rc == TG_OK) rc = download(cell_data, num_cell_data, bufs, cd, s);
  if (conv) {
    const hipError_t e = tg::work_free(conv, s);
    if (rc == TG_OK && e != hipSuccess) rc = tg::hip_fail(e, "VTU equispaced nodes");
  }
  const hipError_t e = hipStreamSynchronize(s);
  if (rc != TG_OK) return rc;
  if (e != hipSuccess) return tg::hip_fail(e, "VTU download");
  // Widened, with the origin added (mesh.rs:234-240).
  std::vector<double> nodes(n);
  for (size_t i = 0; i < n; ++i) {
    double x;
    if (f32) {
      float v;
      std::memcpy(&v, &raw[i * 4], 4);
      x = v;
    } else {
      std::memcpy(&x, &raw[i * 8], 8);
    }
    nodes[i] = x + g->origin[i % g->gdim];
  }
  std::vector<tg::vtk::Block> blocks;
  for (size_t b = 0; b < g->blocks.size(); ++b)
    blocks.push_back({g->blocks[b].cell, g->blocks[b].degree, dofmaps[b].data(),
                      dofmaps[b].size()});
  return tg::vtk::write_file(path, g->gdim, nodes.data(), g->num_nodes, blocks,
                             options | (f32 ? TG_VTU_F32_POINTS : 0u), pd.data(), num_point_data,
                             cd.data(), num_cell_data, pyramids);
}

}  // extern "C"
