// C-ABI entry points declared in include/tetraloma_geom.h: handle management, argument checks
// with the reference's error variants, cached basis tables, and kernel dispatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/tetraloma_geom.h"
#include "dense_mfma.hpp"
#include "geom_internal.hpp"
#include "handles.hpp"
#include "refcell.hpp"

namespace {
thread_local std::string g_last_error;
}  // namespace

namespace tg {

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

const char* cell_name(int c) {
  static const char* names[] = {"Point",       "Interval",   "Triangle", "Quadrilateral",
                                "Tetrahedron", "Hexahedron", "Prism",    "Pyramid"};
  return valid_cell(c) ? names[c] : "?";
}

int hip_fail(hipError_t e, const char* what) {
  return fail(TG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {
std::mutex g_pool_mutex;
std::map<int, hipMemPool_t> g_pools;  // kept until the process ends
}  // namespace

// The work-memory pool of the current device: stream-ordered, with an unlimited release
// threshold, so that blocks freed by one call stay mapped for the next. The default pool's
// threshold is 0: every synchronisation handed its free blocks back to the driver and the next
// call mapped them again (about 0.1 ms for the 201 MB lattice field of a Q3 64^3 conversion).
static hipError_t work_pool(hipMemPool_t* out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(g_pool_mutex);
  auto it = g_pools.find(dev);
  if (it != g_pools.end()) {
    *out = it->second;
    return hipSuccess;
  }
  hipMemPoolProps props{};
  props.allocType = hipMemAllocationTypePinned;
  props.handleTypes = hipMemHandleTypeNone;
  props.location.type = hipMemLocationTypeDevice;
  props.location.id = dev;
  hipMemPool_t pool = nullptr;
  e = hipMemPoolCreate(&pool, &props);
  if (e != hipSuccess) return e;
  uint64_t keep = UINT64_MAX;
  e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  if (e != hipSuccess) {
    (void)hipMemPoolDestroy(pool);
    return e;
  }
  g_pools.emplace(dev, pool);
  *out = pool;
  return hipSuccess;
}

hipError_t work_alloc(void** p, size_t bytes, hipStream_t s) {
  hipMemPool_t pool = nullptr;
  const hipError_t e = work_pool(&pool);
  if (e != hipSuccess) return e;
  return hipMallocFromPoolAsync(p, bytes, pool, s);
}

hipError_t work_free(void* p, hipStream_t s) { return hipFreeAsync(p, s); }

DeviceTables::~DeviceTables() {
  if (tab) (void)hipFree(tab);
  if (tab64) (void)hipFree(tab64);
  if (ref_normals) (void)hipFree(ref_normals);
  if (sf_tab) (void)hipFree(sf_tab);
  if (sf_tab64) (void)hipFree(sf_tab64);
  if (sf_lex) (void)hipFree(sf_lex);
  if (sf_mfma) (void)hipFree(sf_mfma);
  if (dense_mfma) (void)hipFree(dense_mfma);
}

}  // namespace tg

tg_pointset::~tg_pointset() {
  if (weights_device) (void)hipFree(weights_device);
}

using tg::Block;
using tg::cell_name;
using tg::DeviceTables;
using tg::fail;
using tg::hip_fail;

namespace {

// Host tables of (cell, degree, family) at the points, interleaved [q][k][4] in f64.
int host_tables(const tg_pointset* p, unsigned degree, int family, std::vector<double>& inter,
                size_t* k_out) {
  tg::LagrangeElement e;
  std::string err;
  if (!e.init(p->cell, degree, static_cast<tg::Family>(family), &err)) {
    if (degree == 0 || err == "unsupported degree")
      return fail(TG_ERR_UNSUPPORTED_DEGREE, std::string("UnsupportedDegree { cell: ") +
                                                 cell_name(p->cell) +
                                                 ", degree: " + std::to_string(degree) + " }");
    return fail(TG_ERR_UNSUPPORTED_NODE_FAMILY,
                std::string("UnsupportedNodeFamily { cell: ") + cell_name(p->cell) + " }: " + err);
  }
  const size_t k = e.dim(), nq = p->nq;
  std::vector<double> t;
  e.tabulate(p->points.data(), nq, 1, t);
  inter.assign(nq * k * 4, 0.0);
  for (size_t q = 0; q < nq; ++q)
    for (size_t i = 0; i < k; ++i) {
      inter[(q * k + i) * 4] = t[q * k + i];
      for (int d = 0; d < p->tdim; ++d) inter[(q * k + i) * 4 + 1 + d] = t[((1 + d) * nq + q) * k + i];
    }
  *k_out = k;
  return TG_OK;
}

template <typename T>
int upload(const std::vector<double>& src, void** dst) {
  std::vector<T> narrowed(src.begin(), src.end());
  const size_t bytes = std::max<size_t>(narrowed.size(), 1) * sizeof(T);
  hipError_t e = hipMalloc(dst, bytes);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(tables)");
  if (!narrowed.empty()) {
    e = hipMemcpy(*dst, narrowed.data(), narrowed.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(tables)");
  }
  return TG_OK;
}

// Sum-factorisation tables of a hexahedron of `degree` at the points of p, when the point set is
// an m x m x m tensor product of one 1D set u (point i = (u[i % m], u[i / m % m], u[i / m^2]))
// and the element's nodes lie on an n1^3 lattice: tab1d [2][m][n1] (1D Lagrange values, then
// derivatives, on the sorted 1D node positions) and lex [k]. Returns false when not applicable.
bool sumfact_tables(const tg_pointset* p, unsigned degree, int family, std::vector<double>& tab1d,
                    std::vector<uint32_t>& lex, int* n1_out, int* m_out) {
  if (p->cell != tg::kHexahedron || p->tdim != 3) return false;
  std::vector<double> nodes;
  if (!tg::lagrange_points(tg::kHexahedron, degree, static_cast<tg::Family>(family), nodes))
    return false;
  const size_t k = nodes.size() / 3;
  std::vector<double> t1;
  for (size_t i = 0; i < k; ++i) t1.push_back(nodes[i * 3]);
  std::sort(t1.begin(), t1.end());
  t1.erase(std::unique(t1.begin(), t1.end(), [](double a, double b) { return std::fabs(a - b) < 1e-12; }),
           t1.end());
  const int n1 = static_cast<int>(t1.size());
  if (n1 < 2 || n1 > 4 || static_cast<size_t>(n1 * n1 * n1) != k) return false;
  auto index_of = [&](double v) {
    for (int i = 0; i < n1; ++i)
      if (std::fabs(t1[i] - v) < 1e-12) return i;
    return -1;
  };
  lex.assign(k, 0);
  std::vector<char> seen(k, 0);
  for (size_t i = 0; i < k; ++i) {
    const int a = index_of(nodes[i * 3]), b = index_of(nodes[i * 3 + 1]), c = index_of(nodes[i * 3 + 2]);
    if (a < 0 || b < 0 || c < 0) return false;
    const uint32_t l = static_cast<uint32_t>((a * n1 + b) * n1 + c);
    if (seen[l]) return false;
    seen[l] = 1;
    lex[i] = l;
  }
  const size_t nq = p->nq;
  int m = 0;
  while (static_cast<size_t>(m * m * m) < nq) ++m;
  if (m < 1 || m > 4 || static_cast<size_t>(m * m * m) != nq) return false;
  std::vector<double> u(m);
  for (int i = 0; i < m; ++i) u[i] = p->points[i * 3];
  for (size_t i = 0; i < nq; ++i) {
    const double want[3] = {u[i % m], u[(i / m) % m], u[i / (m * m)]};
    for (int d = 0; d < 3; ++d)
      if (std::fabs(p->points[i * 3 + d] - want[d]) > 1e-14) return false;
  }
  // 1D Lagrange polynomial n on the nodes t1 and its derivative, in the product form.
  tab1d.assign(2 * m * n1, 0.0);
  for (int q = 0; q < m; ++q)
    for (int n = 0; n < n1; ++n) {
      double v = 1.0, dv = 0.0;
      for (int j = 0; j < n1; ++j) {
        if (j == n) continue;
        const double f = (u[q] - t1[j]) / (t1[n] - t1[j]);
        dv = dv * f + v / (t1[n] - t1[j]);
        v *= f;
      }
      tab1d[q * n1 + n] = v;
      tab1d[(m + q) * n1 + n] = dv;
    }
  *n1_out = n1;
  *m_out = m;
  return true;
}

}  // namespace

namespace tg {

// The cached device tables of (degree, family, dtype), computed and uploaded on first use
// (point_set.rs:355-365).
int device_tables(tg_pointset* p, unsigned degree, int family, int dtype, DeviceTables** out) {
  auto key = std::make_tuple(degree, family, dtype);
  auto it = p->cache.find(key);
  if (it != p->cache.end()) {
    *out = it->second.get();
    return TG_OK;
  }
  std::vector<double> inter;
  size_t k = 0;
  int rc = host_tables(p, degree, family, inter, &k);
  if (rc) return rc;
  auto t = std::make_unique<DeviceTables>();
  rc = dtype == TG_F32 ? upload<float>(inter, &t->tab) : upload<double>(inter, &t->tab);
  if (rc) return rc;
  rc = upload<double>(inter, reinterpret_cast<void**>(&t->tab64));
  if (rc) return rc;
  if (p->has_normals) {
    rc = dtype == TG_F32 ? upload<float>(p->normals, &t->ref_normals)
                         : upload<double>(p->normals, &t->ref_normals);
    if (rc) return rc;
  }
  const uint32_t ku = static_cast<uint32_t>(k);
  if (p->tdim == 3 && p->nq <= tg::kDenseMfmaMaxPoints &&
      (dtype == TG_F32 ? tg::has_dense_mfma(3, 3, ku) : tg::has_dense_mfma_fp64(3, 3, ku))) {
    // The table in matrix-core operand order (dense_mfma.hip): block qb, kind v, step s, lane l
    // holds T_v[16 qb + l % 16][4 s + l / 16].
    const uint32_t kk = static_cast<uint32_t>(k), nq = static_cast<uint32_t>(p->nq);
    const uint32_t S = (kk + 3) / 4, nqb = (nq + 15) / 16;
    std::vector<double> a(tg::dense_mfma_table_elems(kk, nq), 0.0);
    for (uint32_t qb = 0; qb < nqb; ++qb)
      for (uint32_t v = 0; v < 4; ++v)
        for (uint32_t st = 0; st < S; ++st)
          for (uint32_t l = 0; l < 64; ++l) {
            const uint32_t q = qb * 16 + (l & 15u), node = 4 * st + (l >> 4);
            if (q < nq && node < kk)
              a[((qb * 4 + v) * S + st) * 64 + l] = inter[(q * kk + node) * 4 + v];
          }
    rc = dtype == TG_F32 ? upload<float>(a, &t->dense_mfma) : upload<double>(a, &t->dense_mfma);
    if (rc) return rc;
  }
  std::vector<double> tab1d;
  std::vector<uint32_t> lex;
  int n1 = 0, m = 0;
  if (sumfact_tables(p, degree, family, tab1d, lex, &n1, &m)) {
    rc = dtype == TG_F32 ? upload<float>(tab1d, &t->sf_tab) : upload<double>(tab1d, &t->sf_tab);
    if (rc) return rc;
    rc = upload<double>(tab1d, reinterpret_cast<void**>(&t->sf_tab64));
    if (rc) return rc;
    // lex, then its inverse (the A.5 index at each lexicographic position).
    const size_t k = lex.size();
    std::vector<uint32_t> both(2 * k);
    for (size_t i = 0; i < k; ++i) {
      both[i] = lex[i];
      both[k + lex[i]] = static_cast<uint32_t>(i);
    }
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&t->sf_lex), both.size() * sizeof(uint32_t));
    if (e == hipSuccess)
      e = hipMemcpy(t->sf_lex, both.data(), both.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "sum-factorisation tables");
    if (dtype == TG_F32 && tg::has_sumfact_mfma(n1, m)) {
      // A operand of lane l at b for kind kd: the product of the 1D tables of (p, a) and (q, b),
      // (p, q) = (l % 4, l % 16 / 4), a = l / 16, formed in f64 and rounded once.
      std::vector<double> t2(3 * n1 * 64, 0.0);
      for (int kd = 0; kd < 3; ++kd)
        for (int b = 0; b < n1; ++b)
          for (int l = 0; l < 64; ++l) {
            const int pp = l & 3, qq = (l & 15) >> 2, a = l >> 4;
            if (pp >= m || qq >= m || a >= n1) continue;
            const double fa = tab1d[((kd == 1 ? m : 0) + pp) * n1 + a];
            const double fb = tab1d[((kd == 2 ? m : 0) + qq) * n1 + b];
            t2[(kd * n1 + b) * 64 + l] = fa * fb;
          }
      rc = upload<float>(t2, &t->sf_mfma);
      if (rc) return rc;
    }
    t->sf_n1 = n1;
    t->sf_m = m;
  }
  *out = t.get();
  p->cache.emplace(key, std::move(t));
  return TG_OK;
}

int resident_tables(tg_pointset* p, unsigned degree, int family, int dtype, DeviceTables** out) {
  auto it = p->cache.find(std::make_tuple(degree, family, dtype));
  if (it == p->cache.end())
    return fail(TG_ERR_INVALID_ARGUMENT,
                "the point set's device tables for this degree, node family and precision are "
                "not resident (an evaluation or tg_measure_block with them uploads them)");
  *out = it->second.get();
  return TG_OK;
}

int resident_weights(const tg_pointset* p, const double** out) {
  if (!p->weights_device)
    return fail(TG_ERR_INVALID_ARGUMENT,
                "the point set's weights are not resident (tg_measure_block uploads them)");
  *out = p->weights_device;
  return TG_OK;
}

int device_weights(tg_pointset* p, const double** out) {
  if (!p->weights_device) {
    void* d = nullptr;
    int rc = upload<double>(p->weights, &d);
    if (rc) return rc;
    p->weights_device = static_cast<double*>(d);
  }
  *out = p->weights_device;
  return TG_OK;
}

}  // namespace tg

using tg::device_tables;

namespace {

int make_pointset(tg::Cell cell, std::vector<double> pts, std::vector<double> wts,
                  std::vector<double> normals, bool has_normals, tg_pointset** out) {
  auto* p = new tg_pointset;
  p->cell = cell;
  p->tdim = tg::tdim(cell);
  p->nq = wts.size();
  p->points = std::move(pts);
  p->weights = std::move(wts);
  p->normals = std::move(normals);
  p->has_normals = has_normals;
  *out = p;
  return TG_OK;
}

bool valid_tile(uint32_t tile) {
  return tile == 1 || tile == 2 || tile == 4 || tile == 8 || tile == 16;
}

uint32_t log2u(uint32_t v) {
  uint32_t r = 0;
  while ((1u << r) < v) ++r;
  return r;
}

}  // namespace

namespace tg {

// Whether the sum-factorised kernel is the faster one for hexahedra with n1 nodes and m points
// per direction (measured on MI355X, X + J + det J, tiles of 16, 0.26-0.88 M cells):
//   FP32: degree 3 at every m; degrees 1 and 2 at m = 4, and degree 2 at m = 3 (alone the dense
//         kernel is up to 8 % faster there, but overlapped with other blocks in a plan's graph
//         the sum-factorised one gives the shorter total); otherwise the dense register kernels
//         (1.1-4x faster at m <= 2, and for degree 1 at m = 3);
//   FP64: degrees 2 and 3 at every m; degree 1 never (within 5 % either way).
static bool sumfact_pays(bool f32, int n1, int m) {
  if (n1 >= 4) return true;
  return f32 ? (m >= 4 || (n1 == 3 && m == 3)) : n1 >= 3;
}

// Whether the FP32 sum-factorised hexahedra take their second and third sums on the matrix cores
// (measured on MI355X, X + J + det J, us per launch, matrix cores / vector ALU):
//   degree 2 at 3^3 points 139 / 170 (373,248 cells), degree 3 at 4^3 points 173 / 172
//   (262,144 cells) and 690 / 700 (884,736 cells);
//   elsewhere the vector ALU is as fast or faster (degree 3 at 2^3 and 3^3 points 106 / 73 and
//   122 / 114, degree 2 at 4^3 points 255 / 237, degree 1 at 4^3 points 574 / 497).
static bool mfma_pays(int n1, int m) { return (n1 == 3 && m == 3) || (n1 == 4 && m == 4); }

// The scalars of one factor launch over block b.
template <typename T>
static tg::FactorParams<T> block_params(const tg_geometry* g, const Block& b,
                                        const tg_pointset* pts, uint32_t factors, uint32_t tile,
                                        uint32_t options) {
  tg::FactorParams<T> p{};
  p.num_cells = b.num_cells;
  p.padded_cells = (b.num_cells + tile - 1) / tile * tile;
  p.last_node = static_cast<uint32_t>(g->num_nodes - 1);
  p.nq = static_cast<uint32_t>(pts->nq);
  p.k = b.k;
  p.tile_log2 = log2u(tile);
  p.mask = factors;
  const bool f32 = sizeof(T) == 4;
  p.fixup = (f32 && !(options & TG_OPT_NO_FP64_FIXUP) && !(options & TG_OPT_NO_CELL_SHIFT)) ? 1u
                                                                                            : 0u;
  p.tau = g->tau;
  p.mapping = tg::choose_mapping(g->dtype, b.tdim, g->gdim, b.k, p.nq, tile, options, p.mask);
  p.xcd_remap = tg::choose_xcd_remap();
  const bool relative = (options & TG_OPT_POINTS_RELATIVE) != 0;
  for (int i = 0; i < 3; ++i)
    p.origin[i] = static_cast<T>(i < g->gdim && !relative ? g->origin[i] : 0.0);
  return p;
}

// Whether the factors of a block go to the sum-factorised kernel (hexahedra of gdim 3 whose
// factors it computes, tiles of 16, where it pays or is forced).
static bool takes_sumfact(const tg_geometry* g, const DeviceTables* tabs, uint32_t factors,
                          uint32_t tile, uint32_t options, bool f32) {
  const uint32_t sf_factors = TG_POINTS | TG_JACOBIAN | TG_DETJ;
  return tabs->sf_m > 0 && tile == 16 && g->gdim == 3 && !(factors & ~sf_factors) &&
         !(options & (TG_OPT_FORCE_GENERIC | TG_OPT_NO_SUMFACT)) &&
         ((options & TG_OPT_SUMFACT_ALWAYS) || sumfact_pays(f32, tabs->sf_n1, tabs->sf_m));
}

// Whether the factors of a block go to the dense matrix-core kernel (dense_mfma.hip): a 3D cell
// of a node count it has (FP32 with the cell shift, on FP32 nodes or the FP64 copy: 20 to 84
// nodes in registers and more in the runtime-k form; FP64: 20 nodes and more), the point set's
// operand-order table (at most 128 points), tiles of 16, any of X, J, det J, K and the facet
// normals (these with the point set's reference normals) and nothing else,
// a block the sum-factorised kernels do not take, unless a strategy is forced. Measured on
// MI355X, X + J + det J: P3 tetrahedra on 663,552 cells 0.120 against 0.152 ms at 14 points,
// 0.414 against 0.513 ms at 64; the cells without register kernels 2.4-45x faster than the
// runtime-K and cooperative kernels (dense_mfma.hip lists them). X alone and det J alone too: P3
// tetrahedra at 14 points 55 / 73 us against 90 / 112 us on the register kernel's lane-per-cell
// mappings (FP32; FP64 95 / 143 against 265 / 280 us), Q2 hexahedra 1.2-2.4x.
bool takes_dense_mfma(const tg_geometry* g, const Block& b, const DeviceTables* tabs,
                      uint32_t factors, uint32_t tile, uint32_t options) {
  const bool f32 = g->dtype == TG_F32;
  const uint32_t taken = TG_POINTS | TG_JACOBIAN | TG_DETJ | TG_PINV | TG_FACET_NORMALS;
  if ((f32 && (options & TG_OPT_NO_CELL_SHIFT)) || !tabs->dense_mfma || tile != 16 ||
      !(f32 ? tg::has_dense_mfma(b.tdim, g->gdim, b.k)
            : tg::has_dense_mfma_fp64(b.tdim, g->gdim, b.k)) ||
      !(factors & taken) || (factors & ~taken) ||
      ((factors & TG_FACET_NORMALS) && !tabs->ref_normals) ||
      (options & (TG_OPT_NO_MFMA | TG_OPT_FORCE_GENERIC | TG_OPT_COOPERATIVE)))
    return false;
  return !takes_sumfact(g, tabs, factors, tile, options, f32);
}

template <typename T>
int enqueue_block(const tg_geometry* g, const Block& b, tg_pointset* pts, DeviceTables* tabs,
                  uint32_t factors, uint32_t tile, uint32_t options, const tg_outputs* out,
                  hipStream_t stream) {
  const tg::FactorParams<T> p = block_params<T>(g, b, pts, factors, tile, options);
  const bool f32 = sizeof(T) == 4;
  tg::FactorOutputs o{out->points, out->jacobians, out->detj, out->pinv, out->facet_normals,
                      out->normals};
  const bool shift = f32 && !(options & TG_OPT_NO_CELL_SHIFT);
  // FP32 arithmetic on the FP64 copy of the nodes, where the geometry has one.
  const double* n64 = shift ? g->nodes64 : nullptr;
  if (takes_sumfact(g, tabs, factors, tile, options, f32)) {
    if constexpr (sizeof(T) == 4) {
      const bool mfma = tabs->sf_mfma && !(options & TG_OPT_NO_MFMA) &&
                        ((options & TG_OPT_SUMFACT_ALWAYS) || mfma_pays(tabs->sf_n1, tabs->sf_m));
      if (n64 && mfma) {
        const size_t k = static_cast<size_t>(b.k);
        hipError_t e = tg::launch_sumfact_mfma_f64_nodes(
            tabs->sf_n1, tabs->sf_m, n64, b.dofmap, tabs->sf_tab64, tabs->sf_lex + k,
            static_cast<const float*>(tabs->sf_mfma), tabs->tab64, p, o, stream);
        if (e == hipSuccess) return TG_OK;
        if (e != hipErrorNotSupported) return hip_fail(e, "sum-factorised matrix-core launch");
      }
      if (n64) {
        hipError_t e = tg::launch_sumfact_f64_nodes(
            tabs->sf_n1, tabs->sf_m, n64, b.dofmap, tabs->sf_tab64,
            tabs->sf_lex, tabs->tab64, p, o, stream);
        if (e == hipSuccess) return TG_OK;
        if (e != hipErrorNotSupported) return hip_fail(e, "sum-factorised kernel launch");
      }
      if (!n64 && mfma) {
        const size_t k = static_cast<size_t>(b.k);
        hipError_t e = tg::launch_sumfact_mfma(
            tabs->sf_n1, tabs->sf_m, shift, static_cast<const float*>(g->nodes), b.dofmap,
            static_cast<const float*>(tabs->sf_tab), tabs->sf_lex + k,
            static_cast<const float*>(tabs->sf_mfma), tabs->tab64, p, o, stream);
        if (e == hipSuccess) return TG_OK;
        if (e != hipErrorNotSupported) return hip_fail(e, "sum-factorised matrix-core launch");
      }
    }
    if (!n64) {
      hipError_t e = tg::launch_sumfact<T>(
          tabs->sf_n1, tabs->sf_m, shift, static_cast<const T*>(g->nodes), b.dofmap,
          static_cast<const T*>(tabs->sf_tab), tabs->sf_lex, tabs->tab64, p, o, stream);
      if (e == hipSuccess) return TG_OK;
      if (e != hipErrorNotSupported) return hip_fail(e, "sum-factorised kernel launch");
    }
  }
  if constexpr (sizeof(T) == 4) {
    if (takes_dense_mfma(g, b, tabs, factors, tile, options)) {
      const float* ta = static_cast<const float*>(tabs->dense_mfma);
      const float* rn = static_cast<const float*>(tabs->ref_normals);
      hipError_t e = n64 ? tg::launch_dense_mfma_f64_nodes(b.k, n64, b.dofmap, ta, tabs->tab64,
                                                           rn, p, o, stream)
                         : tg::launch_dense_mfma(b.k, static_cast<const float*>(g->nodes),
                                                 b.dofmap, ta, tabs->tab64, rn, p, o, stream);
      if (e == hipSuccess) return TG_OK;
      if (e != hipErrorNotSupported) return hip_fail(e, "dense matrix-core launch");
    }
  } else {
    if (takes_dense_mfma(g, b, tabs, factors, tile, options)) {
      hipError_t e = tg::launch_dense_mfma_fp64(
          b.k, static_cast<const double*>(g->nodes), b.dofmap,
          static_cast<const double*>(tabs->dense_mfma), tabs->tab64,
          static_cast<const double*>(tabs->ref_normals), p, o, stream);
      if (e == hipSuccess) return TG_OK;
      if (e != hipErrorNotSupported) return hip_fail(e, "dense matrix-core launch");
    }
  }
  // The cooperative strategy: forced, or by default for cells of many nodes (56 in FP32, 84 in
  // FP64: tetrahedra from degree 5 and 6) that have no register-resident kernel, where it
  // measured faster than the runtime-K kernel (P5 tetrahedra FP32 1.07 against 1.47 ms, P6 0.62
  // against 1.65 ms, P8 0.77 against 6.2 ms; FP64 P6 1.07 against 2.57 ms; P4 and FP64 P5
  // favour the runtime-K kernel). Q3 hexahedra (64 nodes) keep their register kernel: 4.9 ms
  // against 17.9 ms cooperative on 96^3 cells at 4^3 points. The reference's GPU policy takes
  // it for tetrahedra from degree 7 (FP32) and 6 (FP64), degree-policy.toml:77,86.
  if (!n64 && !(options & TG_OPT_FORCE_GENERIC) &&
      ((options & TG_OPT_COOPERATIVE) ||
       (b.k >= (f32 ? 56u : 84u) && !tg::has_direct_factor<T>(b.tdim, g->gdim, b.k, shift)))) {
    hipError_t e = tg::launch_factor_coop<T>(
        b.tdim, g->gdim, shift, static_cast<const T*>(g->nodes), b.dofmap, b.flags,
        static_cast<const T*>(tabs->tab), tabs->tab64, static_cast<const T*>(tabs->ref_normals),
        p, o, stream);
    if (e == hipSuccess) return TG_OK;
    if (e != hipErrorNotSupported) return hip_fail(e, "cooperative kernel launch");
  }
  if constexpr (sizeof(T) == 4) {
    if (n64) {
      hipError_t e = tg::launch_factor_f64_nodes(
          b.tdim, g->gdim, (options & TG_OPT_FORCE_GENERIC) != 0, n64, b.dofmap, b.flags,
          static_cast<const float*>(tabs->tab), tabs->tab64,
          static_cast<const float*>(tabs->ref_normals), p, o, stream);
      if (e == hipErrorNotSupported)
        return fail(TG_ERR_UNSUPPORTED_DEGREE, "no kernel for this (tdim, gdim)");
      if (e != hipSuccess) return hip_fail(e, "factor kernel launch");
      return TG_OK;
    }
  }
  hipError_t e = tg::launch_factor<T>(
      b.tdim, g->gdim, shift, (options & TG_OPT_FORCE_GENERIC) != 0,
      static_cast<const T*>(g->nodes), b.dofmap, b.flags, static_cast<const T*>(tabs->tab),
      tabs->tab64, static_cast<const T*>(tabs->ref_normals), p, o, stream);
  if (e == hipErrorNotSupported)
    return fail(TG_ERR_UNSUPPORTED_DEGREE, "no kernel for this (tdim, gdim)");
  if (e != hipSuccess) return hip_fail(e, "factor kernel launch");
  return TG_OK;
}

template <typename T>
bool fused_cell_measures(const tg_geometry* g, const Block& b, const DeviceTables* tabs,
                         uint32_t options) {
  const bool f32 = sizeof(T) == 4;
  // With FP64 nodes the measure takes det J from the evaluation kernels (the field path).
  if (f32 && g->nodes64 && !(options & TG_OPT_NO_CELL_SHIFT)) return false;
  if ((options & TG_OPT_FORCE_GENERIC) || takes_sumfact(g, tabs, TG_DETJ, 16, options, f32))
    return false;
  return tg::has_measure_cells<T>(b.tdim, g->gdim, b.k, f32 && !(options & TG_OPT_NO_CELL_SHIFT));
}

template <typename T>
int enqueue_cell_measures(const tg_geometry* g, const Block& b, tg_pointset* pts,
                          DeviceTables* tabs, const double* weights, uint32_t options, T* sums,
                          hipStream_t stream) {
  const tg::FactorParams<T> p = block_params<T>(g, b, pts, TG_DETJ, 16, options);
  const bool shift = sizeof(T) == 4 && !(options & TG_OPT_NO_CELL_SHIFT);
  hipError_t e = tg::launch_measure_cells<T>(b.tdim, g->gdim, shift,
                                             static_cast<const T*>(g->nodes), b.dofmap,
                                             static_cast<const T*>(tabs->tab), tabs->tab64,
                                             weights, p, sums, stream);
  if (e == hipErrorNotSupported)
    return fail(TG_ERR_UNSUPPORTED_DEGREE, "no measure kernel for this (tdim, gdim)");
  if (e != hipSuccess) return hip_fail(e, "measure kernel launch");
  return TG_OK;
}

template int enqueue_block<float>(const tg_geometry*, const Block&, tg_pointset*, DeviceTables*,
                                  uint32_t, uint32_t, uint32_t, const tg_outputs*, hipStream_t);
template int enqueue_block<double>(const tg_geometry*, const Block&, tg_pointset*, DeviceTables*,
                                   uint32_t, uint32_t, uint32_t, const tg_outputs*, hipStream_t);
template bool fused_cell_measures<float>(const tg_geometry*, const Block&, const DeviceTables*,
                                         uint32_t);
template bool fused_cell_measures<double>(const tg_geometry*, const Block&, const DeviceTables*,
                                          uint32_t);
template int enqueue_cell_measures<float>(const tg_geometry*, const Block&, tg_pointset*,
                                          DeviceTables*, const double*, uint32_t, float*,
                                          hipStream_t);
template int enqueue_cell_measures<double>(const tg_geometry*, const Block&, tg_pointset*,
                                           DeviceTables*, const double*, uint32_t, double*,
                                           hipStream_t);

}  // namespace tg

extern "C" {

const char* tg_last_error_message(void) { return g_last_error.c_str(); }
const char* tg_version(void) { return "gridcl_amd 0.1 (gfx950)"; }

uint64_t tg_field_elems(uint64_t num_cells, uint64_t nq, uint32_t len, uint32_t tile) {
  if (!valid_tile(tile)) return 0;
  return (num_cells + tile - 1) / tile * tile * nq * len;
}

uint64_t tg_num_lagrange_nodes(int32_t cell, uint32_t degree) {
  if (!tg::valid_cell(cell)) return 0;
  return tg::num_lagrange_nodes(static_cast<tg::Cell>(cell), degree);
}

int tg_lagrange_points(int32_t cell, uint32_t degree, int32_t family, double* out, size_t cap,
                       size_t* len) {
  if (!tg::valid_cell(cell) || (family != TG_EQUISPACED && family != TG_GLL))
    return fail(TG_ERR_INVALID_ARGUMENT, "bad cell or family");
  std::vector<double> pts;
  if (!tg::lagrange_points(static_cast<tg::Cell>(cell), degree, static_cast<tg::Family>(family),
                           pts)) {
    if (degree == 0)
      return fail(TG_ERR_UNSUPPORTED_DEGREE, std::string("UnsupportedDegree { cell: ") +
                                                 cell_name(cell) + ", degree: 0 }");
    return fail(TG_ERR_UNSUPPORTED_NODE_FAMILY,
                std::string("UnsupportedNodeFamily { cell: ") + cell_name(cell) + " }");
  }
  if (len) *len = pts.size();
  if (out) {
    if (cap < pts.size()) return fail(TG_ERR_LENGTH_MISMATCH, "output capacity too small");
    std::memcpy(out, pts.data(), pts.size() * sizeof(double));
  }
  return TG_OK;
}

int tg_geometry_create(int32_t dtype, int32_t gdim, const void* nodes, uint64_t num_nodes,
                       const double* origin, const tg_block_desc* blocks, int32_t num_blocks,
                       tg_geometry** out) {
  if (!out) return fail(TG_ERR_INVALID_ARGUMENT, "null output handle");
  *out = nullptr;
  if (dtype != TG_F32 && dtype != TG_F64)
    return fail(TG_ERR_UNSUPPORTED_PRECISION, "UnsupportedPrecision: dtype must be F32 or F64");
  if (gdim < 1 || gdim > 3) return fail(TG_ERR_INVALID_ARGUMENT, "gdim must be 1, 2 or 3");
  if (num_blocks < 0 || (num_blocks > 0 && !blocks))
    return fail(TG_ERR_INVALID_ARGUMENT, "bad block list");
  if (num_nodes > 0 && !nodes) return fail(TG_ERR_INVALID_ARGUMENT, "null nodes");
  if (num_nodes > 0xffffffffull) return fail(TG_ERR_TOO_LARGE, "TooLarge: more than 2^32 nodes");
  auto g = std::make_unique<tg_geometry>();
  g->dtype = dtype;
  g->gdim = gdim;
  g->nodes = nodes;
  g->num_nodes = num_nodes;
  for (int i = 0; i < 3; ++i) g->origin[i] = (origin && i < gdim) ? origin[i] : 0.0;
  for (int bi = 0; bi < num_blocks; ++bi) {
    const tg_block_desc& d = blocks[bi];
    if (!tg::valid_cell(d.cell) || d.cell == TG_POINT)
      return fail(TG_ERR_INVALID_ARGUMENT, "block " + std::to_string(bi) + ": bad cell type");
    const tg::Cell cell = static_cast<tg::Cell>(d.cell);
    if (tg::tdim(cell) > gdim)
      return fail(TG_ERR_INVALID_ARGUMENT, "block " + std::to_string(bi) + ": tdim > gdim");
    if (d.degree == 0)
      return fail(TG_ERR_UNSUPPORTED_DEGREE, std::string("UnsupportedDegree { cell: ") +
                                                 cell_name(d.cell) + ", degree: 0 }");
    if (d.family != TG_EQUISPACED && d.family != TG_GLL)
      return fail(TG_ERR_INVALID_ARGUMENT, "bad node family");
    if (d.family == TG_GLL && cell != tg::kInterval && cell != tg::kQuadrilateral &&
        cell != tg::kHexahedron)
      return fail(TG_ERR_UNSUPPORTED_NODE_FAMILY, std::string("UnsupportedNodeFamily { cell: ") +
                                                      cell_name(d.cell) + ", family: Gll }");
    const uint64_t k = tg::num_lagrange_nodes(cell, d.degree);
    // geometry.rs:122-140: the stride must equal k and the length be a multiple of it.
    if (d.dofmap_stride != k || d.dofmap_len % k != 0)
      return fail(TG_ERR_LENGTH_MISMATCH, "LengthMismatch { expected: " + std::to_string(k) +
                                              ", found: " + std::to_string(d.dofmap_stride) + " }");
    if (num_nodes == 0 && d.dofmap_len != 0)
      return fail(TG_ERR_LENGTH_MISMATCH, "LengthMismatch { expected: 0, found: " +
                                              std::to_string(d.dofmap_len) + " }");
    if (d.dofmap_len > 0 && !d.dofmap) return fail(TG_ERR_INVALID_ARGUMENT, "null dofmap");
    Block b{cell, d.degree, static_cast<tg::Family>(d.family), d.dofmap, d.dofmap_len / k,
            d.flags, static_cast<uint32_t>(k), tg::tdim(cell)};
    g->blocks.push_back(b);
  }
  *out = g.release();
  return TG_OK;
}

int tg_geometry_destroy(tg_geometry* g) {
  delete g;
  return TG_OK;
}

int tg_geometry_num_blocks(const tg_geometry* g) {
  return g ? static_cast<int>(g->blocks.size()) : -1;
}

uint64_t tg_geometry_block_cells(const tg_geometry* g, int32_t block) {
  if (!g || block < 0 || block >= static_cast<int>(g->blocks.size())) return 0;
  return g->blocks[block].num_cells;
}

int tg_geometry_set_fixup_threshold(tg_geometry* g, double tau) {
  if (!g || !(tau >= 0.0)) return fail(TG_ERR_INVALID_ARGUMENT, "bad geometry or threshold");
  g->tau = tau;
  return TG_OK;
}

int tg_geometry_set_nodes_f64(tg_geometry* g, const double* nodes_f64) {
  if (!g) return fail(TG_ERR_INVALID_ARGUMENT, "null geometry");
  if (g->dtype != TG_F32 && nodes_f64)
    return fail(TG_ERR_UNSUPPORTED_PRECISION,
                "UnsupportedPrecision: FP64 nodes only widen an FP32 geometry");
  g->nodes64 = nodes_f64;
  return TG_OK;
}

int tg_pointset_create(int32_t cell, const double* points, const double* weights, uint64_t nq,
                       const double* facet_normals, tg_pointset** out) {
  if (!out) return fail(TG_ERR_INVALID_ARGUMENT, "null output handle");
  *out = nullptr;
  if (!tg::valid_cell(cell)) return fail(TG_ERR_INVALID_ARGUMENT, "bad cell type");
  const int td = tg::tdim(static_cast<tg::Cell>(cell));
  if (nq > 0 && (!weights || (td > 0 && !points)))
    return fail(TG_ERR_INVALID_ARGUMENT, "null points or weights");
  std::vector<double> p(points, points + (points ? nq * td : 0));
  std::vector<double> w(weights, weights + (weights ? nq : 0));
  std::vector<double> n;
  if (facet_normals) n.assign(facet_normals, facet_normals + nq * td);
  return make_pointset(static_cast<tg::Cell>(cell), std::move(p), std::move(w), std::move(n),
                       facet_normals != nullptr, out);
}

int tg_pointset_gauss(int32_t cell, uint32_t degree, tg_pointset** out) {
  if (!out) return fail(TG_ERR_INVALID_ARGUMENT, "null output handle");
  if (!tg::valid_cell(cell)) return fail(TG_ERR_INVALID_ARGUMENT, "bad cell type");
  std::vector<double> p, w;
  tg::gauss_rule(static_cast<tg::Cell>(cell), degree, p, w);
  return make_pointset(static_cast<tg::Cell>(cell), std::move(p), std::move(w), {}, false, out);
}

int tg_pointset_facet_gauss(int32_t cell, uint32_t degree, tg_pointset** out) {
  if (!out) return fail(TG_ERR_INVALID_ARGUMENT, "null output handle");
  if (!tg::valid_cell(cell) || cell == TG_POINT)
    return fail(TG_ERR_INVALID_ARGUMENT, "bad cell type");
  const tg::Cell c = static_cast<tg::Cell>(cell);
  const int td = tg::tdim(c);
  const std::vector<double> fn = tg::reference_facet_normals(c);
  std::vector<double> pts, wts, nrm;
  for (int f = 0; f < tg::num_sub_entities(c, td - 1); ++f) {
    std::vector<double> fp, fw;
    tg::gauss_rule(tg::sub_entity_type(c, td - 1, f), degree, fp, fw);
    std::vector<double> mapped = tg::facet_points_to_cell(c, f, fp, fw.size());
    pts.insert(pts.end(), mapped.begin(), mapped.end());
    for (size_t q = 0; q < fw.size(); ++q)
      nrm.insert(nrm.end(), fn.begin() + f * td, fn.begin() + (f + 1) * td);
    wts.insert(wts.end(), fw.begin(), fw.end());
  }
  return make_pointset(c, std::move(pts), std::move(wts), std::move(nrm), true, out);
}

int tg_pointset_tet_rule(int32_t rule, tg_pointset** out) {
  if (!out) return fail(TG_ERR_INVALID_ARGUMENT, "null output handle");
  std::vector<double> p, w;
  if (rule == 0)
    tg::tet_rule_keast4(p, w);
  else if (rule == 1)
    tg::tet_rule_14(p, w);
  else
    return fail(TG_ERR_INVALID_ARGUMENT, "unknown tetrahedron rule");
  return make_pointset(tg::kTetrahedron, std::move(p), std::move(w), {}, false, out);
}

int tg_pointset_destroy(tg_pointset* p) {
  delete p;
  return TG_OK;
}

uint64_t tg_pointset_len(const tg_pointset* p) { return p ? p->nq : 0; }

int tg_pointset_get(const tg_pointset* p, double* points, double* weights, double* facet_normals,
                    uint64_t* num_normals) {
  if (!p) return fail(TG_ERR_INVALID_ARGUMENT, "null point set");
  if (points && !p->points.empty())
    std::memcpy(points, p->points.data(), p->points.size() * sizeof(double));
  if (weights && !p->weights.empty())
    std::memcpy(weights, p->weights.data(), p->weights.size() * sizeof(double));
  const size_t nn = p->has_normals ? p->normals.size() : 0;
  if (facet_normals && nn) std::memcpy(facet_normals, p->normals.data(), nn * sizeof(double));
  if (num_normals) *num_normals = nn;
  return TG_OK;
}

int tg_pointset_tabulate(const tg_pointset* p, uint32_t degree, int32_t family, double* values,
                         double* derivatives) {
  if (!p) return fail(TG_ERR_INVALID_ARGUMENT, "null point set");
  std::vector<double> inter;
  size_t k = 0;
  int rc = host_tables(p, degree, family, inter, &k);
  if (rc) return rc;
  for (size_t q = 0; q < p->nq; ++q)
    for (size_t i = 0; i < k; ++i) {
      if (values) values[q * k + i] = inter[(q * k + i) * 4];
      if (derivatives)
        for (int d = 0; d < p->tdim; ++d)
          derivatives[(q * k + i) * p->tdim + d] = inter[(q * k + i) * 4 + 1 + d];
    }
  return TG_OK;
}

int tg_eval_block(const tg_geometry* g, int32_t block, tg_pointset* pts, uint32_t factors,
                  uint32_t tile, uint32_t options, const tg_outputs* out, void* stream) {
  if (!g || !pts || !out) return fail(TG_ERR_INVALID_ARGUMENT, "null geometry, point set or outputs");
  if (block < 0 || block >= static_cast<int>(g->blocks.size()))
    return fail(TG_ERR_INVALID_ARGUMENT, "block index out of range");
  const Block& b = g->blocks[block];
  if (pts->cell != b.cell)
    return fail(TG_ERR_CELL_NOT_IN_RULE, std::string("CellNotInRule(\"") + cell_name(b.cell) + "\")");
  if (!valid_tile(tile)) return fail(TG_ERR_INVALID_ARGUMENT, "tile must be 1, 2, 4, 8 or 16");
  const uint32_t all = TG_POINTS | TG_JACOBIAN | TG_DETJ | TG_PINV | TG_FACET_NORMALS | TG_NORMALS;
  if (factors == 0 || (factors & ~all)) return fail(TG_ERR_INVALID_ARGUMENT, "bad factor mask");
  if ((factors & TG_NORMALS) && b.tdim + 1 != g->gdim)
    return fail(TG_ERR_INVALID_ARGUMENT, "normals need a block of codimension 1");
  if ((factors & TG_FACET_NORMALS) && (b.tdim != g->gdim || !pts->has_normals))
    return fail(TG_ERR_INVALID_ARGUMENT,
                "facet normals need a block of codimension 0 and a facet point set");
  const void* ptrs[6] = {out->points, out->jacobians, out->detj, out->pinv, out->facet_normals,
                         out->normals};
  for (int i = 0; i < 6; ++i)
    if ((factors & (1u << i)) && !ptrs[i])
      return fail(TG_ERR_INVALID_ARGUMENT, "missing output for a requested factor");
  if (b.num_cells == 0 || pts->nq == 0) return TG_OK;  // layout.rs:509-511: nothing to launch
  if (pts->nq > 0xffffffffull) return fail(TG_ERR_TOO_LARGE, "TooLarge: too many points");
  DeviceTables* tabs = nullptr;
  int rc = device_tables(pts, b.degree, b.family, g->dtype, &tabs);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return g->dtype == TG_F32
             ? tg::enqueue_block<float>(g, b, pts, tabs, factors, tile, options, out, s)
             : tg::enqueue_block<double>(g, b, pts, tabs, factors, tile, options, out, s);
}

uint64_t tg_evector_elems(const tg_geometry* g, int32_t block) {
  if (!g || block < 0 || block >= static_cast<int>(g->blocks.size())) return 0;
  const Block& b = g->blocks[block];
  return static_cast<uint64_t>(b.k) * g->gdim * ((b.num_cells + 15) / 16 * 16);
}

int tg_evector_build(const tg_geometry* g, int32_t block, void* out, void* stream) {
  if (!g || !out) return fail(TG_ERR_INVALID_ARGUMENT, "null geometry or output");
  if (block < 0 || block >= static_cast<int>(g->blocks.size()))
    return fail(TG_ERR_INVALID_ARGUMENT, "block index out of range");
  const Block& b = g->blocks[block];
  if (b.num_cells == 0) return TG_OK;
  const uint64_t padded = (b.num_cells + 15) / 16 * 16;
  const uint32_t last = static_cast<uint32_t>(g->num_nodes - 1);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e =
      g->dtype == TG_F32
          ? tg::launch_evector<float>(static_cast<const float*>(g->nodes), b.dofmap, b.num_cells,
                                      padded, b.k, g->gdim, last, static_cast<float*>(out), s)
          : tg::launch_evector<double>(static_cast<const double*>(g->nodes), b.dofmap,
                                       b.num_cells, padded, b.k, g->gdim, last,
                                       static_cast<double*>(out), s);
  if (e != hipSuccess) return hip_fail(e, "evector kernel launch");
  return TG_OK;
}

}  // extern "C"
