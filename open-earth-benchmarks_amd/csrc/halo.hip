// halo.hip -- halo update between the stencils of a chain, and HIP-graph
// capture of a chain step.
//
// The reference has neither: each kernel writes only its output region, and
// only the drivers' fillMath (include/util.h:155-163) ever writes a halo.  A
// chain of stencils, whose outputs feed the next stencil's inputs, has to
// refill the halo a stencil reads before it runs.  oeb_halo_fill_* does that
// in two launches on the caller's stream:
//   1. the i/j ring [-h, nx+h) x [-h, ny+h) minus the interior, on levels
//      [0, nz): periodic (wrapped index) or zero-gradient (clamped index);
//   2. the k-halo planes [-hk, 0) and [nz, nz+hk) over the whole i/j box:
//      copies of planes 0 and nz-1 (zero-gradient in k whatever the mode,
//      since launch 1 has already filled those planes' rings).
// Both are pure copies, HBM-bound and small next to the stencils: the ring is
// 2h(nx + ny + 2h) points per level.
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ int wrap(int a, int n) {
  const int r = a % n;
  return r < 0 ? r + n : r;
}

__device__ __forceinline__ int clampi(int a, int n) { return a < 0 ? 0 : (a >= n ? n - 1 : a); }

// Ring point e of one level -> (i, j): h rows below the domain, h rows above
// (each nx + 2h wide), then 2h points per interior row (h left, h right).
__device__ __forceinline__ void ring_point(int64_t e, int nx, int ny, int h, int &i, int &j) {
  const int64_t w = nx + 2 * h;
  const int64_t band = int64_t(h) * w;
  if (e < band) {
    j = -h + int(e / w), i = -h + int(e % w);
  } else if (e < 2 * band) {
    const int64_t f = e - band;
    j = ny + int(f / w), i = -h + int(f % w);
  } else {
    const int64_t f = e - 2 * band;
    j = int(f / (2 * h));
    const int c = int(f % (2 * h));
    i = c < h ? -h + c : nx + (c - h);
  }
}

template <class T>
struct RingArgs {
  T *f;
  int nx, ny, nz, h, sj, periodic;
  int64_t sk, ring;
};

template <class T>
struct KArgs {
  T *f;
  int nx, ny, nz, h, hk, sj;
  int64_t sk;
};

// Logical block `id`: ring elements [id*256, id*256 + 256) of all levels.
template <class T>
__device__ __forceinline__ void halo_ring_body(const RingArgs<T> &a, const Geo &, const Tiles &, int id) {
  const int64_t e = int64_t(id) * kThreads + threadIdx.x;
  if (e >= a.ring * a.nz) return;
  const int k = int(e / a.ring);
  int i, j;
  ring_point(e % a.ring, a.nx, a.ny, a.h, i, j);
  const int si = a.periodic ? wrap(i, a.nx) : clampi(i, a.nx);
  const int sjj = a.periodic ? wrap(j, a.ny) : clampi(j, a.ny);
  const int64_t plane = int64_t(k) * a.sk;
  a.f[plane + i + int64_t(j) * a.sj] = a.f[plane + si + int64_t(sjj) * a.sj];
}

// Logical block `id`: elements [id*256, id*256 + 256) of the 2*hk k-halo
// planes (each the whole i/j box).
template <class T>
__device__ __forceinline__ void halo_k_body(const KArgs<T> &a, const Geo &, const Tiles &, int id) {
  const int64_t w = a.nx + 2 * a.h, box = w * (a.ny + 2 * a.h);
  const int64_t e = int64_t(id) * kThreads + threadIdx.x;
  if (e >= box * 2 * a.hk) return;
  const int64_t p = e % box;
  const int q = int(e / box);  // 0 .. 2hk-1: planes -hk..-1, then nz..nz+hk-1
  const int k = q < a.hk ? q - a.hk : a.nz + (q - a.hk);
  const int ks = q < a.hk ? 0 : a.nz - 1;
  const int i = -a.h + int(p % w), j = -a.h + int(p / w);
  const int64_t at = i + int64_t(j) * a.sj;
  a.f[int64_t(k) * a.sk + at] = a.f[int64_t(ks) * a.sk + at];
}

template <class A, void (*Body)(const A &, const Geo &, const Tiles &, int)>
__global__ void __launch_bounds__(kThreads) halo_kernel(A a, int blocks) {
  for (int b = int(blockIdx.x); b < blocks; b += int(gridDim.x)) Body(a, Geo{}, Tiles{}, b);
}

template <class T>
hipError_t halo_ring_batch(const BatchMember *tab, int n, int total, hipStream_t s) {
  return launch_batch<RingArgs<T>, halo_ring_body<T>>(tab, n, total, s, dim3(kThreads));
}
template <class T>
hipError_t halo_k_batch(const BatchMember *tab, int n, int total, hipStream_t s) {
  return launch_batch<KArgs<T>, halo_k_body<T>>(tab, n, total, s, dim3(kThreads));
}
template <class T>
const BatchKind kHaloRingBatch{"halo_fill (ring)", halo_ring_batch<T>};
template <class T>
const BatchKind kHaloKBatch{"halo_fill (k planes)", halo_k_batch<T>};

int grid_for(int64_t blocks) { return int(blocks < 8192 ? (blocks < 1 ? 1 : blocks) : 8192); }

template <class T, class S>
int halo_run(const oeb_domain *d, const S *field, int halo_ij, int halo_k, int mode, void *stream) {
  const char *K = "halo_fill";
  OEB_TRY(check_domain(K, d));
  if (halo_ij < 0 || halo_ij > 16 || halo_k < 0 || halo_k > 16)
    return fail(OEB_ERR_ARG, "halo_fill: halo_ij and halo_k must lie in [0, 16]");
  if (mode != OEB_HALO_PERIODIC && mode != OEB_HALO_ZERO_GRADIENT)
    return fail(OEB_ERR_ARG, "halo_fill: mode must be OEB_HALO_PERIODIC or OEB_HALO_ZERO_GRADIENT");
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz, h = halo_ij, hk = halo_k;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box b{-h, nx + h, -h, ny + h, -hk, nz + hk};
  const T *p;
  OEB_TRY(check3(K, "field", field, b, g, true, &p));
  T *f = const_cast<T *>(p);
  if (h > 0) {
    const int64_t ring = 2 * h * (nx + 2 * h) + ny * 2 * h;
    const int64_t blocks = (ring * nz + kThreads - 1) / kThreads;
    if (blocks > (int64_t(1) << 30)) return fail(OEB_ERR_DOMAIN, "halo_fill: domain too large");
    const RingArgs<T> a{f, int(nx), int(ny), int(nz), int(h), g.sj, mode == OEB_HALO_PERIODIC ? 1 : 0, g.sk, ring};
    if (!batch_record(kHaloRingBatch<T>, a, g, Tiles{}, int(blocks))) {
      hipLaunchKernelGGL((halo_kernel<RingArgs<T>, halo_ring_body<T>>), dim3(grid_for(blocks)), dim3(kThreads), 0,
                         hipStream_t(stream), a, int(blocks));
      OEB_TRY(launch_status(K));
    }
  }
  if (hk > 0) {
    const int64_t blocks = ((nx + 2 * h) * (ny + 2 * h) * 2 * hk + kThreads - 1) / kThreads;
    if (blocks > (int64_t(1) << 30)) return fail(OEB_ERR_DOMAIN, "halo_fill: domain too large");
    const KArgs<T> a{f, int(nx), int(ny), int(nz), int(h), int(hk), g.sj, g.sk};
    if (!batch_record(kHaloKBatch<T>, a, g, Tiles{}, int(blocks))) {
      hipLaunchKernelGGL((halo_kernel<KArgs<T>, halo_k_body<T>>), dim3(grid_for(blocks)), dim3(kThreads), 0,
                         hipStream_t(stream), a, int(blocks));
      OEB_TRY(launch_status(K));
    }
  }
  return OEB_OK;
}

}  // namespace
}  // namespace oeb

extern "C" int oeb_halo_fill_f64(const oeb_domain *d, const oeb_s3 *field, int halo_ij, int halo_k, int mode,
                                 void *stream) {
  return oeb::halo_run<double>(d, field, halo_ij, halo_k, mode, stream);
}

extern "C" int oeb_halo_fill_f32(const oeb_domain *d, const oeb_s3f *field, int halo_ij, int halo_k, int mode,
                                 void *stream) {
  return oeb::halo_run<float>(d, field, halo_ij, halo_k, mode, stream);
}

// ------------------------------------------------------------------ graphs
struct oeb_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

extern "C" int oeb_graph_begin(void *stream) {
  using namespace oeb;
  if (stream == nullptr) return fail(OEB_ERR_ARG, "graph_begin: the null stream cannot be captured; pass a stream");
  const hipError_t e = hipStreamBeginCapture(hipStream_t(stream), hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return fail(OEB_ERR_HIP, std::string("graph_begin: ") + hipGetErrorString(e));
  return OEB_OK;
}

extern "C" int oeb_graph_end(void *stream, oeb_graph **graph) {
  using namespace oeb;
  if (graph == nullptr) return fail(OEB_ERR_NULL, "graph_end: graph is NULL");
  *graph = nullptr;
  if (stream == nullptr) return fail(OEB_ERR_ARG, "graph_end: the null stream cannot be captured; pass a stream");
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(hipStream_t(stream), &g);
  if (e != hipSuccess) return fail(OEB_ERR_HIP, std::string("graph_end: ") + hipGetErrorString(e));
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(OEB_ERR_HIP, std::string("graph_end: instantiate: ") + hipGetErrorString(e));
  }
  *graph = new oeb_graph{g, x};
  return OEB_OK;
}

extern "C" int oeb_graph_launch(const oeb_graph *graph, void *stream) {
  using namespace oeb;
  if (graph == nullptr) return fail(OEB_ERR_NULL, "graph_launch: graph is NULL");
  const hipError_t e = hipGraphLaunch(graph->exec, hipStream_t(stream));
  if (e != hipSuccess) return fail(OEB_ERR_HIP, std::string("graph_launch: ") + hipGetErrorString(e));
  return OEB_OK;
}

extern "C" int oeb_graph_destroy(oeb_graph *graph) {
  using namespace oeb;
  if (graph == nullptr) return OEB_OK;
  const hipError_t e1 = hipGraphExecDestroy(graph->exec);
  const hipError_t e2 = hipGraphDestroy(graph->graph);
  delete graph;
  if (e1 != hipSuccess || e2 != hipSuccess)
    return fail(OEB_ERR_HIP, std::string("graph_destroy: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2));
  return OEB_OK;
}
