// halo.hip -- halo update between the stencils of a chain, and HIP-graph
// capture of a chain step.
//
// The reference has neither: each kernel writes only its output region, and
// only the drivers' fillMath (include/util.h:155-163) ever writes a halo.  A
// chain of stencils, whose outputs feed the next stencil's inputs, has to
// refill the halo a stencil reads before it runs.  oeb_halo_fill_* does that
// for up to kMaxFields fields of one layout in ONE launch on the caller's
// stream, over
//   1. the i/j ring [-h, nx+h) x [-h, ny+h) minus the interior, on levels
//      [0, nz), and
//   2. the k-halo planes [-hk, 0) and [nz, nz+hk) over the whole i/j box.
// Every halo point copies the interior point at its wrapped (periodic) or
// clamped (zero-gradient) i/j index on its clamped level (zero-gradient in k
// whatever the mode).  Sources are interior points only, so no point of the
// launch reads what another writes and one launch does both parts.  It is a
// pure copy of 2h(nx + ny + 2h) points per level plus 2hk planes, HBM-bound
// and small next to the stencils; at the sizes of a chain its time is the
// fixed cost of a launch, which one launch for all fields of a refill pays
// once.
#include "oeb_common.hpp"

namespace oeb {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxFields = 8;

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
struct HaloArgs {
  T *f[kMaxFields];
  int nx, ny, nz, h, hk, sj, periodic;
  int bpf;         // logical blocks per field
  int64_t sk;
  int64_t ring;    // ring points per level
  int64_t nring;   // ring * nz: ring points of all levels, before the k planes
  int64_t npts;    // halo points per field
};

// Logical block `id`: halo points [b*256, b*256 + 256) of field id / bpf,
// b = id % bpf; the ring points of levels 0..nz-1 first, then the 2hk k-halo
// planes (-hk..-1, then nz..nz+hk-1), each the whole i/j box.
template <class T>
__device__ __forceinline__ void halo_body(const HaloArgs<T> &a, const Geo &, const Tiles &, int id) {
  const int fld = id / a.bpf;
  const int64_t e = int64_t(id - fld * a.bpf) * kThreads + threadIdx.x;
  if (e >= a.npts) return;
  int i, j, k;
  if (e < a.nring) {
    k = int(e / a.ring);
    ring_point(e - int64_t(k) * a.ring, a.nx, a.ny, a.h, i, j);
  } else {
    const int64_t w = a.nx + 2 * a.h, box = w * (a.ny + 2 * a.h);
    const int64_t f = e - a.nring;
    const int q = int(f / box);
    const int64_t p = f - int64_t(q) * box;
    k = q < a.hk ? q - a.hk : a.nz + (q - a.hk);
    i = -a.h + int(p % w), j = -a.h + int(p / w);
  }
  const int si = a.periodic ? wrap(i, a.nx) : clampi(i, a.nx);
  const int sjj = a.periodic ? wrap(j, a.ny) : clampi(j, a.ny);
  const int ks = clampi(k, a.nz);
  T *f = a.f[fld];
  f[int64_t(k) * a.sk + i + int64_t(j) * a.sj] = f[int64_t(ks) * a.sk + si + int64_t(sjj) * a.sj];
}

template <class A, void (*Body)(const A &, const Geo &, const Tiles &, int)>
__global__ void __launch_bounds__(kThreads) halo_kernel(A a, int blocks) {
  for (int b = int(blockIdx.x); b < blocks; b += int(gridDim.x)) Body(a, Geo{}, Tiles{}, b);
}

template <class T>
hipError_t halo_batch(const BatchMember *tab, int n, int total, hipStream_t s) {
  return launch_batch<HaloArgs<T>, halo_body<T>>(tab, n, total, s, dim3(kThreads));
}
template <class T>
const BatchKind kHaloBatch{"halo_fill", halo_batch<T>};

int grid_for(int64_t blocks) { return int(blocks < 8192 ? (blocks < 1 ? 1 : blocks) : 8192); }

const char *const kFieldName[kMaxFields] = {"fields[0]", "fields[1]", "fields[2]", "fields[3]",
                                            "fields[4]", "fields[5]", "fields[6]", "fields[7]"};

template <class T, class S>
int halo_run(const char *K, const oeb_domain *d, const S *const *fields, int nfields, int halo_ij, int halo_k,
             int mode, void *stream) {
  OEB_TRY(check_domain(K, d));
  if (fields == nullptr) return fail(OEB_ERR_NULL, std::string(K) + ": fields is NULL");
  if (nfields < 1 || nfields > kMaxFields)
    return fail(OEB_ERR_ARG, std::string(K) + ": nfields must lie in [1, " + std::to_string(kMaxFields) + "]");
  if (halo_ij < 0 || halo_ij > 16 || halo_k < 0 || halo_k > 16)
    return fail(OEB_ERR_ARG, std::string(K) + ": halo_ij and halo_k must lie in [0, 16]");
  if (mode != OEB_HALO_PERIODIC && mode != OEB_HALO_ZERO_GRADIENT)
    return fail(OEB_ERR_ARG, std::string(K) + ": mode must be OEB_HALO_PERIODIC or OEB_HALO_ZERO_GRADIENT");
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz, h = halo_ij, hk = halo_k;
  Geo g{int(nx), int(ny), int(nz), 0, 0};
  const Box b{-h, nx + h, -h, ny + h, -hk, nz + hk};
  HaloArgs<T> a{};
  AliasSet as;
  for (int n = 0; n < nfields; ++n) {
    const T *p;
    OEB_TRY(check3(K, nfields == 1 ? "field" : kFieldName[n], fields[n], b, g, n == 0, &p));
    as.add(kFieldName[n], p, b, g, true);
    a.f[n] = const_cast<T *>(p);
  }
  OEB_TRY(as.check(K));
  const int64_t ring = 2 * h * (nx + 2 * h) + ny * 2 * h;
  const int64_t npts = ring * nz + (nx + 2 * h) * (ny + 2 * h) * 2 * hk;
  if (npts == 0) return OEB_OK;
  const int64_t bpf = (npts + kThreads - 1) / kThreads;
  if (bpf * nfields > (int64_t(1) << 30)) return fail(OEB_ERR_DOMAIN, std::string(K) + ": domain too large");
  a.nx = int(nx), a.ny = int(ny), a.nz = int(nz), a.h = int(h), a.hk = int(hk), a.sj = g.sj;
  a.periodic = mode == OEB_HALO_PERIODIC ? 1 : 0;
  a.bpf = int(bpf), a.sk = g.sk, a.ring = ring, a.nring = ring * nz, a.npts = npts;
  const int blocks = int(bpf * nfields);
  if (!batch_record(kHaloBatch<T>, a, g, Tiles{}, blocks, as)) {
    hipLaunchKernelGGL((halo_kernel<HaloArgs<T>, halo_body<T>>), dim3(grid_for(blocks)), dim3(kThreads), 0,
                       hipStream_t(stream), a, blocks);
    OEB_TRY(launch_status(K));
  }
  return OEB_OK;
}

}  // namespace
}  // namespace oeb

extern "C" int oeb_halo_fill_f64(const oeb_domain *d, const oeb_s3 *field, int halo_ij, int halo_k, int mode,
                                 void *stream) {
  return oeb::halo_run<double>("halo_fill", d, &field, 1, halo_ij, halo_k, mode, stream);
}

extern "C" int oeb_halo_fill_f32(const oeb_domain *d, const oeb_s3f *field, int halo_ij, int halo_k, int mode,
                                 void *stream) {
  return oeb::halo_run<float>("halo_fill", d, &field, 1, halo_ij, halo_k, mode, stream);
}

extern "C" int oeb_halo_fill_n_f64(const oeb_domain *d, const oeb_s3 *const *fields, int nfields, int halo_ij,
                                   int halo_k, int mode, void *stream) {
  return oeb::halo_run<double>("halo_fill_n", d, fields, nfields, halo_ij, halo_k, mode, stream);
}

extern "C" int oeb_halo_fill_n_f32(const oeb_domain *d, const oeb_s3f *const *fields, int nfields, int halo_ij,
                                   int halo_k, int mode, void *stream) {
  return oeb::halo_run<float>("halo_fill_n", d, fields, nfields, halo_ij, halo_k, mode, stream);
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
