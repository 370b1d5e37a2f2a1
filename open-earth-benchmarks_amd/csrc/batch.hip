// batch.hip -- recording entry-point launches for an ensemble and replaying
// each run of one kernel as a single persistent launch (oeb_batch_*).
//
// While a batch records on a host thread, the batchable entry points
// validate their arguments as usual but, instead of launching, append a
// BatchMember (kernel arguments, geometry, tiles) to the batch.
// oeb_batch_stop splits the records into runs of consecutive records of one
// kernel, numbers each run's blocks member after member, and uploads each
// run's member table once.  The records of a run execute concurrently, so a
// record that reads what an earlier record of its run writes, or writes what
// an earlier record reads or writes, starts a new run: the launches then
// keep the order in which they were recorded.  oeb_batch_launch then enqueues one persistent
// launch per run (batch_kernel in oeb_common.hpp): kernels only, so a launch
// can be captured in a graph and replayed.
#include <cstring>
#include <vector>

#include "oeb_common.hpp"

struct oeb_batch {
  struct Range {
    uintptr_t lo, hi;
    bool out;
  };
  struct Run {
    const oeb::BatchKind *kind;
    oeb::BatchMember *table;  // device memory, n entries
    int n, total;
  };
  std::vector<oeb::BatchMember> records;
  std::vector<const oeb::BatchKind *> kinds;
  std::vector<std::vector<Range>> ranges;  // bytes each record reads / writes
  std::vector<Run> runs;
  bool recording = false, stopped = false;
};

namespace oeb {
namespace {
thread_local oeb_batch *g_recording = nullptr;
}

bool batch_record(const BatchKind &kind, const void *args, size_t bytes, const Geo &g, const Tiles &t, int blocks,
                  const AliasSet &as) {
  oeb_batch *b = g_recording;
  if (b == nullptr) return false;
  std::vector<oeb_batch::Range> rs;
  for (int i = 0; i < as.n; ++i) rs.push_back({as.lo[i], as.hi[i], as.out[i]});
  b->ranges.push_back(std::move(rs));
  BatchMember m{};
  m.g = g;
  m.t = t;
  m.first = 0;
  m.blocks = blocks;
  std::memcpy(m.args, args, bytes);
  b->records.push_back(m);
  b->kinds.push_back(&kind);
  return true;
}

bool batch_recording() { return g_recording != nullptr; }

int refuse_while_recording(const char *kname) {
  if (g_recording == nullptr) return OEB_OK;
  return fail(OEB_ERR_ARG, std::string(kname) +
                               ": has no batched form and cannot be called while this thread records a batch "
                               "(call it before oeb_batch_record or after oeb_batch_stop)");
}

namespace {
// True when record b depends on record a: b writes what a reads or writes,
// or reads what a writes.
bool depends(const std::vector<oeb_batch::Range> &a, const std::vector<oeb_batch::Range> &b) {
  for (const auto &x : a)
    for (const auto &y : b)
      if ((x.out || y.out) && x.lo < y.hi && y.lo < x.hi) return true;
  return false;
}
}  // namespace

}  // namespace oeb

extern "C" int oeb_batch_create(oeb_batch **batch) {
  using namespace oeb;
  if (batch == nullptr) return fail(OEB_ERR_NULL, "batch_create: batch is NULL");
  *batch = new oeb_batch();
  return OEB_OK;
}

extern "C" int oeb_batch_record(oeb_batch *batch) {
  using namespace oeb;
  if (batch == nullptr) return fail(OEB_ERR_NULL, "batch_record: batch is NULL");
  if (batch->stopped) return fail(OEB_ERR_ARG, "batch_record: the batch is already stopped");
  if (g_recording != nullptr) return fail(OEB_ERR_ARG, "batch_record: this thread is already recording a batch");
  batch->recording = true;
  g_recording = batch;
  return OEB_OK;
}

extern "C" int oeb_batch_stop(oeb_batch *batch) {
  using namespace oeb;
  if (batch == nullptr) return fail(OEB_ERR_NULL, "batch_stop: batch is NULL");
  if (g_recording != batch) return fail(OEB_ERR_ARG, "batch_stop: this thread is not recording this batch");
  g_recording = nullptr;
  batch->recording = false;
  batch->stopped = true;
  const size_t n = batch->records.size();
  for (size_t a = 0; a < n;) {
    size_t b = a;
    int64_t total = 0;
    while (b < n && batch->kinds[b] == batch->kinds[a]) {
      bool dep = false;
      for (size_t c = a; c < b && !dep; ++c) dep = depends(batch->ranges[c], batch->ranges[b]);
      if (dep) break;
      batch->records[b].first = int(total);
      total += batch->records[b].blocks;
      ++b;
    }
    if (total > (int64_t(1) << 30)) return fail(OEB_ERR_DOMAIN, "batch_stop: a run has more than 2^30 blocks");
    BatchMember *d = nullptr;
    hipError_t e = hipMalloc(&d, (b - a) * sizeof(BatchMember));
    if (e == hipSuccess)
      e = hipMemcpy(d, batch->records.data() + a, (b - a) * sizeof(BatchMember), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      if (d != nullptr) (void)hipFree(d);
      return fail(OEB_ERR_HIP, std::string("batch_stop: ") + hipGetErrorString(e));
    }
    batch->runs.push_back({batch->kinds[a], d, int(b - a), int(total)});
    a = b;
  }
  return OEB_OK;
}

extern "C" int oeb_batch_launch(const oeb_batch *batch, void *stream) {
  using namespace oeb;
  if (batch == nullptr) return fail(OEB_ERR_NULL, "batch_launch: batch is NULL");
  if (!batch->stopped) return fail(OEB_ERR_ARG, "batch_launch: stop the batch before launching it");
  for (const auto &r : batch->runs) {
    const hipError_t e = r.kind->launch(r.table, r.n, r.total, hipStream_t(stream));
    if (e != hipSuccess)
      return fail(OEB_ERR_HIP, std::string("batch_launch: ") + r.kind->name + ": " + hipGetErrorString(e));
  }
  return OEB_OK;
}

extern "C" int oeb_batch_info(const oeb_batch *batch, int *records, int *launches) {
  using namespace oeb;
  if (batch == nullptr) return fail(OEB_ERR_NULL, "batch_info: batch is NULL");
  if (records != nullptr) *records = int(batch->records.size());
  if (launches != nullptr) *launches = int(batch->runs.size());
  return OEB_OK;
}

extern "C" int oeb_batch_destroy(oeb_batch *batch) {
  using namespace oeb;
  if (batch == nullptr) return OEB_OK;
  if (g_recording == batch) g_recording = nullptr;
  hipError_t first = hipSuccess;
  for (auto &r : batch->runs) {
    const hipError_t e = hipFree(r.table);
    if (first == hipSuccess) first = e;
  }
  delete batch;
  if (first != hipSuccess) return fail(OEB_ERR_HIP, std::string("batch_destroy: ") + hipGetErrorString(first));
  return OEB_OK;
}
