// oeb_common.hpp -- shared host and device pieces of liboeb.
//
// Geometry.  Every 3-D array of one call shares the j stride `sj` and the k
// stride `sk` (in elements, i stride 1).  A kernel receives a pointer to each
// array's element (0,0,0), so element (i,j,k) is p[i + j*sj + k*sk].  The
// in-plane part i + j*sj is 32-bit (the host checks that a plane fits).  The
// plane part k*sk is 64-bit.
//
// Work decomposition.  A workgroup owns a tile of TILE_I x TILE_J columns and
// a chunk of levels.  Each lane owns one column (i, j) and marches up its
// chunk.  Vertical neighbours are carried in registers from one level to the
// next.  Horizontal neighbours come from L1/L2, or from LDS where a stencil
// stages a tile.  Tiles are numbered i fastest, then j, then k.  The hardware
// deals consecutive workgroup ids round-robin to the 8 XCDs, so neighbouring
// tiles run at the same time on different XCDs.  Their shared halo lines are
// then served by the MALL, which all XCDs share.  Remapping ids so that each
// XCD owns a contiguous run measured 3-10% slower on every stencil shape tried;
// laplace interleaves level chunks between tile rows instead (Tiles::xcd).
//
// Memory access.  Outputs are written once and never read back, so they are
// stored non-temporally (st_nt).  Inputs read only at the centre point are
// loaded non-temporally (ld_nt).  Inputs whose neighbours are read keep the
// normal cache policy, so the neighbour reads hit L1/L2.  Non-temporal
// streaming avoids the burst of dirty-line write-back at kernel end.  It
// measured 10-25% faster than normal stores for plain copies of 84-336 MB.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <initializer_list>
#include <string>
#include <type_traits>

#include "../../include/oeb.h"

namespace oeb {

// ------------------------------------------------------------------ geometry
struct Geo {
  int nx, ny, nz;
  int sj;      // j stride, elements
  int64_t sk;  // k stride, elements
};

// Half-open index box [ilo,ihi) x [jlo,jhi) x [klo,khi), relative to the
// domain origin, for which a stencil reads or writes an array.
struct Box {
  int64_t ilo, ihi, jlo, jhi, klo, khi;
};

// ------------------------------------------------------------------ knobs
// Launch parameters that the autotuner may set per kernel, element size and
// domain (oeb_tune_knob_set).  Each entry point asks for the knobs it has with
// knob(), giving its built-in default; the tuning table's most specific entry
// (this domain, then every domain; this element size, then both) wins over it.
//   KNOB_KCHUNK  levels per workgroup of a tile launch
//   KNOB_VECTOR  1: the 16-byte vector / column-pair body where it applies
//   KNOB_FORM    1: the strip (row march) form, 0: the tile form, 2: fvtp2d's
//                LDS-staged tiles
//   KNOB_ROWS    rows per strip chunk
//   KNOB_DEPTH   rows a strip loads ahead
//   KNOB_ORDER   tile order: 0 rows in order, 1 Tiles::xcd, 2 Tiles::irow,
//                3 Tiles::klev
//   KNOB_TILE    tile shape (kernel-specific index)
//   KNOB_LANES   workgroup size of a strip or tile
//   KNOB_GRID    workgroups of a batched launch
enum Knob {
  KNOB_KCHUNK,
  KNOB_VECTOR,
  KNOB_FORM,
  KNOB_ROWS,
  KNOB_DEPTH,
  KNOB_ORDER,
  KNOB_TILE,
  KNOB_LANES,
  KNOB_GRID,
  KNOB_COUNT
};
int knob(const char *kname, size_t esize, const oeb_domain *d, Knob k, int dflt);
inline bool knob_on(const char *kname, size_t esize, const oeb_domain *d, Knob k, bool dflt) {
  return knob(kname, esize, d, k, dflt ? 1 : 0) != 0;
}

// Tile layout of a column launch: logical id -> (tile_i, tile_j, chunk),
// tile_i fastest.  Workgroup ids go round-robin to the 8 XCDs (id mod 8), each
// with its own 4 MB L2; the orders below place tiles that share halo lines on
// one XCD.
//   ORDER_ROWS  rows of tiles in order, level chunks last.
//   ORDER_XCD   (xcd = g > 1) g consecutive level chunks interleaved between
//               the rows of tiles: id = ci + ti (sub + g (cj + tj kk)),
//               chunk = g kk + sub.  With ti g = 8 a tile and the tile below
//               it (id + 8) run on one XCD and fetch their shared halo rows
//               into its L2 once.
//   ORDER_IROW  (irow) the ti tiles of one row of one level chunk go to one
//               XCD, in consecutive slots of its dispatch order:
//               id = x + 8 (ci + ti q) for row r = x + 8 q (r = cj + tj chunk),
//               so that tiles side by side share their i-halo lines.
//   ORDER_KLEV  (klev) XCD x takes the level chunks x, x + 8, x + 16, ...,
//               each one whole, tiles in row order: id = x + 8 s with
//               s = ci + ti (cj + tj m), chunk = x + 8 m.  Tiles side by side
//               and one above the other then run on one XCD a few slots
//               apart, so both their i- and j-halo lines are shared in its L2.
//               The last nchunk mod 8 chunks follow in row order.
enum TileOrder { ORDER_ROWS = 0, ORDER_XCD = 1, ORDER_IROW = 2, ORDER_KLEV = 3 };
struct Tiles {
  int ti, tj, nchunk, kchunk, n;
  int xcd = 1;
  int irow = 0;
  int klev = 0;
};

// Tiles of TILE_I x TILE_J columns and `kchunk` levels in tile order `order`
// (TileOrder).  An order whose condition the shape does not meet falls back
// to ORDER_ROWS.
inline Tiles make_tiles(int64_t ni, int64_t nj, int nz, int tile_i, int tile_j, int kchunk, int order = ORDER_ROWS) {
  Tiles t;
  t.ti = int((ni + tile_i - 1) / tile_i);
  t.tj = int((nj + tile_j - 1) / tile_j);
  t.kchunk = kchunk < 1 ? 1 : kchunk;
  t.nchunk = (nz + t.kchunk - 1) / t.kchunk;
  t.n = t.ti * t.tj * t.nchunk;
  if (order == ORDER_XCD && t.ti < 8 && 8 % t.ti == 0 && t.nchunk % (8 / t.ti) == 0) t.xcd = 8 / t.ti;
  if (order == ORDER_IROW && (int64_t(t.tj) * t.nchunk) % 8 == 0) t.irow = 1;
  if (order == ORDER_KLEV && t.nchunk >= 8) t.klev = 1;
  return t;
}

// Non-temporal (streaming) store and load of one element (float or double).
template <class T>
__device__ __forceinline__ void st_nt(T *p, T v) {
  __builtin_nontemporal_store(v, p);
}
template <class T>
__device__ __forceinline__ T ld_nt(const T *p) {
  return __builtin_nontemporal_load(p);
}

// An array read or written a row at a time as a raw buffer (stride 0, range
// 2^31 - 1 bytes from its base).  An access takes the row's byte offset in a
// scalar register (soffset) and the lane's byte offset in a vector register
// (voffset), so a march down rows spends no vector instructions on addresses
// and holds one 32-bit offset per lane instead of a 64-bit pointer per array.
// The host keeps every offset below 2^31.  Word 3 of the descriptor is the
// gfx9 family's raw-buffer format (32-bit data, no swizzle).
struct RowBuf {
  __amdgpu_buffer_rsrc_t r;
};
template <class T>
__device__ __forceinline__ RowBuf row_buf(const T *base) {
  return RowBuf{__builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(base), 0, 0x7fffffff, 0x00020000)};
}
// Cache policy bit of the load and store: non-temporal.
constexpr int kBufNT = 2;
template <class T, int kAux = 0>
__device__ __forceinline__ T buf_ld(RowBuf b, unsigned voff, int soff) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(b.r, int(voff), soff, kAux));
  } else {
    static_assert(sizeof(T) == 8, "float or double");
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(b.r, int(voff), soff, kAux));
  }
}
template <class T>
__device__ __forceinline__ void buf_st_nt(RowBuf b, unsigned voff, int soff, T v) {
  if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), b.r, int(voff), soff, kBufNT);
  } else {
    static_assert(sizeof(T) == 8, "float or double");
    using u2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), b.r, int(voff), soff, kBufNT);
  }
}

// Pairs of adjacent doubles along i, moved with one 16-byte access.  Used by
// the paired kernels, which the host selects only when every 3-D array's
// element (0,0,0) is 16-byte aligned and both strides are even, so that
// p + (even i) + j*sj + k*sk is always 16-byte aligned.  A pair moves twice
// the bytes per memory instruction.  At 256x256x80 and 512x512x80 laplace
// measured 9-11% faster that way.
typedef double dvec2 __attribute__((ext_vector_type(2)));
struct d2 {
  double x, y;
};
__device__ __forceinline__ d2 ld2(const double *p) {
  const dvec2 v = *reinterpret_cast<const dvec2 *>(p);
  return d2{v.x, v.y};
}
__device__ __forceinline__ d2 ld2_nt(const double *p) {
  const dvec2 v = __builtin_nontemporal_load(reinterpret_cast<const dvec2 *>(p));
  return d2{v.x, v.y};
}
__device__ __forceinline__ void st2_nt(double *p, double x, double y) {
  const dvec2 v = {x, y};
  __builtin_nontemporal_store(v, reinterpret_cast<dvec2 *>(p));
}

// The FP64 fields of the FP32 forms ("islands": pressure, geopotential and
// heights, whose small differences between neighbouring points and levels
// cancel most of their digits), in one of two storages:
//   IslandF64  an FP64 array (oeb_<name>_f32);
//   IslandDev  FP32 deviations from an FP64 per-level base profile,
//              v(i,j,k) = base(k) + dev(i,j,k), added in FP64
//              (oeb_<name>_f32d): 4 bytes per point, and the FP32 rounding
//              applies to the deviation only.
// level(k) reads what a level contributes (IslandDev: the base value; read
// once per level, so that the deviation loads of the level issue together);
// at(p, lv) is then the value at 3-D index p of that level, and pair(p, lv)
// the values at p and p + 1 (p even: 16-byte aligned for doubles, 8-byte for
// floats, as pairs_ok() guarantees).
struct IslandF64 {
  const double *p;
  struct Level {};
  __device__ __forceinline__ Level level(int) const { return Level{}; }
  __device__ __forceinline__ double at(int64_t i, Level) const { return p[i]; }
  __device__ __forceinline__ d2 pair(int64_t i, Level) const { return ld2(p + i); }
  const void *ptr() const { return p; }
};
struct IslandDev {
  const float *p;
  const double *base;  // base(k) = base[k * bs]
  int bs;
  struct Level {
    double b;
  };
  __device__ __forceinline__ Level level(int k) const { return Level{base[int64_t(k) * bs]}; }
  __device__ __forceinline__ double at(int64_t i, Level lv) const { return lv.b + double(p[i]); }
  __device__ __forceinline__ d2 pair(int64_t i, Level lv) const {
    typedef float fvec2 __attribute__((ext_vector_type(2)));
    const fvec2 v = *reinterpret_cast<const fvec2 *>(p + i);
    return d2{lv.b + double(v.x), lv.b + double(v.y)};
  }
  const void *ptr() const { return p; }
};

// True when the paired (16-byte) kernels may be used for these arrays.
inline bool pairs_ok(const Geo &g, std::initializer_list<const void *> bases) {
  if ((g.sj & 1) || (g.sk & 1)) return false;
  for (const void *p : bases)
    if (reinterpret_cast<uintptr_t>(p) & 15) return false;
  return true;
}

// W consecutive elements along i, moved with one W*sizeof(T)-byte access
// (16 bytes for W = 2 doubles or 4 floats).  The host selects the vector
// kernels only when vec_ok() holds, so p + (i multiple of W) + j*sj + k*sk is
// always aligned.
template <class T, int W>
struct Vec {
  typedef T type __attribute__((ext_vector_type(W)));
};
template <class T, int W>
__device__ __forceinline__ typename Vec<T, W>::type ldv(const T *p) {
  return *reinterpret_cast<const typename Vec<T, W>::type *>(p);
}
template <class T, int W>
__device__ __forceinline__ typename Vec<T, W>::type ldv_nt(const T *p) {
  return __builtin_nontemporal_load(reinterpret_cast<const typename Vec<T, W>::type *>(p));
}
template <class T, int W>
__device__ __forceinline__ void stv_nt(T *p, typename Vec<T, W>::type v) {
  __builtin_nontemporal_store(v, reinterpret_cast<typename Vec<T, W>::type *>(p));
}

// True when the W-wide vector kernels may be used: nx, both strides and every
// base (element (0,0,0)) aligned to W elements of `esize` bytes.
inline bool vec_ok(const Geo &g, int w, size_t esize, std::initializer_list<const void *> bases) {
  if (g.nx % w || g.sj % w || g.sk % w) return false;
  for (const void *p : bases)
    if (reinterpret_cast<uintptr_t>(p) % (size_t(w) * esize)) return false;
  return true;
}

// Stages one level of a field into LDS: s[r][c] = f(i0 - R + c, j0 - R + r)
// for the (TI + 2R) x (TJ + 2R) tile, read row by row so that consecutive
// lanes read consecutive doubles.  Points with i >= ilim or j >= jlim lie
// outside the field's valid box and are set to 0; the caller never uses them.
// `f` points at element (0,0) of the level.
template <int TI, int TJ, int R, int NT, class T>
__device__ __forceinline__ void stage_tile(T (*s)[TI + 2 * R], const T *__restrict__ f, int i0, int j0, int sj,
                                           int ilim, int jlim, int tid) {
  constexpr int W = TI + 2 * R, H = TJ + 2 * R;
  for (int e = tid; e < W * H; e += NT) {
    const int r = e / W, c = e % W;
    const int gi = i0 - R + c, gj = j0 - R + r;
    s[r][c] = (gi < ilim && gj < jlim) ? f[gi + gj * sj] : T(0);
  }
}

struct TileCoord {
  int i0, j0, k0, k1;
};

// Tile coordinates of logical block `id` of a launch (tile_i fastest).
__device__ __forceinline__ TileCoord tile_coord_of(const Tiles &t, int id, int tile_i, int tile_j, int nz) {
  int ci, cj, ck;
  if (t.klev) {
    const int per = t.ti * t.tj, full = 8 * per * (t.nchunk / 8);
    if (id < full) {
      const int s = id >> 3;
      ci = s % t.ti;
      cj = (s / t.ti) % t.tj;
      ck = (id & 7) + 8 * (s / per);
    } else {  // the last nchunk mod 8 chunks, in row order
      const int r = id - full;
      ci = r % t.ti;
      cj = (r / t.ti) % t.tj;
      ck = 8 * (t.nchunk / 8) + r / per;
    }
  } else if (t.irow) {
    const int s = id >> 3, r = (id & 7) + 8 * (s / t.ti);
    ci = s % t.ti;
    cj = r % t.tj;
    ck = r / t.tj;
  } else {
    ci = id % t.ti;
    int rest = id / t.ti;
    int sub = 0;
    if (t.xcd > 1) {
      sub = rest % t.xcd;
      rest /= t.xcd;
    }
    cj = rest % t.tj;
    ck = (rest / t.tj) * t.xcd + sub;
  }
  TileCoord c;
  c.i0 = ci * tile_i;
  c.j0 = cj * tile_j;
  c.k0 = ck * t.kchunk;
  c.k1 = min(nz, c.k0 + t.kchunk);
  return c;
}

__device__ __forceinline__ TileCoord tile_coord(const Tiles &t, int tile_i, int tile_j, int nz) {
  return tile_coord_of(t, int(blockIdx.x), tile_i, tile_j, nz);
}

// Cross-lane moves within a wave (DPP on gfx9): from_right(v) is v of lane
// n+1 (wave_shl:1) and from_left(v) v of lane n-1 (wave_shr:1).  Lane 63
// (from_right), lane 0 (from_left) and a lane whose source lane is inactive
// get their own v back; the caller replaces those values.
template <int kCtrl, class T>
__device__ __forceinline__ T wave_shift(T v) {
  if constexpr (sizeof(T) == 8) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp(int(b), kCtrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(int(b >> 32), kCtrl, 0xf, 0xf, false);
    return __builtin_bit_cast(T, (long long)(unsigned)lo | ((long long)hi << 32));
  } else {
    static_assert(sizeof(T) == 4, "float or double");
    return __builtin_bit_cast(T, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), kCtrl, 0xf, 0xf, false));
  }
}
template <class T>
__device__ __forceinline__ T from_right(T v) {
  return wave_shift<0x130>(v);
}
template <class T>
__device__ __forceinline__ T from_left(T v) {
  return wave_shift<0x138>(v);
}

// Workgroup barrier that orders LDS accesses only.  __syncthreads() is a
// workgroup fence on all memory: before its s_barrier it waits for every
// outstanding global load and store of the wave (s_waitcnt vmcnt(0)), which
// drains the rows or levels a kernel has prefetched and makes each barrier
// wait for the write-back of the outputs just stored.  The kernels that use
// lds_barrier() share data between lanes through LDS only; global memory is
// never written by one lane and read by another within a launch.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ------------------------------------------------------------------ strips
// Row-march layout of the horizontal stencils (those with no k reach).  A
// workgroup owns the output columns [i0, i0 + w) of one level k over a chunk
// of rows [j0, j1), with one lane per column of [i0 - h, i0 + w + h): the
// strip and its i-halo.  It walks the rows in order, keeping each column's
// j-neighbours in registers and exchanging i-neighbours through LDS, so an
// input line is fetched once per row chunk.  A 64 x 4 tile fetches the lines
// of its i- and j-halo again for every tile it borders, and those tiles run on
// other XCDs (other L2s): 1.9-2.6x the algorithmic bytes for reach 1-3.  A
// strip re-fetches only the rows of its j-halo, (rows + 2h) / rows, and one
// line at each end of the row.  One strip spans nx when nx + 2h <= 1024.
struct Strips {
  int ns;     // strips across nx
  int w;      // output columns per strip (the last may hold fewer)
  int v;      // columns per lane
  int h;      // i-halo lanes each side
  int lanes;  // workgroup size: w / v + 2h rounded up to a wavefront
  int rows;   // rows per chunk
  int nj;     // chunks across ny
  int n;      // workgroups: ns * nj * nz
};

// Strips of at most max_lanes (<= 1024) lanes of v columns (w a multiple of
// v), h halo lanes each side.
inline Strips make_strips(int64_t nx, int64_t ny, int nz, int h, int rows, int v = 1, int max_lanes = 1024) {
  Strips s;
  s.h = h;
  s.v = v < 1 ? 1 : v;
  const int64_t maxw = int64_t(s.v) * (max_lanes - 2 * h);
  s.ns = int((nx + maxw - 1) / maxw);
  s.w = int(((nx + s.ns - 1) / s.ns + s.v - 1) / s.v * s.v);
  s.lanes = ((s.w / s.v + 2 * h + 63) / 64) * 64;
  s.rows = rows < 1 ? 1 : rows;
  s.nj = int((ny + s.rows - 1) / s.rows);
  s.n = s.ns * s.nj * nz;
  return s;
}

// Default rows per chunk for a strip of `lanes` lanes: ny * lanes / 4096,
// between 4 and 32, so that narrow strips (few waves per workgroup) get
// shorter chunks and more workgroups.  Measured at 256x256x80 and
// 512x512x80: longer chunks re-fetch fewer halo rows but leave CUs idle at
// the end of the launch; shorter ones re-fetch more.
inline int default_strip_rows(int64_t ny, int lanes) {
  const int64_t r = ny * lanes / 4096;
  return int(r < 4 ? 4 : (r > 32 ? 32 : r));
}

struct StripCoord {
  int i0, w, j0, j1, k;
};

__device__ __forceinline__ StripCoord strip_coord(const Strips &s, int id, int nx, int ny) {
  StripCoord c;
  const int si = id % s.ns;
  const int rest = id / s.ns;
  const int cj = rest % s.nj;
  c.k = rest / s.nj;
  c.i0 = si * s.w;
  c.w = min(s.w, nx - c.i0);
  c.j0 = cj * s.rows;
  c.j1 = min(ny, c.j0 + s.rows);
  return c;
}

// Fails with OEB_ERR_DOMAIN when a strip launch does not fit one grid.
int check_strips(const char *kname, const Strips &s);


// ------------------------------------------------------------------ batches
// A batch (oeb_batch_*) records the launches the entry points would make,
// for many members of an ensemble, and replays each run of consecutive
// records of one kernel as ONE launch.  The run's blocks are numbered member
// after member (member 0's blocks, then member 1's, ...): that list is the
// tile queue.  A grid of at most KNOB_GRID workgroups walks it with stride
// gridDim.x, each workgroup taking every gridDim.x-th tile, so no atomic
// counter has to be reset between replays; with the default cap the grid
// covers the whole list and each workgroup takes one tile.  Each record keeps the
// member's kernel arguments (a trivially copyable struct), geometry and tiles.
constexpr int kBatchArgBytes = 224;
// Default cap on the workgroups of a batched launch: none, i.e. one
// workgroup per block of every member.  A persistent grid of 2048 workgroups
// walking the blocks measured 1.16x slower on the 48-member ensemble (each
// workgroup then runs its blocks one after another, exposing every block's
// start-up latency); 8192 was 1.08x and 32768 1.04x slower.
constexpr int kBatchGrid = 1 << 30;

struct alignas(16) BatchMember {
  Geo g;
  Tiles t;
  int first;   // first flattened block of this member in its run
  int blocks;  // blocks of this member
  alignas(16) unsigned char args[kBatchArgBytes];
};

// Launches one run: `table` (device memory) holds n members, `total` blocks.
typedef hipError_t (*BatchLauncher)(const BatchMember *table, int n, int total, hipStream_t stream);

struct BatchKind {
  const char *name;
  BatchLauncher launch;
};

struct AliasSet;

// When this host thread is recording a batch, appends the launch, with the
// byte ranges the call reads and writes (`as`), and returns true; the caller
// then launches nothing.
bool batch_record(const BatchKind &kind, const void *args, size_t bytes, const Geo &g, const Tiles &t, int blocks,
                  const AliasSet &as);

template <class A>
bool batch_record(const BatchKind &kind, const A &args, const Geo &g, const Tiles &t, int blocks,
                  const AliasSet &as) {
  static_assert(sizeof(A) <= kBatchArgBytes, "batch arguments too large");
  static_assert(std::is_trivially_copyable<A>::value, "batch arguments must be trivially copyable");
  return batch_record(kind, &args, sizeof(A), g, t, blocks, as);
}

// True while this host thread records a batch.
bool batch_recording();

// Fails with OEB_ERR_ARG when this host thread is recording a batch: for the
// entry points that have no batched form, which would otherwise launch at
// once while the batch's earlier records have not run.
int refuse_while_recording(const char *kname);

// The member whose blocks include flattened block b (members sorted by first).
__device__ __forceinline__ int batch_find(const BatchMember *tab, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].first <= b) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return lo;
}

// Members whose first blocks a workgroup of a persistent launch keeps in LDS
// for the member lookup of each of its blocks.  A search through the table in
// global memory is a chain of dependent loads ahead of every block.
constexpr int kBatchLdsMembers = 256;

__device__ __forceinline__ int batch_find_lds(const int *firsts, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (firsts[mid] <= b) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return lo;
}

// The persistent kernel of a run: Body(args, geo, tiles, logical block id).
template <class A, void (*Body)(const A &, const Geo &, const Tiles &, int), int NT = 256>
__global__ void __launch_bounds__(NT) batch_kernel(const BatchMember *__restrict__ tab, int n, int total) {
  __shared__ int firsts[kBatchLdsMembers];
  const bool cached = n <= kBatchLdsMembers;
  if (cached) {
    const int nt = int(blockDim.x * blockDim.y), tid = int(threadIdx.x + blockDim.x * threadIdx.y);
    for (int m = tid; m < n; m += nt) firsts[m] = tab[m].first;
    lds_barrier();
  }
  for (int b = int(blockIdx.x); b < total; b += int(gridDim.x)) {
    const int m = cached ? batch_find_lds(firsts, n, b) : batch_find(tab, n, b);
    Body(*reinterpret_cast<const A *>(tab[m].args), tab[m].g, tab[m].t, b - tab[m].first);
  }
}


template <class A, void (*Body)(const A &, const Geo &, const Tiles &, int), int NT = 256>
hipError_t launch_batch(const BatchMember *table, int n, int total, hipStream_t stream, dim3 block) {
  const int cap = knob("batch", 0, nullptr, KNOB_GRID, kBatchGrid);
  const int grid = total < cap ? total : cap;
  hipLaunchKernelGGL((batch_kernel<A, Body, NT>), dim3(grid), block, 0, stream, table, n, total);
  return hipGetLastError();
}

// ------------------------------------------------------------- host checks
// Thread-local message of the last failing call (oeb_last_error).
void set_error(const std::string &msg);
int fail(int status, const std::string &msg);

// Validates the 3-D descriptor `s` of array `name` for kernel `kname`: the
// box `b` must lie inside the allocation it describes and its strides must
// match `g` (the first array checked fills g.sj / g.sk).  On success *base is
// the device address of element (0,0,0).
// Strides are counted in elements, so FP32 and FP64 arrays of one call share
// them.
int check3(const char *kname, const char *name, const oeb_s3 *s, const Box &b, Geo &g, bool first,
           const double **base);
int check3(const char *kname, const char *name, const oeb_s3f *s, const Box &b, Geo &g, bool first,
           const float **base);
// Validates a 1-D descriptor over the index range [lo, hi).  *base is the
// address of element 0 and *stride the element stride.
int check1(const char *kname, const char *name, const oeb_s1 *s, int64_t lo, int64_t hi, const double **base,
           int *stride);
int check1(const char *kname, const char *name, const oeb_s1f *s, int64_t lo, int64_t hi, const float **base,
           int *stride);
int check_domain(const char *kname, const oeb_domain *d);

// Tracks byte ranges of the arrays of one call so that an output overlapping
// any other array is rejected.
struct AliasSet {
  static constexpr int kMax = 32;
  const char *name[kMax];
  uintptr_t lo[kMax], hi[kMax];
  bool out[kMax];
  int n = 0;
  template <class T>
  void add(const char *nm, const T *base, const Box &b, const Geo &g, bool is_out) {
    add_range(nm, reinterpret_cast<uintptr_t>(base + (b.klo * g.sk + b.jlo * int64_t(g.sj) + b.ilo)),
              reinterpret_cast<uintptr_t>(base + ((b.khi - 1) * g.sk + (b.jhi - 1) * int64_t(g.sj) + (b.ihi - 1)) + 1),
              is_out);
  }
  template <class T>
  void add1(const char *nm, const T *base, int64_t lo_i, int64_t hi_i, int stride) {
    add_range(nm, reinterpret_cast<uintptr_t>(base + lo_i * stride),
              reinterpret_cast<uintptr_t>(base + (hi_i - 1) * stride + 1), false);
  }
  bool overflow = false;  // more than kMax arrays were added
  void add_range(const char *nm, uintptr_t lo_b, uintptr_t hi_b, bool is_out);
  int check(const char *kname) const;
};

// Converts the state of the last launch into a status.
int launch_status(const char *kname);

// Compute units of the current device (queried once per device).
int cu_count();

// Levels per workgroup in [lo, hi] for a tile launch of tiles_per_level
// workgroups per level chunk, wg_per_cu of which fit one CU at a time: the
// largest whose workgroup count fills the last round of the device's CUs
// best (ties to the larger chunk); dflt when none fills 90% of it.  A launch
// whose last round is half empty leaves half the CUs idle for a whole
// workgroup's time.
int balanced_kchunk(int64_t tiles_per_level, int nz, int wg_per_cu, int lo, int hi, int dflt);

// Fails with OEB_ERR_DOMAIN when a launch would need more workgroups than one
// grid dimension holds.
int check_tiles(const char *kname, const Tiles &t, int64_t ni, int64_t nj, int nz, int tile_i, int tile_j);



#define OEB_TRY(expr)              \
  do {                             \
    const int oeb_s_ = (expr);     \
    if (oeb_s_ != OEB_OK) return oeb_s_; \
  } while (0)

}  // namespace oeb
