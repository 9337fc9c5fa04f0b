// libshci_hip.so: the SHCI variational hot path written for MI355X (gfx950).
//
//   * Hamiltonian build (reference: src/solver/hamiltonian.h:87-598): packed
//     uint64 alpha/beta strings, radix-sorted into (alpha, beta) order; alpha and
//     beta groups; single-excitation partner lists of unique alpha strings from
//     sorted "minus one electron" keys; one wavefront per row tests candidates
//     with XOR/popcount, evaluates Slater-Condon elements in FP64 with the
//     reference's exact operation order (bit-identical values, so the
//     |H| >= 1e-12 cut keeps the same pattern), and writes a full-symmetric CSR
//     (off-diagonal; FP64 diagonal kept apart) in two passes: count -> scan -> fill.
//   * SpMV (reference: src/solver/sparse_matrix.cc:28-56): full-symmetric CSR,
//     L lanes per row (L chosen from the mean row length), shuffle reductions, no
//     atomics; FP64, FP32 (FP32 values + FP32 row accumulation, FP64 diagonal) and
//     FP32-values/FP64-accumulation variants, plus a 4-vector block product.
//   * Davidson (reference: src/solver/davidson.cc:8-196) with device vectors,
//     device-resident Gram-Schmidt scalars and a host Jacobi Rayleigh-Ritz, in
//     FP64 or mixed precision.
//   * RDM adjoint (reference: src/chem/rdm.cc:12-131, 374-395, 435-657) over the
//     same CSR with FP64 atomics.
//
// Everything runs on the context's single stream.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "shci_hip.h"

namespace {

constexpr double kEps = 1.0e-12;  // src/util.h:50
constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Device buffers and context
// ---------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

template <class T>
inline T* P(const DevBuf& b) {
  return static_cast<T*>(b.p);
}

struct SubspaceMat;

}  // namespace

struct shci_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  size_t mem_in_use = 0;

  // Integrals.
  uint32_t n_orb = 0;
  double core = 0.0;
  bool have_ints = false;
  DevBuf h1, eri, tD, tX;
  std::vector<double> h_h1, h_eri;

  // Determinants in caller order and in internal (alpha, beta)-sorted order.
  uint64_t n = 0;
  DevBuf up_c, dn_c;
  DevBuf up_s, dn_s, perm, iperm;
  uint32_t n_up = 0, n_dn = 0;
  uint32_t n_ugroups = 0, n_dgroups = 0;
  DevBuf up_gid, up_gstart, dn_gid, dn_gstart, dnord_idx, dnord_up;
  // "Minus one electron" tables of the unique alpha / beta strings and the
  // mixed-double segments ((alpha-1, beta-1) cores) of the determinants.
  uint32_t r_alpha = 0, r_beta = 0;  // distinct minus-one strings
  DevBuf m1a_id, m1b_id;             // [group * n_e + k] -> minus-one string id
  uint64_t n_dm1 = 0, n_dm1_runs = 0;
  DevBuf dm1_vals, dm1_rstart, dm1_seg;  // sorted (det * K + slot); run starts; slot -> run
  DevBuf flag;                           // device status words (fill overflow)

  // Matrix.
  bool built = false;
  uint32_t build_flags = 0;
  DevBuf diag, cnt, row_ptr, row_end, col, v64, v32, len;
  uint64_t nnz = 0;        // stored off-diagonal entries (full symmetric)
  uint64_t n_dropped = 0;  // candidates removed by the |H| >= 1e-12 cut
  uint64_t nnz_alloc = 0;  // allocated (padded) CSR slots
  int lanes = 16;

  // Scratch.
  DevBuf tmp, ka, kb, va, vb, scal, part;
  DevBuf xv, yv, x32, y32;
  DevBuf dav;  // Davidson vectors

  // Segment layout (SHCI_LAYOUT_SEGMENTS): every connected pair stored once,
  // as a value slot of the segment (shared core) that owns it.
  bool segl = false;
  uint32_t ra = 0, rb = 0;              // electrons the A / B cores remove (min(n, 2))
  uint32_t r2_alpha = 0, r2_beta = 0;   // distinct minus-two strings
  DevBuf mA_id, mB_id;                  // [group * C(n_e, 2) + pair] -> minus-two string id
  uint64_t n_seg = 0, n_mem = 0, n_slots = 0, n_pairs = 0;
  uint64_t seg_cs[28] = {};             // segment ranges: size class x type (C, A, B)
  uint64_t seg_kept[27] = {};           // kept (owned, |H| >= cut) pairs of each range, last fill
  uint32_t seg_maxm[27] = {};           // largest segment of each range
  DevBuf seg_desc, seg_type, seg_core, seg_mem, sv32, sv64;
  DevBuf dnpos;                         // internal det -> position in beta order
  DevBuf xb, yb;                        // beta-order x and partial y of the A segments
  // Tiles of the segments with m <= 64: runs of one spin group's segments
  // whose x and y a workgroup keeps in LDS (see k_seg_spmv_tiles).
  uint64_t n_tiles = 0;
  DevBuf tile_rb, tile_re;              // [tile * 6 + class - 3] -> segment range of the class run
  DevBuf tile_u0, tile_nu;              // first entry and size of the tile's det table
  DevBuf tile_tab;                      // tile det tables (internal det ids)
  DevBuf seg_loc;                       // membership -> index in its tile's table (build only)
  DevBuf tile_roff, tile_rec;           // per-tile record offset (16-byte units) and the records
  DevBuf kept;                          // kept-pair counters of the segment fill
  // Time-reversal symmetry (SHCI_TIME_SYM): the matrix is built over the
  // expanded list E (each det, then its reverse when up != dn); calls take and
  // return vectors in the reduced basis (one entry per det).
  bool ts = false;
  uint64_t n_red = 0;                   // dets of the reduced basis
  std::vector<uint64_t> ts_e;           // reduced det -> position in E (caller order)
  std::vector<uint8_t> ts_pair;         // 1: up != dn (E holds the reverse next)
  DevBuf ts_rev;                        // internal E index -> internal index of its reverse
  DevBuf ts_diag;                       // reduced diagonal of H_ts (caller order)
  // Heat-bath selection: the singles and double-excitation queues of the
  // current integrals (rebuilt when the integrals or the electron count
  // change) and the new determinants of the last sweep.
  uint32_t sel_ne = 0;                  // electron count of the queues (0 = not built)
  DevBuf sq_S, sq_r, sq_len, sq_max;    // n singles lists of n entries: S descending, target r
  DevBuf hq_H, hq_rs, hq_len, hq_max;   // n*n double lists of n*n entries: |H| descending, (r, s)
  DevBuf sel_up, sel_dn;                // last sweep's new dets, (up, dn) ascending
  uint64_t sel_n = 0;
  float sel_ms[4] = {};                 // last sweep: set-up, walk, repeats + unique, output
  uint32_t pt_batches = 0;              // batches of the last PT2 pass
  // Released device blocks kept for reuse (builds allocate and drop the same
  // temporaries every time; hipMalloc/hipFree of GB blocks per build would
  // cost more than the kernels).  Freed on memory pressure, clear and destroy.
  std::vector<DevBuf> cache;
  // Blocks of one chunk (1 GiB) or more: physical chunks mapped into a
  // reserved address range, by start address; unmapped chunks wait in
  // `pool` for the next block (see ensure).
  struct ChunkMap {
    size_t reserved = 0;                                // chunks of address range
    std::vector<hipMemGenericAllocationHandle_t> used;  // mapped chunks, in address order
  };
  std::unordered_map<void*, ChunkMap> chunked;
  std::vector<hipMemGenericAllocationHandle_t> pool;
  size_t chunk = 1ull << 30;  // chunk size (SHCI_CHUNK_MIB at context creation, 1 GiB by default)

  // Ritz-vector capture (shci_capture_ritz): the ground-state Ritz vector of
  // every Rayleigh-Ritz step of the latest Davidson call, widened to FP64, in
  // internal order.
  DevBuf cap;
  uint32_t cap_max = 0, cap_n = 0;
  // Device address of this device's copy of the allocation-free product's
  // scratch array (hipGetSymbolAddress at context creation).
  unsigned char* noalloc_scratch = nullptr;

  hipEvent_t ev[9] = {};
  float phase_ms[8] = {};
};

namespace {

int fail(shci_ctx* c, int code, const std::string& msg) {
  c->err = msg;
  return code;
}

#define HIPC(...)                                                                         \
  do {                                                                                    \
    hipError_t _e = (__VA_ARGS__);                                                        \
    if (_e != hipSuccess) {                                                               \
      return fail(ctx, _e == hipErrorOutOfMemory ? SHCI_ERR_OOM : SHCI_ERR_HIP,           \
                  std::string(#__VA_ARGS__) + ": " + hipGetErrorString(_e));              \
    }                                                                                     \
  } while (0)

#define CHECK_RC(...)        \
  do {                       \
    int _rc = (__VA_ARGS__); \
    if (_rc != SHCI_OK) return _rc; \
  } while (0)

// Gives b back to the context's block cache (its memory stays counted in use).
void release(shci_ctx* ctx, DevBuf& b) {
  if (b.p) ctx->cache.push_back(b);
  b.p = nullptr;
  b.bytes = 0;
}

// SHCI_LOG_ALLOC=1: every device allocation and cache release on stderr.
bool log_alloc() {
  static const bool on = getenv("SHCI_LOG_ALLOC") && atoi(getenv("SHCI_LOG_ALLOC")) != 0;
  return on;
}

// ---------------------------------------------------------------------------
// Large blocks: physical chunks kept by the context.
//
// Device memory that a process has used and returned to the driver is slow
// to get back: on MI355X a hipMalloc or hipMemCreate that needs such memory
// costs ~40 ms per GiB, against well under a millisecond for memory never
// used.  A growing det list asks for larger blocks at every build (a 1e7-det
// FP64 value store is 156 GB), so freeing and reallocating them cost seconds
// per build.  Blocks of at least one chunk are therefore made of physical
// chunks (hipMemCreate; 1 GiB, or SHCI_CHUNK_MIB at context creation) mapped
// into a reserved address range.  A block that no longer fits is resized in
// place: chunks are mapped at its end (the range reserves room to grow) or
// unmapped from it into the context's pool; a block given up altogether
// returns its chunks to the pool.  Mapping and unmapping cost ~0.01 and
// ~0.25 ms per 1 GiB chunk.  The pool goes back to the driver only on
// shci_clear, shci_ctx_destroy, or when a smaller allocation finds no memory.
// ---------------------------------------------------------------------------

bool is_chunked(const shci_ctx* ctx, const void* p) { return p && ctx->chunked.count(const_cast<void*>(p)) != 0; }

// Unmaps the chunks of block p past the first `keep` into the pool.  Kernels
// still queued on the stream may use them, so the stream is drained first.
int chunks_unmap(shci_ctx* ctx, void* p, size_t keep) {
  shci_ctx::ChunkMap& m = ctx->chunked.at(p);
  if (m.used.size() <= keep) return SHCI_OK;
  HIPC(hipStreamSynchronize(ctx->stream));
  if (log_alloc())
    fprintf(stderr, "[shci] unmapped %zu chunks; block %zu chunks, pool %zu chunks\n", m.used.size() - keep, keep,
            ctx->pool.size() + m.used.size() - keep);
  while (m.used.size() > keep) {
    const size_t i = m.used.size() - 1;
    HIPC(hipMemUnmap(static_cast<char*>(p) + i * ctx->chunk, ctx->chunk));
    ctx->pool.push_back(m.used[i]);
    m.used.pop_back();
  }
  return SHCI_OK;
}

// Unmaps block p entirely and frees its address range.
int chunks_drop(shci_ctx* ctx, void* p) {
  CHECK_RC(chunks_unmap(ctx, p, 0));
  const size_t reserved = ctx->chunked.at(p).reserved;
  ctx->chunked.erase(p);
  HIPC(hipMemAddressFree(p, reserved * ctx->chunk));
  return SHCI_OK;
}

// Gives up to `count` of the pool's chunks (all by default) back to the driver.
void pool_release(shci_ctx* ctx, size_t count = SIZE_MAX) {
  const size_t k = std::min(count, ctx->pool.size());
  for (size_t i = 0; i < k; i++) {
    (void)hipMemRelease(ctx->pool.back());
    ctx->pool.pop_back();
  }
  if (log_alloc() && k) fprintf(stderr, "[shci] released %zu pooled chunks\n", k);
  ctx->mem_in_use -= k * ctx->chunk;
}

// Frees the cached blocks: hipMalloc blocks go back to the driver, chunked
// blocks to the pool.
int cache_trim(shci_ctx* ctx) {
  size_t total = 0;
  int rc = SHCI_OK;
  for (DevBuf& c : ctx->cache) {
    if (is_chunked(ctx, c.p)) {
      const int r = chunks_drop(ctx, c.p);
      if (r != SHCI_OK) rc = r;
    } else {
      (void)hipFree(c.p);
      ctx->mem_in_use -= c.bytes;
    }
    total += c.bytes;
  }
  if (log_alloc() && !ctx->cache.empty())
    fprintf(stderr, "[shci] cache trim: %zu blocks, %.2f GB\n", ctx->cache.size(), total / 1e9);
  ctx->cache.clear();
  return rc;
}

// Device memory the context could hand out: free memory, its cached blocks
// and its unmapped chunks.
size_t available_bytes(shci_ctx* ctx) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    free_b = 0;
  }
  for (const DevBuf& c : ctx->cache) free_b += c.bytes;
  return free_b + ctx->pool.size() * ctx->chunk;
}

// Maps k more chunks at the end of block p: chunks of the pool first, then
// new ones; when the device has no memory left for them, the cached chunked
// blocks are taken apart (largest first: a replaced store rather than the
// working buffers the same build asks for again), then the cached hipMalloc
// blocks freed.
int chunks_map(shci_ctx* ctx, void* p, size_t k) {
  if (k == 0) return SHCI_OK;
  hipMemAllocationProp prop = {};
  prop.type = hipMemAllocationTypePinned;
  prop.location.type = hipMemLocationTypeDevice;
  prop.location.id = ctx->device;
  size_t created = 0;
  const auto t0 = std::chrono::steady_clock::now();
  bool trimmed = false;
  while (ctx->pool.size() < k) {
    hipMemGenericAllocationHandle_t h;
    const hipError_t e = hipMemCreate(&h, ctx->chunk, &prop, 0);
    if (e == hipSuccess) {
      ctx->pool.push_back(h);
      ctx->mem_in_use += ctx->chunk;
      created++;
      continue;
    }
    (void)hipGetLastError();
    if (e == hipErrorOutOfMemory) {
      size_t big = ctx->cache.size();
      for (size_t i = 0; i < ctx->cache.size(); i++)
        if (is_chunked(ctx, ctx->cache[i].p) && (big == ctx->cache.size() || ctx->cache[i].bytes > ctx->cache[big].bytes))
          big = i;
      if (big < ctx->cache.size()) {
        void* q = ctx->cache[big].p;
        ctx->cache.erase(ctx->cache.begin() + big);
        CHECK_RC(chunks_drop(ctx, q));
        continue;
      }
      if (!trimmed && !ctx->cache.empty()) {
        CHECK_RC(cache_trim(ctx));
        trimmed = true;
        continue;
      }
    }
    return fail(ctx, e == hipErrorOutOfMemory ? SHCI_ERR_OOM : SHCI_ERR_HIP,
                std::string("hipMemCreate(chunk) with ") + std::to_string(ctx->chunked.at(p).used.size()) +
                    " of a block's chunks mapped and " + std::to_string(k - ctx->pool.size()) +
                    " more needed: " + hipGetErrorString(e));
  }
  shci_ctx::ChunkMap& m = ctx->chunked.at(p);
  if (m.used.size() + k > m.reserved) return fail(ctx, SHCI_ERR_HIP, "chunked block: address range exceeded");
  const size_t first = m.used.size();
  for (size_t i = 0; i < k; i++) {
    const hipMemGenericAllocationHandle_t h = ctx->pool.back();
    HIPC(hipMemMap(static_cast<char*>(p) + (first + i) * ctx->chunk, ctx->chunk, 0, h, 0));
    ctx->pool.pop_back();
    m.used.push_back(h);
  }
  hipMemAccessDesc acc = {};
  acc.location.type = hipMemLocationTypeDevice;
  acc.location.id = ctx->device;
  acc.flags = hipMemAccessFlagsProtReadWrite;
  HIPC(hipMemSetAccess(static_cast<char*>(p) + first * ctx->chunk, k * ctx->chunk, &acc, 1));
  if (log_alloc())
    fprintf(stderr, "[shci] mapped %zu chunks (%zu new) in %.1f ms; block %zu chunks, pool %zu chunks\n", k, created,
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(), m.used.size(),
            ctx->pool.size());
  return SHCI_OK;
}

// b holds at least `bytes`: kept if large enough, else replaced.
// Blocks of at least one chunk are chunked: b itself is extended in place if
// its address range has room, else the smallest cached chunked block of
// n .. n + n/8 + 1 chunks is taken as it is, else a new range of twice the
// chunks (room to grow) is reserved and mapped.  Smaller blocks: the smallest cached block of
// bytes .. 9/8 bytes + 64 MiB (a close fit, so that the same sequence of
// requests finds the same blocks build after build), else hipMalloc (if the
// device is full, the cache and then the pool are freed and it is retried).
int ensure(shci_ctx* ctx, DevBuf& b, size_t bytes) {
  if (bytes <= b.bytes && b.p) return SHCI_OK;
  if (bytes >= ctx->chunk) {
    const size_t n = (bytes + ctx->chunk - 1) / ctx->chunk;
    if (is_chunked(ctx, b.p) && ctx->chunked.at(b.p).reserved >= n) {
      CHECK_RC(chunks_map(ctx, b.p, n - ctx->chunked.at(b.p).used.size()));
      b.bytes = n * ctx->chunk;
      return SHCI_OK;
    }
    release(ctx, b);
    // the smallest cached chunked block of n .. n + n/8 + 1 chunks, as it is
    size_t fit = ctx->cache.size();
    for (size_t i = 0; i < ctx->cache.size(); i++) {
      void* q = ctx->cache[i].p;
      if (!is_chunked(ctx, q)) continue;
      const size_t u = ctx->chunked.at(q).used.size();
      if (u >= n && u <= n + n / 8 + 1 && (fit == ctx->cache.size() || ctx->cache[i].bytes < ctx->cache[fit].bytes))
        fit = i;
    }
    if (fit < ctx->cache.size()) {
      b = ctx->cache[fit];
      ctx->cache.erase(ctx->cache.begin() + fit);
      return SHCI_OK;
    }
    void* p = nullptr;
    const size_t reserved = 2 * n;
    HIPC(hipMemAddressReserve(&p, reserved * ctx->chunk, ctx->chunk, nullptr, 0));
    ctx->chunked[p].reserved = reserved;
    const int rc = chunks_map(ctx, p, n);
    if (rc != SHCI_OK) {
      (void)chunks_drop(ctx, p);
      return rc;
    }
    b.p = p;
    b.bytes = n * ctx->chunk;
    return SHCI_OK;
  }
  release(ctx, b);
  if (bytes == 0) return SHCI_OK;
  size_t best = ctx->cache.size();
  for (size_t i = 0; i < ctx->cache.size(); i++) {
    const size_t cb = ctx->cache[i].bytes;
    if (cb >= bytes && cb <= bytes + bytes / 8 + (64ull << 20) &&
        (best == ctx->cache.size() || cb < ctx->cache[best].bytes))
      best = i;
  }
  if (best < ctx->cache.size()) {
    b = ctx->cache[best];
    ctx->cache.erase(ctx->cache.begin() + best);
    return SHCI_OK;
  }
  const auto t_alloc = std::chrono::steady_clock::now();
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e == hipErrorOutOfMemory && !ctx->cache.empty()) {
    (void)hipGetLastError();
    CHECK_RC(cache_trim(ctx));
    e = hipMalloc(&b.p, bytes);
  }
  if (e == hipErrorOutOfMemory && !ctx->pool.empty()) {
    (void)hipGetLastError();
    pool_release(ctx, bytes / ctx->chunk + 2);
    e = hipMalloc(&b.p, bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    return fail(ctx, e == hipErrorOutOfMemory ? SHCI_ERR_OOM : SHCI_ERR_HIP,
                std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
  }
  b.bytes = bytes;
  ctx->mem_in_use += bytes;
  if (log_alloc())
    fprintf(stderr, "[shci] hipMalloc %.3f GB in %.1f ms\n", bytes / 1e9,
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_alloc).count());
  return SHCI_OK;
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ int ctz64(uint64_t x) { return __builtin_ctzll(x); }

// Workgroup id remapped so each XCD (8 on MI355X, assigned round-robin) sees a
// contiguous range of rows: neighbouring rows share alpha groups in L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t G) {
  const uint32_t X = 8, per = G / X, rem = G % X, x = b % X, idx = b / X;
  return x < rem ? x * (per + 1) + idx : rem * (per + 1) + (x - rem) * per + idx;
}

__device__ __forceinline__ uint64_t c2(uint64_t a, uint64_t b) {
  return a > b ? (a * (a + 1)) / 2 + b : (b * (b + 1)) / 2 + a;
}

struct Ints {
  const double* __restrict__ h1;   // n*n
  const double* __restrict__ eri;  // combine4-packed
  const double* __restrict__ tD;   // tD[(p*n+q)*n+k] = (pq|kk)
  const double* __restrict__ tX;   // tX[(p*n+q)*n+k] = (pk|kq)
  int n;
  double core;
};

__device__ __forceinline__ double eri4(const Ints& I, int p, int q, int r, int s) {
  return I.eri[c2(c2(p, q), c2(r, s))];
}

__device__ __forceinline__ int below(uint64_t s, int o) { return __popcll(s & ((1ull << o) - 1ull)); }

// Single excitation p -> q in one spin, det_i on the left
// (src/chem/chem_system.cc:544-577, same operation order).
__device__ double single_elem(const Ints& I, uint64_t si, uint64_t sj, uint64_t oppo) {
  const uint64_t lo = si & ~sj, ro = sj & ~si;
  const int p = ctz64(lo), q = ctz64(ro);
  const int cnt = below(si, p) + below(sj, q);
  const double pf = (cnt & 1) ? -1.0 : 1.0;
  const int n = I.n;
  const double one = pf * I.h1[p * n + q];
  const double* D = I.tD + (size_t)(p * n + q) * n;
  const double* X = I.tX + (size_t)(p * n + q) * n;
  double e = 0.0;
  // Same-spin orbitals k != p in ascending order, then the opposite-spin
  // orbitals not among them, four at a time: the table loads of a group are
  // issued before its additions, which keep the reference's order.
  const uint64_t m0 = si & ~(1ull << p);
  for (uint64_t m = m0; m;) {
    int k[4];
    bool v[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      v[t] = m != 0;
      k[t] = v[t] ? ctz64(m) : 0;
      m &= m - 1;
    }
    double xv[4], dv[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      xv[t] = X[k[t]];
      dv[t] = D[k[t]];
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
      if (v[t]) {
        e -= xv[t];
        if ((oppo >> k[t]) & 1ull) {
          e += 2 * dv[t];
        } else {
          e += dv[t];
        }
      }
    }
  }
  for (uint64_t op = oppo & ~m0; op;) {
    int k[4];
    bool v[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      v[t] = op != 0;
      k[t] = v[t] ? ctz64(op) : 0;
      op &= op - 1;
    }
    double dv[4];
#pragma unroll
    for (int t = 0; t < 4; t++) dv[t] = D[k[t]];
#pragma unroll
    for (int t = 0; t < 4; t++)
      if (v[t]) e += dv[t];
  }
  e *= pf;
  return one + e;
}

// Same-spin double (src/chem/chem_system.cc:581-598).
__device__ double double_same(const Ints& I, uint64_t si, uint64_t sj) {
  const uint64_t lo = si & ~sj, ro = sj & ~si;
  const int i1 = ctz64(lo), i2 = ctz64(lo & (lo - 1));
  const int j1 = ctz64(ro), j2 = ctz64(ro & (ro - 1));
  const int cnt = below(si, i1) + below(si, i2) + below(sj, j1) + below(sj, j2);
  double e = eri4(I, i1, j1, i2, j2) - eri4(I, i1, j2, i2, j1);
  e *= (cnt & 1) ? -1.0 : 1.0;
  return e;
}

// Opposite-spin double (src/chem/chem_system.cc:599-607).
__device__ double double_mixed(const Ints& I, uint64_t ui, uint64_t uj, uint64_t di, uint64_t dj) {
  const uint64_t lu = ui & ~uj, ru = uj & ~ui, ld = di & ~dj, rd = dj & ~di;
  const int pu = ctz64(lu), qu = ctz64(ru), pd = ctz64(ld), qd = ctz64(rd);
  const int cu = below(ui, pu) + below(uj, qu);
  const int cd = below(di, pd) + below(dj, qd);
  const int pf = ((cu & 1) ? -1 : 1) * ((cd & 1) ? -1 : 1);
  double e = eri4(I, pu, qu, pd, qd);
  e *= pf;
  return e;
}

// Diagonal (src/chem/chem_system.cc:444-457 and 495-542, same order).
__device__ double diag_elem(const Ints& I, uint64_t up, uint64_t dn) {
  const int n = I.n;
  double e1 = 0.0;
  for (uint64_t m = up; m; m &= m - 1) {
    const int k = ctz64(m);
    e1 += I.h1[k * n + k];
  }
  if (up == dn) {
    e1 *= 2;
  } else {
    for (uint64_t m = dn; m; m &= m - 1) {
      const int k = ctz64(m);
      e1 += I.h1[k * n + k];
    }
  }
  double direct = 0.0, exchange = 0.0;
  for (uint64_t a = up; a; a &= a - 1) {
    const int i = ctz64(a);
    const double* D = I.tD + (size_t)(i * n + i) * n;
    const double* X = I.tX + (size_t)(i * n + i) * n;
    for (uint64_t b = a & (a - 1); b; b &= b - 1) {
      const int j = ctz64(b);
      direct += D[j];
      exchange -= X[j];
    }
  }
  if (up == dn) {
    direct *= 2;
    exchange *= 2;
  } else {
    for (uint64_t a = dn; a; a &= a - 1) {
      const int i = ctz64(a);
      const double* D = I.tD + (size_t)(i * n + i) * n;
      const double* X = I.tX + (size_t)(i * n + i) * n;
      for (uint64_t b = a & (a - 1); b; b &= b - 1) {
        const int j = ctz64(b);
        direct += D[j];
        exchange -= X[j];
      }
    }
  }
  for (uint64_t a = up; a; a &= a - 1) {
    const int i = ctz64(a);
    const double* D = I.tD + (size_t)(i * n + i) * n;
    for (uint64_t b = dn; b; b &= b - 1) direct += D[ctz64(b)];
  }
  return e1 + (direct + exchange) + I.core;
}

// Any pair with excitation <= 2 (dispatch of chem_system.cc:409-442).
__device__ double elem_any(const Ints& I, uint64_t ui, uint64_t di, uint64_t uj, uint64_t dj) {
  const int cu = __popcll(ui ^ uj), cd = __popcll(di ^ dj);
  if (cu == 0 && cd == 0) return diag_elem(I, ui, di);
  if (cd == 0) return cu == 2 ? single_elem(I, ui, uj, di) : double_same(I, ui, uj);
  if (cu == 0) return cd == 2 ? single_elem(I, di, dj, ui) : double_same(I, di, dj);
  return double_mixed(I, ui, uj, di, dj);
}

// ---------------------------------------------------------------------------
// Sort / grouping kernels
// ---------------------------------------------------------------------------
__global__ void k_iota(uint32_t* a, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = (uint32_t)i;
}

__global__ void k_gather_u64(const uint64_t* __restrict__ src, const uint32_t* __restrict__ idx,
                             uint64_t* __restrict__ dst, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

__global__ void k_inverse_perm(const uint32_t* __restrict__ perm, uint32_t* __restrict__ iperm, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) iperm[perm[i]] = (uint32_t)i;
}

// flag[i] = 1 where a new run of equal keys starts.
__global__ void k_head_flags(const uint64_t* __restrict__ keys, uint32_t* __restrict__ flag, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// From inclusive scan of head flags: group id per element (scan-1) and group starts.
__global__ void k_group_starts(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ incl,
                               uint32_t* __restrict__ gid, uint32_t* __restrict__ gstart, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t g = incl[i] - 1;
  gid[i] = g;
  if (flag[i]) gstart[g] = (uint32_t)i;
  if (i == n - 1) gstart[g + 1] = (uint32_t)n;
}

// Run starts only (no per-element id): rstart[run] = first position, rstart[R] = n.
__global__ void k_run_starts(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ incl,
                             uint32_t* __restrict__ rstart, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag[i]) rstart[incl[i] - 1] = (uint32_t)i;
  if (i == n - 1) rstart[incl[i]] = (uint32_t)n;
}

// dn-ordered arrays: dnord_up[k] = up_s[dnord_idx[k]], dn_gid[dnord_idx[k]] = group of k.
__global__ void k_dn_order(const uint32_t* __restrict__ dnord_idx, const uint64_t* __restrict__ up_s,
                           const uint32_t* __restrict__ kgid, uint64_t* __restrict__ dnord_up,
                           uint32_t* __restrict__ dn_gid, uint64_t n) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = dnord_idx[k];
  dnord_up[k] = up_s[i];
  dn_gid[i] = kgid[k];
}

// Unique strings of a grouped, sorted key array: out[g] = keys[gstart[g]].
__global__ void k_unique_strings(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ gstart, uint32_t G,
                                 uint64_t* __restrict__ out) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < G) out[g] = keys[gstart[g]];
}

// Minus-one keys of unique spin strings: key = string with its k-th electron
// removed, val = g * n_e + k (the analogue of the reference's abm1 map,
// src/solver/hamiltonian.h:256-305).
__global__ void k_minus_one(const uint64_t* __restrict__ uniq, uint32_t G, uint32_t n_e, uint64_t* __restrict__ keys,
                            uint32_t* __restrict__ vals) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const uint64_t s = uniq[g];
  uint64_t m = s;
  for (uint32_t e = 0; e < n_e; e++) {
    const uint64_t bit = m & (~m + 1);
    m &= m - 1;
    keys[g * n_e + e] = s & ~bit;
    vals[g * n_e + e] = (uint32_t)(g * n_e + e);
  }
}

// id[vals[p]] = run id of sorted position p (incl = inclusive scan of head flags).
__global__ void k_scatter_run_id(const uint32_t* __restrict__ vals, const uint32_t* __restrict__ incl, uint64_t m,
                                 uint32_t* __restrict__ id) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < m) id[vals[p]] = incl[p] - 1;
}

// Mixed-double keys: for det i (alpha group g, beta group h) and electrons
// k1 < n_up, k2 < n_dn, key = (alpha-minus-one id << bb) | beta-minus-one id,
// val = i * K + k1 * n_dn + k2.  Two dets share a key iff they differ by one
// alpha and one beta excitation (or are equal in one spin): every
// alpha-single x beta-single connection (hamiltonian.h:482-532) shares exactly
// one key.
__global__ void k_dm1_keys(const uint32_t* __restrict__ up_gid, const uint32_t* __restrict__ dn_gid,
                           const uint32_t* __restrict__ m1a, const uint32_t* __restrict__ m1b, uint64_t n,
                           uint32_t nu, uint32_t nd, int bb, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t K = nu * nd;
  if (t >= n * K) return;
  const uint64_t i = t / K;
  const uint32_t slot = (uint32_t)(t - i * K);
  const uint32_t k1 = slot / nd, k2 = slot - k1 * nd;
  const uint64_t a = m1a[(uint64_t)up_gid[i] * nu + k1];
  const uint64_t b = m1b[(uint64_t)dn_gid[i] * nd + k2];
  keys[t] = (a << bb) | b;
  vals[t] = (uint32_t)t;
}

__global__ void k_widen(const uint32_t* __restrict__ a, int64_t* __restrict__ b, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// Input validation: every det within n_orb orbitals with the same electron
// counts as det 0 (the search and the integral tables assume both).
__global__ void k_validate(const uint64_t* __restrict__ up, const uint64_t* __restrict__ dn, uint64_t n,
                           uint64_t mask, int nu, int nd, unsigned int* __restrict__ bad) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t u = up[i], d = dn[i];
  if ((u & mask) || (d & mask) || __popcll(u) != nu || __popcll(d) != nd) atomicAdd(&bad[0], 1u);
}

// After the (alpha, beta) sort equal neighbours are duplicate determinants.
__global__ void k_duplicates(const uint64_t* __restrict__ up, const uint64_t* __restrict__ dn, uint64_t n,
                             unsigned int* __restrict__ bad) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 || i >= n) return;
  if (up[i] == up[i - 1] && dn[i] == dn[i - 1]) atomicAdd(&bad[1], 1u);
}

__global__ void k_diag(Ints I, const uint64_t* __restrict__ up, const uint64_t* __restrict__ dn,
                       double* __restrict__ diag, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) diag[i] = diag_elem(I, up[i], dn[i]);
}

// ---------------------------------------------------------------------------
// Row kernels: connection search (count / fill), one wavefront per row.
// ---------------------------------------------------------------------------
struct RowParams {
  uint64_t n;
  const uint64_t* __restrict__ up;
  const uint64_t* __restrict__ dn;
  const uint32_t* __restrict__ up_gid;
  const uint32_t* __restrict__ up_gstart;
  const uint32_t* __restrict__ dn_gid;
  const uint32_t* __restrict__ dn_gstart;
  const uint32_t* __restrict__ dnord_idx;
  const uint64_t* __restrict__ dnord_up;
  const uint32_t* __restrict__ dm1_vals;    // sorted det * K + slot
  const uint32_t* __restrict__ dm1_rstart;  // segment starts
  const uint32_t* __restrict__ dm1_seg;     // det * K + slot -> segment
  uint32_t K;                               // n_up * n_dn
  Ints I;
  uint32_t* __restrict__ cnt;
  const int64_t* __restrict__ row_ptr;
  int64_t* __restrict__ row_end;
  uint32_t* __restrict__ col;
  double* __restrict__ v64;
  float* __restrict__ v32;
  uint32_t* __restrict__ len;
  unsigned int* __restrict__ overflow;
  int keep_zeros;
};

template <bool FILL>
struct Emitter {
  uint64_t base;  // FILL: next output slot; COUNT: running count
  uint64_t end;   // FILL: end of this row's slots (writes beyond are dropped and flagged)
  __device__ __forceinline__ void hit_only(bool hit) { base += __popcll(__ballot(hit)); }
  __device__ __forceinline__ void emit(const RowParams& Pm, bool hit, uint32_t j, double v) {
    const bool keep = hit && (Pm.keep_zeros || fabs(v) >= kEps);
    const uint64_t m = __ballot(keep);
    if (keep) {
      const uint64_t pos = base + mbcnt(m);
      if (pos < end) {
        Pm.col[pos] = j;
        if (Pm.v64) Pm.v64[pos] = v;
        if (Pm.v32) Pm.v32[pos] = (float)v;
      } else {
        atomicOr(Pm.overflow, 1u);
      }
    }
    base += __popcll(m);
  }
};

template <bool FILL>
__global__ __launch_bounds__(256) void k_rows(RowParams Pm) {
  const uint32_t lane = lane_id();
  const uint64_t r = (uint64_t)xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  if (r >= Pm.n) return;
  const uint64_t up_i = Pm.up[r], dn_i = Pm.dn[r];
  Emitter<FILL> em;
  em.base = FILL ? (uint64_t)Pm.row_ptr[r] : 0;
  em.end = FILL ? (uint64_t)Pm.row_ptr[r + 1] : 0;

  // (A) same beta string: single/double alpha excitations (hamiltonian.h:430-442).
  {
    const uint32_t g = Pm.dn_gid[r];
    const uint32_t b = Pm.dn_gstart[g], e = Pm.dn_gstart[g + 1];
    for (uint32_t k0 = b; k0 < e; k0 += kWave) {
      const uint32_t k = k0 + lane;
      const bool valid = k < e;
      const uint64_t u = valid ? Pm.dnord_up[k] : up_i;
      const int c = __popcll(u ^ up_i);
      const bool hit = valid && (c == 2 || c == 4);
      if (!FILL) {
        em.hit_only(hit);
      } else {
        double v = 0.0;
        uint32_t j = 0;
        if (hit) {
          j = Pm.dnord_idx[k];
          v = (c == 2) ? single_elem(Pm.I, up_i, u, dn_i) : double_same(Pm.I, up_i, u);
        }
        em.emit(Pm, hit, j, v);
      }
    }
  }
  // (B) same alpha string: single/double beta excitations (hamiltonian.h:456-468).
  const uint32_t ga = Pm.up_gid[r];
  {
    const uint32_t b = Pm.up_gstart[ga], e = Pm.up_gstart[ga + 1];
    for (uint32_t k0 = b; k0 < e; k0 += kWave) {
      const uint32_t k = k0 + lane;
      const bool valid = k < e;
      const uint64_t d = valid ? Pm.dn[k] : dn_i;
      const int c = __popcll(d ^ dn_i);
      const bool hit = valid && (c == 2 || c == 4);
      if (!FILL) {
        em.hit_only(hit);
      } else {
        double v = 0.0;
        if (hit) v = (c == 2) ? single_elem(Pm.I, dn_i, d, up_i) : double_same(Pm.I, dn_i, d);
        em.emit(Pm, hit, k, v);
      }
    }
  }
  // (C) alpha single x beta single (hamiltonian.h:482-532): the members of
  // this det's K = n_up * n_dn mixed-double segments (dets sharing its
  // (alpha - 1 electron, beta - 1 electron) core), flattened across segments
  // so all 64 lanes stay busy.  A member differing in both spins is a mixed
  // double; members equal in one spin are singles found by (A)/(B).
  {
    const uint32_t K = Pm.K;
    const uint64_t sb = r * (uint64_t)K;
    for (uint32_t s0 = 0; s0 < K; s0 += kWave) {
      const uint32_t sl = s0 + lane;
      uint32_t st = 0, ln = 0;
      if (sl < K) {
        const uint32_t seg = Pm.dm1_seg[sb + sl];
        st = Pm.dm1_rstart[seg];
        ln = Pm.dm1_rstart[seg + 1] - st;
      }
      uint32_t incl = ln;
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d, kWave);
        if ((int)lane >= d) incl += y;
      }
      const uint32_t excl = incl - ln;
      const uint32_t total = __shfl(incl, kWave - 1, kWave);
      for (uint32_t t0 = 0; t0 < total; t0 += kWave) {
        const uint32_t t = t0 + lane;
        const bool valid = t < total;
        int lo = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
          const int cand = lo + step;
          const uint32_t ec = __shfl(excl, cand < kWave ? cand : kWave - 1, kWave);
          if (cand < kWave && ec <= t) lo = cand;
        }
        const uint32_t stl = __shfl(st, lo, kWave);
        const uint32_t exl = __shfl(excl, lo, kWave);
        uint32_t j = 0;
        uint64_t uj = up_i, dj = dn_i;
        if (valid) {
          j = Pm.dm1_vals[stl + (t - exl)] / K;
          uj = Pm.up[j];
          dj = Pm.dn[j];
        }
        const bool hit = valid && uj != up_i && dj != dn_i;
        if (!FILL) {
          em.hit_only(hit);
        } else {
          double v = 0.0;
          if (hit) v = double_mixed(Pm.I, up_i, uj, dn_i, dj);
          em.emit(Pm, hit, j, v);
        }
      }
    }
  }
  if (FILL) {
    // Padding slots (rounding to 4 and entries dropped by the |H| cut): a
    // zero-valued self reference, which every consumer skips or adds as 0.
    for (uint64_t k = em.base + lane; k < em.end; k += kWave) {
      Pm.col[k] = (uint32_t)r;
      if (Pm.v64) Pm.v64[k] = 0.0;
      if (Pm.v32) Pm.v32[k] = 0.0f;
    }
  }
  if (lane == 0) {
    if (FILL) {
      Pm.len[r] = (uint32_t)(em.base - (uint64_t)Pm.row_ptr[r]);
      Pm.row_end[r] = (int64_t)em.end;
    } else {
      Pm.cnt[r] = (uint32_t)((em.base + 3) & ~3ull);  // rows padded to a multiple of 4 entries
    }
  }
}

// Sum of per-row lengths (the stored off-diagonal count); `mask` selects the
// bits summed.
__global__ void k_sum_u32(const uint32_t* __restrict__ a, uint64_t n, unsigned long long* __restrict__ out,
                          uint32_t mask = 0xFFFFFFFFu) {
  __shared__ unsigned long long red[256];
  unsigned long long v = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) v += a[i] & mask;
  red[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, red[0]);
}

// Recompute values on a fixed pattern (update_existing_elems).
__global__ __launch_bounds__(256) void k_update_values(uint64_t n, Ints I, const uint64_t* __restrict__ up,
                                                       const uint64_t* __restrict__ dn, const int64_t* __restrict__ rp,
                                                       const int64_t* __restrict__ re, const uint32_t* __restrict__ col, double* __restrict__ v64,
                                                       float* __restrict__ v32, double* __restrict__ diag) {
  const uint64_t r = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const uint64_t ui = up[r], di = dn[r];
  if (lane_id() == 0) diag[r] = diag_elem(I, ui, di);
  for (int64_t k = rp[r] + lane_id(); k < re[r]; k += kWave) {
    const uint32_t j = col[k];
    if (j == r) continue;  // padding
    const double v = elem_any(I, ui, di, up[j], dn[j]);
    if (v64) v64[k] = v;
    if (v32) v32[k] = (float)v;
  }
}

// ---------------------------------------------------------------------------
// SpMV kernels: L lanes per row, 256 threads per block.
//   y_i = (d_i - shift) x_i + sum_j A_ij x_j
// VT: stored value type; XT: vector type; AT: row accumulator type.
// ---------------------------------------------------------------------------
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// Four stored values starting at a multiple of 4 (16-byte / 2x16-byte loads).
__device__ __forceinline__ void load4(const float* __restrict__ v, int64_t k, float (&a)[4]) {
  const f32x4 t = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(v + k));
  a[0] = t.x;
  a[1] = t.y;
  a[2] = t.z;
  a[3] = t.w;
}
__device__ __forceinline__ void load4(const double* __restrict__ v, int64_t k, double (&a)[4]) {
  const f64x2 t0 = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(v + k));
  const f64x2 t1 = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(v + k + 2));
  a[0] = t0.x;
  a[1] = t0.y;
  a[2] = t1.x;
  a[3] = t1.y;
}

// Rows start and end at multiples of 4 entries; each lane takes 4 consecutive
// entries per step (one 16-byte column load + the value load), two steps in
// flight, then the dependent x gathers.
template <int L, typename VT, typename XT, typename AT>
__global__ __launch_bounds__(256) void k_spmv(uint64_t n, const int64_t* __restrict__ rp,
                                              const int64_t* __restrict__ re, const uint32_t* __restrict__ col,
                                              const VT* __restrict__ val, const double* __restrict__ diag,
                                              double shift, const XT* __restrict__ x, XT* __restrict__ y) {
  const uint32_t sub = threadIdx.x & (L - 1);
  const uint64_t row = (uint64_t)xcd_remap(blockIdx.x, gridDim.x) * (256 / L) + threadIdx.x / L;
  AT acc0 = 0, acc1 = 0;
  if (row < n) {
    const int64_t b = rp[row], e = re[row];
    int64_t k = b + 4 * (int64_t)sub;
    for (; k + 4 * L < e; k += 8 * L) {
      const u32x4 c0 = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(col + k));
      const u32x4 c1 = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(col + k + 4 * L));
      VT a0[4], a1[4];
      load4(val, k, a0);
      load4(val, k + 4 * L, a1);
      const XT x0 = x[c0.x], x1 = x[c0.y], x2 = x[c0.z], x3 = x[c0.w];
      const XT x4 = x[c1.x], x5 = x[c1.y], x6 = x[c1.z], x7 = x[c1.w];
      acc0 += (AT)a0[0] * (AT)x0 + (AT)a0[1] * (AT)x1;
      acc1 += (AT)a0[2] * (AT)x2 + (AT)a0[3] * (AT)x3;
      acc0 += (AT)a1[0] * (AT)x4 + (AT)a1[1] * (AT)x5;
      acc1 += (AT)a1[2] * (AT)x6 + (AT)a1[3] * (AT)x7;
    }
    if (k < e) {
      const u32x4 c0 = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(col + k));
      VT a0[4];
      load4(val, k, a0);
      acc0 += (AT)a0[0] * (AT)x[c0.x] + (AT)a0[1] * (AT)x[c0.y];
      acc1 += (AT)a0[2] * (AT)x[c0.z] + (AT)a0[3] * (AT)x[c0.w];
    }
  }
  AT acc = acc0 + acc1;
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, L);
  if (row < n && sub == 0) {
    const double dx = (diag[row] - shift) * (double)x[row];
    y[row] = (XT)(dx + (double)acc);
  }
}

// Block product for K vectors stored interleaved (x[i*K + v]).
template <int L, int K>
__global__ __launch_bounds__(256) void k_spmm_f32(uint64_t n, const int64_t* __restrict__ rp,
                                                  const int64_t* __restrict__ re, const uint32_t* __restrict__ col,
                                                  const float* __restrict__ val, const double* __restrict__ diag,
                                                  double shift, const float* __restrict__ x, float* __restrict__ y) {
  const uint32_t sub = threadIdx.x & (L - 1);
  const uint64_t row = (uint64_t)xcd_remap(blockIdx.x, gridDim.x) * (256 / L) + threadIdx.x / L;
  float acc[K];
#pragma unroll
  for (int v = 0; v < K; v++) acc[v] = 0.f;
  if (row < n) {
    const int64_t b = rp[row], e = re[row];
    for (int64_t k = b + sub; k < e; k += L) {
      const uint32_t c = __builtin_nontemporal_load(&col[k]);
      const float a = __builtin_nontemporal_load(&val[k]);
      const float* xc = x + (size_t)c * K;
#pragma unroll
      for (int v = 0; v < K; v++) acc[v] += a * xc[v];
    }
  }
#pragma unroll
  for (int v = 0; v < K; v++) {
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) acc[v] += __shfl_xor(acc[v], o, L);
  }
  if (row < n && sub == 0) {
    const double d = diag[row] - shift;
#pragma unroll
    for (int v = 0; v < K; v++) y[(size_t)row * K + v] = (float)(d * (double)x[(size_t)row * K + v] + (double)acc[v]);
  }
}

// ---------------------------------------------------------------------------
// Vector kernels (Davidson).
// ---------------------------------------------------------------------------
constexpr int kDotBlocks = 1024;
constexpr int kMaxDots = 24;

struct DotArgs {
  const void* b[kMaxDots];
};

// Partial dot products of a against m vectors: part[j*kDotBlocks + block].
template <typename T>
__global__ __launch_bounds__(256) void k_dots_partial(const T* __restrict__ a, DotArgs B, int m, uint64_t n,
                                                      double* __restrict__ part) {
  __shared__ double red[256];
  for (int j = 0; j < m; j++) {
    const T* __restrict__ bj = static_cast<const T*>(B.b[j]);
    double s = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
      s += (double)a[i] * (double)bj[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[j * kDotBlocks + blockIdx.x] = red[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_dots_final(const double* __restrict__ part, int m, int nblk,
                                                    double* __restrict__ out) {
  __shared__ double red[256];
  for (int j = 0; j < m; j++) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 256) s += part[j * kDotBlocks + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
    __syncthreads();
  }
}

// v -= s[0] * w  (s a device scalar; arithmetic in FP64).
template <typename T>
__global__ void k_axpy_dev(T* __restrict__ v, const T* __restrict__ w, const double* __restrict__ s, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (T)((double)v[i] - s[0] * (double)w[i]);
}

// v *= 1/norm with norm a host value.
template <typename T>
__global__ void k_scale(T* __restrict__ v, double inv, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (T)((double)v[i] * inv);
}

// v = x / norm (copy with scaling), possibly converting precision.
template <typename TI, typename TO>
__global__ void k_copy_scale(const TI* __restrict__ x, TO* __restrict__ v, double inv, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (TO)((double)x[i] * inv);
}

// Diagonal preconditioner (davidson.cc:106-114): lam and diag in FP64.
template <typename T>
__global__ void k_precond(T* __restrict__ v, const T* __restrict__ Hw, const T* __restrict__ w,
                          const double* __restrict__ diag, double lam, double shift, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = lam - (diag[i] - shift);
  v[i] = fabs(d) < 1.0e-8 ? (T)0 : (T)(((double)Hw[i] - lam * (double)w[i]) / d);
}

struct RitzArgs {
  const void* v[kMaxDots];
  const void* hv[kMaxDots];
  double c[kMaxDots];
};

// w = sum_i c_i v_i, Hw = sum_i c_i Hv_i (davidson.cc:155-165), FP64 accumulation.
template <typename T, typename TO>
__global__ void k_ritz(RitzArgs R, int m, uint64_t n, TO* __restrict__ w, TO* __restrict__ hw) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double a = 0.0, b = 0.0;
  for (int i = 0; i < m; i++) {
    a += (double)static_cast<const T*>(R.v[i])[j] * R.c[i];
    if (hw) b += (double)static_cast<const T*>(R.hv[i])[j] * R.c[i];
  }
  w[j] = (TO)a;
  if (hw) hw[j] = (TO)b;
}

// caller-order <-> internal-order permutations.
template <typename TI, typename TO>
__global__ void k_to_internal(const TI* __restrict__ xc, const uint32_t* __restrict__ perm, TO* __restrict__ xi,
                              uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xi[i] = (TO)xc[perm[i]];
}
template <typename TI, typename TO>
__global__ void k_to_caller(const TI* __restrict__ xi, const uint32_t* __restrict__ perm, TO* __restrict__ xc,
                            uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xc[perm[i]] = (TO)xi[i];
}
// Time-reversal symmetry: y <- (y + R y) / 2 on the expanded list (R swaps
// each det with its reverse; rev[i] = i for up == dn).  The pair's smaller
// index writes both.
template <typename T>
__global__ void k_ts_sym(T* __restrict__ y, const uint32_t* __restrict__ rev, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t r = rev[i];
  if (r > i) {
    const T m = (y[i] + y[r]) * (T)0.5;
    y[i] = m;
    y[r] = m;
  }
}

// Diagonal of H_ts = P^T H P over the reduced basis (caller order):
// (H_dd + H_rr) / 2 + H_dr for a det with up != dn (r its reverse; H_dr is
// a mixed double when up and dn differ in one orbital, else 0), H_dd else.
// base_system.h:86-101 gives H_dd + H_rd: the same number up to rounding.
__global__ void k_ts_diag(Ints I, const uint64_t* __restrict__ up_c, const uint64_t* __restrict__ dn_c,
                          const uint64_t* __restrict__ e, const uint8_t* __restrict__ pair, uint64_t n,
                          double* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t u = up_c[e[i]], d = dn_c[e[i]];
  const double hdd = diag_elem(I, u, d);
  if (!pair[i]) {
    out[i] = hdd;
    return;
  }
  const double hrr = diag_elem(I, d, u);
  const double hdr = __popcll(u ^ d) == 2 ? double_mixed(I, u, d, d, u) : 0.0;
  out[i] = 0.5 * (hdd + hrr) + hdr;
}

// interleave k vectors (row-major k x n, caller order) into internal x[i*K+v].
__global__ void k_interleave(const double* __restrict__ X, const uint32_t* __restrict__ perm, float* __restrict__ xi,
                             uint32_t K, uint32_t Kp, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = perm[i];
  for (uint32_t v = 0; v < Kp; v++) xi[i * Kp + v] = v < K ? (float)X[(size_t)v * n + c] : 0.f;
}
__global__ void k_deinterleave(const float* __restrict__ yi, const uint32_t* __restrict__ perm, double* __restrict__ Y,
                               uint32_t K, uint32_t Kp, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = perm[i];
  for (uint32_t v = 0; v < K; v++) Y[(size_t)v * n + c] = (double)yi[i * Kp + v];
}
// Up to 4 internal-order FP32 vectors <-> one interleaved block (x[i*4+v]),
// for the Davidson's block products.
struct Vec4 {
  float* p[4];
};
__global__ void k_pack4(Vec4 src, uint32_t k, uint64_t n, float* __restrict__ xi) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (uint32_t v = 0; v < 4; v++) xi[i * 4 + v] = v < k ? src.p[v][i] : 0.f;
}
__global__ void k_unpack4(const float* __restrict__ yi, uint32_t k, uint64_t n, Vec4 dst) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (uint32_t v = 0; v < k; v++) dst.p[v][i] = yi[i * 4 + v];
}

// ---------------------------------------------------------------------------
// RDM adjoint kernel (src/chem/rdm.cc).  Each unordered connected pair (and
// each diagonal) is visited once: row r, entries with col > r.
// ---------------------------------------------------------------------------
struct RdmParams {
  uint64_t n;
  int n_orb;
  uint32_t n_states;
  const uint64_t* __restrict__ up;
  const uint64_t* __restrict__ dn;
  const int64_t* __restrict__ rp;
  const int64_t* __restrict__ re;
  const uint32_t* __restrict__ col;
  const double* __restrict__ c;  // n_states x n, internal order
  double* __restrict__ rdm1;
  double* __restrict__ rdm2;
  int pair1;  // the pairs also write the 1-RDM (its diagonal and singles)
};

__device__ __forceinline__ double cc(const RdmParams& R, uint64_t i, uint64_t j) {
  double v = 0.0;
  for (uint32_t s = 0; s < R.n_states; s++) v += R.c[s * R.n + i] * R.c[s * R.n + j];
  return v;
}

// write_in_2rdm (rdm.cc:644-657): packed a = p*n+s, b = q*n+r, stored when a >= b.
__device__ __forceinline__ void w2(const RdmParams& R, int p, int q, int r, int s, double f, double cij) {
  const uint64_t n = (uint64_t)R.n_orb;
  const uint64_t a = (uint64_t)p * n + s, b = (uint64_t)q * n + r;
  if (a >= b) atomicAdd(&R.rdm2[(a * (a + 1)) / 2 + b], f * cij);
}

// permfac_ccaa (rdm.cc:243-269): sign of c+_p c+_q c_r c_s |h>.
__device__ __forceinline__ double permfac_ccaa(uint64_t h, int p, int q, int r, int s) {
  int c = below(h, s);
  h &= ~(1ull << s);
  c += below(h, r);
  h &= ~(1ull << r);
  c += below(h, q);
  h |= 1ull << q;
  c += below(h, p);
  return (c & 1) ? -1.0 : 1.0;
}

__device__ __forceinline__ double pf_of(uint64_t l, uint64_t rr) {
  // diff(l, rr).permutation_factor (half_det.cc:120-206)
  int c = 0;
  for (uint64_t m = l & ~rr; m; m &= m - 1) c += below(l, ctz64(m));
  for (uint64_t m = rr & ~l; m; m &= m - 1) c += below(rr, ctz64(m));
  return (c & 1) ? -1.0 : 1.0;
}

// 1-RDM terms of a single excitation between det i (this) and det j
// (connected), both orientations (rdm.cc:105-124); nothing for other pairs.
__device__ __forceinline__ void rdm1_single(const RdmParams& R, uint64_t ut, uint64_t dt, uint64_t i, uint64_t uc,
                                            uint64_t dc, uint64_t j) {
  uint64_t st, sc;
  if (dt == dc && __popcll(ut ^ uc) == 2) {
    st = ut;
    sc = uc;
  } else if (ut == uc && __popcll(dt ^ dc) == 2) {
    st = dt;
    sc = dc;
  } else {
    return;
  }
  const int n = R.n_orb;
  const int a1 = ctz64(st & ~sc), a2 = ctz64(sc & ~st);
  const double v = pf_of(sc, st) * cc(R, i, j);
  atomicAdd(&R.rdm1[a1 * n + a2], v);
  atomicAdd(&R.rdm1[a2 * n + a1], v);
}

// One stored pair i != j (rdm.cc:435-642 with connected = (uc, dc) at j and
// this = (ut, dt) at i; the diagonal pairs are k_rdm_diag's).  The
// reference's 2-RDM runs over the Hamiltonian's connections and its 1-RDM
// over the det list; the pairs write the 1-RDM when every single of the list
// is among them (R.pair1), and k_rdm1 does it otherwise.
__device__ void rdm_pair(const RdmParams& R, uint64_t ut, uint64_t dt, uint64_t i, uint64_t uc, uint64_t dc,
                         uint64_t j) {
  const double cij = cc(R, i, j);
  if (R.pair1) rdm1_single(R, ut, dt, i, uc, dc, j);
  if (ut == uc) {
    const int nd = __popcll(dc ^ dt) / 2;
    if (nd == 1) {
      const int b1 = ctz64(dt & ~dc), b2 = ctz64(dc & ~dt);  // from (this), to (connected)
      const double pf = pf_of(dc, dt);
      for (uint64_t m = dt; m; m &= m - 1) {
        const int p = ctz64(m);
        if (p == b1 || p == b2) continue;
        double sf = permfac_ccaa(dt, p, b2, b1, p);
        w2(R, p, b2, b1, p, sf, cij);
        w2(R, b2, p, b1, p, -sf, cij);
        w2(R, p, b2, p, b1, -sf, cij);
        w2(R, b2, p, p, b1, sf, cij);
        sf = permfac_ccaa(dc, p, b1, b2, p);
        w2(R, p, b1, b2, p, sf, cij);
        w2(R, b1, p, b2, p, -sf, cij);
        w2(R, p, b1, p, b2, -sf, cij);
        w2(R, b1, p, p, b2, sf, cij);
      }
      const double sf1 = pf, sf2 = pf_of(dt, dc);
      for (uint64_t m = ut; m; m &= m - 1) {
        const int p = ctz64(m);
        w2(R, p, b2, b1, p, sf1, cij);
        w2(R, b2, p, p, b1, sf1, cij);
        w2(R, p, b1, b2, p, sf2, cij);
        w2(R, b1, p, p, b2, sf2, cij);
      }
    } else if (nd == 2) {
      const uint64_t ro = dt & ~dc, lo = dc & ~dt;
      const int b1 = ctz64(ro), b2 = ctz64(ro & (ro - 1)), b3 = ctz64(lo), b4 = ctz64(lo & (lo - 1));
      double sf = pf_of(dc, dt);
      w2(R, b3, b4, b2, b1, sf, cij);
      w2(R, b3, b4, b1, b2, -sf, cij);
      w2(R, b4, b3, b2, b1, -sf, cij);
      w2(R, b4, b3, b1, b2, sf, cij);
      sf = pf_of(dt, dc);
      w2(R, b1, b2, b4, b3, sf, cij);
      w2(R, b1, b2, b3, b4, -sf, cij);
      w2(R, b2, b1, b4, b3, -sf, cij);
      w2(R, b2, b1, b3, b4, sf, cij);
    }
    return;
  }
  const int na = __popcll(uc ^ ut) / 2;
  if (na == 1) {
    const int a1 = ctz64(ut & ~uc), a2 = ctz64(uc & ~ut);
    if (dt == dc) {
      const double pf = pf_of(uc, ut);
      for (uint64_t m = ut; m; m &= m - 1) {
        const int p = ctz64(m);
        if (p == a1 || p == a2) continue;
        double sf = permfac_ccaa(ut, p, a2, a1, p);
        w2(R, p, a2, a1, p, sf, cij);
        w2(R, p, a2, p, a1, -sf, cij);
        w2(R, a2, p, a1, p, -sf, cij);
        w2(R, a2, p, p, a1, sf, cij);
        sf = permfac_ccaa(uc, p, a1, a2, p);
        w2(R, p, a1, a2, p, sf, cij);
        w2(R, p, a1, p, a2, -sf, cij);
        w2(R, a1, p, a2, p, -sf, cij);
        w2(R, a1, p, p, a2, sf, cij);
      }
      const double sf1 = pf, sf2 = pf_of(ut, uc);
      for (uint64_t m = dt; m; m &= m - 1) {
        const int p = ctz64(m);
        w2(R, a2, p, p, a1, sf1, cij);
        w2(R, p, a2, a1, p, sf1, cij);
        w2(R, a1, p, p, a2, sf2, cij);
        w2(R, p, a1, a2, p, sf2, cij);
      }
    } else if (__popcll(dc ^ dt) == 2) {
      const int b1 = ctz64(dt & ~dc), b2 = ctz64(dc & ~dt);
      double sf = pf_of(uc, ut) * pf_of(dc, dt);
      w2(R, a2, b2, b1, a1, sf, cij);
      w2(R, b2, a2, a1, b1, sf, cij);
      sf = pf_of(ut, uc) * pf_of(dt, dc);
      w2(R, a1, b1, b2, a2, sf, cij);
      w2(R, b1, a1, a2, b2, sf, cij);
    }
  } else if (na == 2 && dt == dc) {
    const uint64_t ro = ut & ~uc, lo = uc & ~ut;
    const int a1 = ctz64(ro), a2 = ctz64(ro & (ro - 1)), a3 = ctz64(lo), a4 = ctz64(lo & (lo - 1));
    double sf = pf_of(uc, ut);
    w2(R, a3, a4, a2, a1, sf, cij);
    w2(R, a3, a4, a1, a2, -sf, cij);
    w2(R, a4, a3, a2, a1, -sf, cij);
    w2(R, a4, a3, a1, a2, sf, cij);
    sf = pf_of(ut, uc);
    w2(R, a1, a2, a4, a3, sf, cij);
    w2(R, a1, a2, a3, a4, -sf, cij);
    w2(R, a2, a1, a4, a3, -sf, cij);
    w2(R, a2, a1, a3, a4, sf, cij);
  }
}

__global__ __launch_bounds__(256) void k_rdm(RdmParams R) {
  const uint64_t r = (uint64_t)xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  if (r >= R.n) return;
  const uint64_t ut = R.up[r], dt = R.dn[r];
  for (int64_t k = R.rp[r] + lane_id(); k < R.re[r]; k += kWave) {
    const uint32_t j = R.col[k];
    if (j <= r) continue;
    rdm_pair(R, ut, dt, r, R.up[j], R.dn[j], j);
  }
}

// Export helpers: count, per caller row, the stored upper entries (diag + j>i).
__global__ __launch_bounds__(256) void k_export_count(uint64_t n, const int64_t* __restrict__ rp,
                                                      const int64_t* __restrict__ re, const uint32_t* __restrict__ col,
                                                      const uint32_t* __restrict__ perm,
                                                      unsigned long long* __restrict__ cnt_caller) {
  const uint64_t r = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const uint32_t ci = perm[r];
  uint32_t c = 0;
  for (int64_t k = rp[r] + lane_id(); k < re[r]; k += kWave) c += perm[col[k]] > ci ? 1u : 0u;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, kWave);
  if (lane_id() == 0) cnt_caller[ci] = c + 1;
}

__global__ __launch_bounds__(256) void k_export_fill(uint64_t n, const int64_t* __restrict__ rp,
                                                     const int64_t* __restrict__ re, const uint32_t* __restrict__ col, const double* __restrict__ v64,
                                                     const double* __restrict__ diag, const uint32_t* __restrict__ perm,
                                                     const unsigned long long* __restrict__ out_rp,
                                                     int64_t* __restrict__ out_col, double* __restrict__ out_val) {
  const uint64_t r = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const uint32_t ci = perm[r];
  uint64_t base = out_rp[ci];
  if (lane_id() == 0) {
    out_col[base] = ci;
    out_val[base] = diag[r];
  }
  base += 1;
  for (int64_t k0 = rp[r]; k0 < re[r]; k0 += kWave) {
    const int64_t k = k0 + lane_id();
    const bool valid = k < re[r];
    const uint32_t cj = valid ? perm[col[k]] : 0;
    const bool keep = valid && cj > ci;
    const uint64_t m = __ballot(keep);
    if (keep) {
      const uint64_t pos = base + mbcnt(m);
      out_col[pos] = cj;
      out_val[pos] = v64 ? v64[k] : 0.0;
    }
    base += __popcll(m);
  }
}


// ---------------------------------------------------------------------------
// Segment layout (SHCI_LAYOUT_SEGMENTS).
//
// Determinants are gathered into "segments" of dets that share a core:
//   C  an (alpha - 1, beta - 1) core: members differ by at most one alpha and
//      one beta excitation (the reference's alpha-single x beta-single search,
//      hamiltonian.h:482-532);
//   A  an alpha core two electrons short (one if n_up == 1) and one beta
//      string: members differ by an alpha single or double (the same-beta
//      scan, hamiltonian.h:430-442);
//   B  the same with the spins exchanged (the same-alpha scan,
//      hamiltonian.h:456-468).
// Every connected pair of distinct dets is owned by exactly one segment: a C
// segment owns the pairs that differ in both spins; an A segment owns its
// alpha doubles (their common part is the core) and the alpha singles whose
// lowest common orbital lies outside the core (a single shares n_up - 1 such
// cores); B likewise.  A segment of m members keeps m(m-1)/2 value slots in
// rotation order: step d = 1..m/2 holds the pairs (j, (j + d) mod m) for
// j < m, or for j < m/2 in the last step of an even m.  Slots of pairs the
// segment does not own, and of pairs under the |H| >= 1e-12 cut, hold 0.
// No column indices are stored: a lane group of S >= m lanes (m <= 64), or
// one wavefront with the member x staged in LDS (m > 64), walks the steps,
// applies each value in both directions (the partner's x and the backward
// value arrive by cross-lane permutes or from LDS) and adds each member's sum
// to y with one global atomic per member and segment.
// ---------------------------------------------------------------------------
constexpr uint32_t kSegMaxM = 4096;  // C: (n_orb-n_up+1)(n_orb-n_dn+1) <= 63*63; A, B: C(n_orb-n+2, 2) <= 2016
constexpr int kSegClasses = 9;       // m > 1024, > 256, > 64 (one wavefront each); lane groups of 64 .. 2
constexpr int kSegRanges = kSegClasses * 3;  // x segment type; range index = class * 3 + type
// Zero slots after the last value slot: the lane-group product reads a few
// rows past a segment's last step instead of clamping every address.
constexpr uint64_t kSlotPad = 4096;

enum : uint32_t { kSegC = 0, kSegA = 1, kSegB = 2 };

struct SegDesc {
  uint32_t moff;  // first member in the member array
  uint32_t m;     // members
  uint64_t voff;  // first value slot
};

__device__ __forceinline__ uint32_t seg_class(uint32_t m) {
  return m > 1024 ? 0 : m > 256 ? 1 : m > 64 ? 2 : m > 32 ? 3 : m > 16 ? 4 : m > 8 ? 5 : m > 4 ? 6 : m > 2 ? 7 : 8;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, kWave);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, kWave);
  return ((uint64_t)hi << 32) | lo;
}

// Does a segment of this type and core own the pair of dets (j, l)?
// n_rm: electrons the A (or B) core removes from its spin string.
__device__ __forceinline__ bool seg_owns(uint32_t type, uint64_t core, uint32_t n_rm, uint64_t uj, uint64_t dj,
                                         uint64_t ul, uint64_t dl) {
  if (type == kSegC) return uj != ul && dj != dl;
  const uint64_t sj = type == kSegA ? uj : dj, sl = type == kSegA ? ul : dl;
  if (__popcll(sj ^ sl) != 2 || n_rm < 2) return true;
  const uint64_t c = sj & sl;
  return ((c & (~c + 1ull)) & ~core) != 0;
}

// String s without the electrons of removal pattern k: the k-th electron
// (rm == 1) or the k-th electron pair in (lower, upper) lexicographic order
// (rm == 2), counting electrons from the lowest orbital.
__device__ __forceinline__ uint64_t drop_electrons(uint64_t s, uint32_t k, uint32_t rm) {
  if (rm == 1) {
    uint64_t m = s;
    for (uint32_t e = 0; e < k; e++) m &= m - 1;
    return s & ~(m & (~m + 1ull));
  }
  uint32_t c = 0;
  for (uint64_t a = s; a; a &= a - 1) {
    const uint64_t ba = a & (~a + 1ull);
    for (uint64_t b = a & (a - 1); b; b &= b - 1) {
      if (c == k) return s & ~ba & ~(b & (~b + 1ull));
      c++;
    }
  }
  return s;
}

// Minus-two keys of unique spin strings: key = string without an electron
// pair, val = g * C(n_e, 2) + pair index (the two-electron analogue of the
// abm1 map, src/solver/hamiltonian.h:256-305).
__global__ void k_minus_two(const uint64_t* __restrict__ uniq, uint32_t G, uint32_t n_e, uint64_t* __restrict__ keys,
                            uint32_t* __restrict__ vals) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G) return;
  const uint64_t s = uniq[g];
  const uint64_t P = (uint64_t)n_e * (n_e - 1) / 2;
  uint64_t k = 0;
  for (uint64_t a = s; a; a &= a - 1) {
    const uint64_t ba = a & (~a + 1ull);
    for (uint64_t b = a & (a - 1); b; b &= b - 1) {
      keys[g * P + k] = s & ~ba & ~(b & (~b + 1ull));
      vals[g * P + k] = (uint32_t)(g * P + k);
      k++;
    }
  }
}

struct SegKeyParams {
  uint64_t n;
  uint32_t PC, PA, PB;  // keys per det in the C, A, B regions
  uint32_t nu, nd, ra, rb;
  const uint32_t* __restrict__ up_gid;
  const uint32_t* __restrict__ dn_gid;
  const uint32_t* __restrict__ m1a;  // minus-one ids of unique alpha strings (C cores)
  const uint32_t* __restrict__ m1b;
  const uint32_t* __restrict__ mA;   // A core ids of unique alpha strings
  const uint32_t* __restrict__ mB;
  const uint64_t* __restrict__ up;
  const uint64_t* __restrict__ dn;
  const uint32_t* __restrict__ dnpos;  // internal det -> position in beta order
  int bC, bA, bB;  // widths of the low key fields
  int tag_shift;
  uint64_t eC, eA;      // first key index of the A and B regions
  uint64_t dC, dA, dB;  // ceil(2^64 / P) of PC, PA, PB (P >= 2), for seg_decode's divisions
};

// ceil(2^64 / d) for d >= 2 (0 otherwise).
inline uint64_t div_mul(uint32_t d) { return d >= 2 ? UINT64_MAX / d + 1 : 0; }

// floor(t / d) for t < 2^33: the high half of t * ceil(2^64 / d) is t / d
// plus less than 2^-31, which cannot reach the next integer since t / d
// falls short of it by at least 1/d (d < 2^31).  A multiply instead of the
// long integer division.
__device__ __forceinline__ uint64_t div_by(uint64_t t, uint32_t d, uint64_t mul) {
  return d >= 2 ? __umul64hi(t, mul) : t;
}

// Region, det and removal pattern of key index t (t < 2^31).
__device__ __forceinline__ void seg_decode(const SegKeyParams& K, uint64_t t, uint32_t& type, uint64_t& i,
                                           uint32_t& k) {
  if (t < K.eC) {
    type = kSegC;
    i = div_by(t, K.PC, K.dC);
    k = (uint32_t)(t - i * K.PC);
  } else if (t < K.eA) {
    type = kSegA;
    const uint64_t u = t - K.eC;
    i = div_by(u, K.PA, K.dA);
    k = (uint32_t)(u - i * K.PA);
  } else {
    type = kSegB;
    const uint64_t u = t - K.eA;
    i = div_by(u, K.PB, K.dB);
    k = (uint32_t)(u - i * K.PB);
  }
}

__global__ void k_seg_keys(SegKeyParams K, uint64_t total, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  uint32_t type, k;
  uint64_t i;
  seg_decode(K, t, type, i, k);
  uint64_t payload;
  if (type == kSegC) {
    const uint32_t k1 = k / K.nd, k2 = k - k1 * K.nd;
    payload = ((uint64_t)K.m1a[(uint64_t)K.up_gid[i] * K.nu + k1] << K.bC) | K.m1b[(uint64_t)K.dn_gid[i] * K.nd + k2];
  } else if (type == kSegA) {  // spin group in the high field: runs of one beta group are adjacent
    payload = ((uint64_t)K.dn_gid[i] << K.bA) | K.mA[(uint64_t)K.up_gid[i] * K.PA + k];
  } else {
    payload = ((uint64_t)K.up_gid[i] << K.bB) | K.mB[(uint64_t)K.dn_gid[i] * K.PB + k];
  }
  keys[t] = ((uint64_t)type << K.tag_shift) | payload;
  vals[t] = (uint32_t)t;
}

// Keys without the region tag, 32 bits wide, when every region's key fits:
// the three regions (C, A, B key indices) are then sorted separately.
__global__ void k_seg_keys32(SegKeyParams K, uint64_t total, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  uint32_t type, k;
  uint64_t i;
  seg_decode(K, t, type, i, k);
  uint64_t payload;
  if (type == kSegC) {
    const uint32_t k1 = k / K.nd, k2 = k - k1 * K.nd;
    payload = ((uint64_t)K.m1a[(uint64_t)K.up_gid[i] * K.nu + k1] << K.bC) | K.m1b[(uint64_t)K.dn_gid[i] * K.nd + k2];
  } else if (type == kSegA) {
    payload = ((uint64_t)K.dn_gid[i] << K.bA) | K.mA[(uint64_t)K.up_gid[i] * K.PA + k];
  } else {
    payload = ((uint64_t)K.up_gid[i] << K.bB) | K.mB[(uint64_t)K.dn_gid[i] * K.PB + k];
  }
  keys[t] = (uint32_t)payload;
  vals[t] = (uint32_t)t;
}

// Run heads of region-wise sorted 32-bit keys (a region start always opens a run).
__global__ void k_head_flags32(const uint32_t* __restrict__ keys, uint32_t* __restrict__ flag, uint64_t n,
                               uint64_t eC, uint64_t eA) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = (i == 0 || i == eC || i == eA || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// A member is kept when the run owns at least one pair of it with another
// member: members whose pairs all belong to other segments (a C core whose
// other members share its alpha or beta string, an alpha single owned by a
// different core) would only cost an x gather and a y update per product.
__global__ void k_seg_useful(SegKeyParams K, const uint32_t* __restrict__ vals_s, const uint32_t* __restrict__ incl,
                             const uint32_t* __restrict__ rstart, uint64_t M, uint32_t* __restrict__ use) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= M) return;
  const uint32_t r = incl[p] - 1;
  const uint32_t a = rstart[r], b = rstart[r + 1];
  uint32_t useful = 0;
  if (b - a >= 2) {
    uint32_t ty, k;
    uint64_t i;
    seg_decode(K, vals_s[p], ty, i, k);
    const uint64_t ui = K.up[i], di = K.dn[i];
    // the run's core: this member's string without its removal pattern k
    const uint64_t core = ty == kSegC ? 0ull : ty == kSegA ? drop_electrons(ui, k, K.ra) : drop_electrons(di, k, K.rb);
    const uint32_t nrm = ty == kSegA ? K.ra : K.rb;
    // four candidates per trip, so that their det loads are in flight together
    constexpr uint32_t Q = 4;
    for (uint32_t q0 = a; q0 < b && !useful; q0 += Q) {
      uint64_t uq[Q], dq[Q];
      bool ok[Q];
#pragma unroll
      for (uint32_t t = 0; t < Q; t++) {
        const uint32_t q = q0 + t;
        ok[t] = q < b && q != p;
        uint32_t tq, kq;
        uint64_t l;
        seg_decode(K, vals_s[q < b ? q : a], tq, l, kq);
        uq[t] = K.up[l];
        dq[t] = K.dn[l];
      }
#pragma unroll
      for (uint32_t t = 0; t < Q; t++)
        if (ok[t] && seg_owns(ty, core, nrm, ui, di, uq[t], dq[t])) useful = 1u;
    }
  }
  use[p] = useful;
}

// Kept members of run r (from the inclusive scan of the member flags) and the
// scan value before the run.
__device__ __forceinline__ uint32_t run_kept(const uint32_t* __restrict__ useinc, const uint32_t* __restrict__ rstart,
                                             uint32_t r, uint32_t& before) {
  const uint32_t a = rstart[r], b = rstart[r + 1];
  before = a ? useinc[a - 1] : 0u;
  return useinc[b - 1] - before;
}

// Runs of equal keys -> kept segments (>= 2 kept members).
__global__ void k_seg_keep(const uint32_t* __restrict__ rstart, const uint32_t* __restrict__ useinc, uint64_t R,
                           uint32_t* __restrict__ keep, unsigned int* __restrict__ bad) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > R) return;
  if (r == R) {
    keep[r] = 0;
    return;
  }
  if (rstart[r + 1] - rstart[r] > kSegMaxM) atomicOr(bad, 1u);
  uint32_t before;
  keep[r] = run_kept(useinc, rstart, (uint32_t)r, before) >= 2 ? 1u : 0u;
}

// Tiles.  The kept runs come out of the key sort in (type, spin group, core)
// order: C runs grouped by their alpha - 1 core, A runs by their beta string,
// B runs by their alpha string.  The segments of one group share members (a
// det sits in up to n_beta C segments of one alpha - 1 core and in up to
// C(n_alpha, 2) A segments of its beta string), so a workgroup that keeps
// the group's x and y in LDS gathers and updates each det once instead of
// once per segment.  A group's segments with m <= 64 are cut into tiles of
// fewer than kTileMax memberships: a segment starts a new tile when the
// memberships before it in its group reach the next multiple of kTileCap.
constexpr uint32_t kTileCap = 1920;
constexpr uint32_t kTileMax = 1984;   // > kTileCap + 64: memberships (and dets) of one tile
constexpr uint32_t kTileHash = 4096;  // LDS hash slots of the table build
static_assert(kTileMax % 4 == 0, "the table ranks read the keys four at a time");
static_assert(kTileHash >= 2 * kTileMax, "hash table below twice the tile size");

// Per kept run (index f in run order): spin-group key and the memberships it
// brings to its tile (m <= 64) or 0 (larger segments have no tile).
__global__ void k_seg_groups(SegKeyParams K, const uint32_t* __restrict__ vals_s, const uint32_t* __restrict__ rstart,
                             const uint32_t* __restrict__ useinc, const uint32_t* __restrict__ keep,
                             const uint32_t* __restrict__ sidx, uint64_t R, uint64_t* __restrict__ gkey,
                             uint32_t* __restrict__ w) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R || !keep[r]) return;
  uint32_t before;
  const uint32_t m = run_kept(useinc, rstart, (uint32_t)r, before);
  uint32_t ty, k;
  uint64_t i;
  seg_decode(K, vals_s[rstart[r]], ty, i, k);
  uint64_t g;
  if (ty == kSegC) g = K.m1a[(uint64_t)K.up_gid[i] * K.nu + k / K.nd];
  else if (ty == kSegA) g = K.dn_gid[i];
  else g = K.up_gid[i];
  const uint32_t f = sidx[r];
  gkey[f] = ((uint64_t)ty << 32) | g;
  w[f] = m <= 64 ? m : 0u;
}

// gflag[f] = run f opens a spin group.
__global__ void k_seg_gflags(const uint64_t* __restrict__ gkey, uint64_t S, uint32_t* __restrict__ gflag) {
  const uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= S) return;
  gflag[f] = (f == 0 || gkey[f] != gkey[f - 1]) ? 1u : 0u;
}

// gW[group] = memberships of the tiles before the group (W = inclusive scan of w).
__global__ void k_seg_gstart(const uint32_t* __restrict__ gflag, const uint32_t* __restrict__ gord,
                             const uint32_t* __restrict__ W, const uint32_t* __restrict__ w, uint64_t S,
                             uint32_t* __restrict__ gW) {
  const uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= S || !gflag[f]) return;
  gW[gord[f] - 1] = W[f] - w[f];
}

// tflag[f] = run f opens a tile.
__global__ void k_seg_tflags(const uint32_t* __restrict__ gflag, const uint32_t* __restrict__ gord,
                             const uint32_t* __restrict__ W, const uint32_t* __restrict__ w,
                             const uint32_t* __restrict__ gW, uint64_t S, uint32_t* __restrict__ tflag) {
  const uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= S) return;
  const uint32_t piece = (W[f] - w[f] - gW[gord[f] - 1]) / kTileCap;
  uint32_t open = gflag[f];
  if (!open) {
    const uint32_t prev = (W[f - 1] - w[f - 1] - gW[gord[f - 1] - 1]) / kTileCap;
    open = piece != prev ? 1u : 0u;
  }
  tflag[f] = open;
}

// Order key of a kept run: size class, type, then (lane-group classes) the
// tile, so that each class range lists the tiles' runs in tile order; largest
// first within a tile and within the one-wavefront classes.
__global__ void k_seg_compact(SegKeyParams K, const uint32_t* __restrict__ vals_s,
                              const uint32_t* __restrict__ rstart, const uint32_t* __restrict__ useinc,
                              const uint32_t* __restrict__ keep, const uint32_t* __restrict__ sidx,
                              const uint32_t* __restrict__ tileinc, uint64_t R, uint64_t* __restrict__ skey,
                              uint32_t* __restrict__ srun) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R || !keep[r]) return;
  uint32_t before;
  const uint32_t m = run_kept(useinc, rstart, (uint32_t)r, before);
  uint32_t ty, k;
  uint64_t i;
  seg_decode(K, vals_s[rstart[r]], ty, i, k);
  const uint64_t cls = seg_class(m);
  const uint32_t f = sidx[r];
  const uint64_t tile = tileinc[f] - 1;
  const uint64_t low = cls < 3 ? (uint64_t)(kSegMaxM - m) : ((tile << 13) | (uint64_t)(kSegMaxM - m));
  skey[f] = (cls << 60) | ((uint64_t)ty << 58) | low;
  srun[f] = (uint32_t)r;
}

// Class runs of the tiles: rb/re[tile * 6 + class - 3] (zero-initialised;
// f0 = first segment of the lane-group classes).
__global__ void k_tile_runs(const uint64_t* __restrict__ skey, uint64_t f0, uint64_t S, uint32_t* __restrict__ rb,
                            uint32_t* __restrict__ re) {
  const uint64_t f = f0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= S) return;
  const uint64_t key = skey[f];
  const uint64_t run = key >> 13;  // class, type, tile
  const uint64_t slot = ((key >> 13) & ((1ull << 45) - 1)) * 6 + ((key >> 60) - 3);
  if (f == f0 || (skey[f - 1] >> 13) != run) rb[slot] = (uint32_t)f;
  if (f + 1 == S || (skey[f + 1] >> 13) != run) re[slot] = (uint32_t)(f + 1);
}

// Memberships of each tile.
__global__ void k_tile_count(const uint32_t* __restrict__ rb, const uint32_t* __restrict__ re,
                             const SegDesc* __restrict__ desc, uint64_t S, uint64_t n_mem, uint64_t T,
                             uint32_t* __restrict__ cnt) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  uint32_t c = 0;
  for (int k = 0; k < 6; k++) {
    const uint32_t b = rb[t * 6 + k], e = re[t * 6 + k];
    if (b < e) c += (uint32_t)((e < S ? (uint64_t)desc[e].moff : n_mem) - desc[b].moff);
  }
  cnt[t] = c;
}

__global__ void k_tile_check(const uint32_t* __restrict__ cnt, uint64_t T, unsigned int* __restrict__ bad) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < T && cnt[t] >= kTileMax) atomicOr(bad, 1u);
}

// Ranks of the keys i = threadIdx.x + 256 j (j < NK, i < total) among the
// total distinct keys in LDS (padded with 0xFFFFFFFF to a multiple of 4):
// rank[i] and out[rank] = key.
template <int NK>
__device__ __forceinline__ void tile_rank(const uint32_t* key, uint32_t total, uint16_t* rank,
                                          uint32_t* __restrict__ out) {
  uint32_t kk[NK], rr[NK];
#pragma unroll
  for (int j = 0; j < NK; j++) {
    const uint32_t i = threadIdx.x + 256u * j;
    kk[j] = i < total ? key[i] : 0u;
    rr[j] = 0;
  }
  const uint4* k4 = reinterpret_cast<const uint4*>(key);
  const uint32_t n4 = (total + 3) / 4;
#pragma unroll 8
  for (uint32_t q = 0; q < n4; q++) {
    const uint4 v = k4[q];
#pragma unroll
    for (int j = 0; j < NK; j++)
      rr[j] += (v.x < kk[j] ? 1u : 0u) + (v.y < kk[j] ? 1u : 0u) + (v.z < kk[j] ? 1u : 0u) + (v.w < kk[j] ? 1u : 0u);
  }
#pragma unroll
  for (int j = 0; j < NK; j++) {
    const uint32_t i = threadIdx.x + 256u * j;
    if (i < total) {
      rank[i] = (uint16_t)rr[j];
      out[rr[j]] = kk[j];
    }
  }
}

// One workgroup per tile: the distinct dets of its memberships (LDS hash
// table with linear probing; at most kTileMax keys in kTileHash slots), their
// table tab[u0 + i] in ascending order (ranks by counting in LDS) and each
// membership's table index.  A tiles list beta-order positions (the SpMV runs
// them on the beta-order copies of x and y, where a beta group's dets are
// adjacent), the others internal det ids; nu[t] = table size | (A tile << 31).
// Ascending tables make the SpMV's gathers and atomics of a tile touch
// neighbouring addresses.
__global__ __launch_bounds__(256) void k_tile_tables(const uint32_t* __restrict__ rb, const uint32_t* __restrict__ re,
                                                     const uint32_t* __restrict__ u0, const SegDesc* __restrict__ desc,
                                                     uint64_t S, uint64_t n_mem, const uint32_t* __restrict__ mem,
                                                     const uint8_t* __restrict__ type, uint32_t* __restrict__ tab,
                                                     uint32_t* __restrict__ nu, uint16_t* __restrict__ loc) {
  __shared__ uint32_t hk[kTileHash];
  __shared__ uint16_t cid[kTileHash];
  __shared__ uint16_t sl[kTileMax];
  __shared__ __align__(16) uint32_t key[kTileMax];  // read four at a time by tile_rank
  __shared__ uint16_t rank[kTileMax];
  __shared__ uint32_t wtot[4];
  const uint32_t t = blockIdx.x, tid = threadIdx.x;
  uint64_t mb[6], me[6];
  bool isA = false;
  uint32_t cnt = 0;
  for (int k = 0; k < 6; k++) {
    const uint32_t b = rb[(uint64_t)t * 6 + k], e = re[(uint64_t)t * 6 + k];
    mb[k] = me[k] = 0;
    if (b < e) {
      mb[k] = desc[b].moff;
      me[k] = e < S ? (uint64_t)desc[e].moff : n_mem;
      isA = type[b] == kSegA;
      cnt += (uint32_t)(me[k] - mb[k]);
    }
  }
  // hash slots: a power of two >= 2 x memberships, 256 .. kTileHash
  uint32_t hbits = 8;
  while ((1u << hbits) < 2 * cnt && (1u << hbits) < kTileHash) hbits++;
  const uint32_t H = 1u << hbits;
  for (uint32_t h = tid; h < H; h += 256) hk[h] = 0xFFFFFFFFu;
  __syncthreads();
  uint32_t k0 = 0;
  for (int k = 0; k < 6; k++) {
    const uint32_t len = (uint32_t)(me[k] - mb[k]);
    for (uint32_t i = tid; i < len; i += 256) {
      const uint32_t d = mem[mb[k] + i];
      uint32_t h = (d * 2654435761u) >> (32 - hbits);
      while (true) {
        const uint32_t old = atomicCAS(&hk[h], 0xFFFFFFFFu, d);
        if (old == 0xFFFFFFFFu || old == d) break;
        h = (h + 1) & (H - 1);
      }
      sl[k0 + i] = (uint16_t)h;
    }
    k0 += len;
  }
  __syncthreads();
  // occupied slots -> dense index: each thread counts PER consecutive slots,
  // then a block-wide exclusive scan of the counts
  const uint32_t PER = H / 256;
  uint32_t c = 0;
  for (uint32_t q = 0; q < PER; q++) c += hk[tid * PER + q] != 0xFFFFFFFFu ? 1u : 0u;
  uint32_t inc = c;
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t v = (uint32_t)__shfl_up((int)inc, o, kWave);
    if (lane_id() >= (uint32_t)o) inc += v;
  }
  if (lane_id() == kWave - 1) wtot[tid / kWave] = inc;
  __syncthreads();
  uint32_t base = inc - c;
  for (uint32_t wv = 0; wv < tid / kWave; wv++) base += wtot[wv];
  const uint32_t total = wtot[0] + wtot[1] + wtot[2] + wtot[3];
  for (uint32_t q = 0; q < PER; q++) {
    const uint32_t h = tid * PER + q;
    if (hk[h] != 0xFFFFFFFFu) {
      cid[h] = (uint16_t)base;
      key[base] = hk[h];
      base++;
    }
  }
  __syncthreads();
  // rank = number of smaller keys (the keys are distinct): each thread ranks
  // its keys i = tid + 256 j against all keys, read four at a time (the key
  // array padded with 0xFFFFFFFF, larger than any det id); every lane reads
  // the same keys, an LDS broadcast
  for (uint32_t i = total + tid; i < ((total + 3) & ~3u); i += 256) key[i] = 0xFFFFFFFFu;
  __syncthreads();
  const uint32_t nk = (total + 255) / 256;  // keys per thread (<= 8), uniform over the workgroup
  switch (nk) {
    case 0: break;
    case 1: tile_rank<1>(key, total, rank, tab + u0[t]); break;
    case 2: tile_rank<2>(key, total, rank, tab + u0[t]); break;
    case 3: tile_rank<3>(key, total, rank, tab + u0[t]); break;
    case 4: tile_rank<4>(key, total, rank, tab + u0[t]); break;
    default: tile_rank<8>(key, total, rank, tab + u0[t]); break;
  }
  if (tid == 0) nu[t] = total | (isA ? 0x80000000u : 0u);
  __syncthreads();
  k0 = 0;
  for (int k = 0; k < 6; k++) {
    const uint32_t len = (uint32_t)(me[k] - mb[k]);
    for (uint32_t i = tid; i < len; i += 256) loc[mb[k] + i] = rank[cid[sl[k0 + i]]];
    k0 += len;
  }
}

// Per-tile records: what the tile product stages in LDS, contiguous and
// 16-byte aligned so that one wide copy brings it in.
//   header (kTileHead bytes): table size | (A tile << 31), first table entry,
//     segments per class run (u16 x 6), memberships per class run (u16 x 6),
//     first value slot of each class run (u64 x 6);
//   body: per segment (class runs in order) a u32 (slot offset in its class
//     run << 11 | member offset in its class run), then per segment its m
//     (u8), then per membership its index in the tile table (u16).
// A class run has fewer than kTileMax memberships, of at most 64 members per
// segment: slot offsets < kTileMax * 63 / 2 < 2^16, member offsets < 2^11.
constexpr uint32_t kTileHead = 80;
__host__ __device__ __forceinline__ uint32_t tile_body_bytes(uint32_t k, uint32_t nm) {
  return ((((5 * k + 1) & ~1u) + 2 * nm) + 15) & ~15u;
}
constexpr uint32_t kTileBodyMax = ((((5 * (kTileMax / 2) + 1) & ~1u) + 2 * kTileMax) + 15) & ~15u;

// Record size of each tile, in 16-byte units.
__global__ void k_tile_rec_size(const uint32_t* __restrict__ rb, const uint32_t* __restrict__ re,
                                const SegDesc* __restrict__ desc, uint64_t T, uint32_t* __restrict__ units) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  uint32_t k = 0, nm = 0;
  for (int c = 0; c < 6; c++) {
    const uint32_t b = rb[t * 6 + c], e = re[t * 6 + c];
    if (b < e) {
      k += e - b;
      nm += desc[e - 1].moff + desc[e - 1].m - desc[b].moff;
    }
  }
  units[t] = (kTileHead + tile_body_bytes(k, nm)) / 16;
}

__global__ __launch_bounds__(256) void k_tile_records(const uint32_t* __restrict__ rb, const uint32_t* __restrict__ re,
                                                      const uint32_t* __restrict__ u0, const uint32_t* __restrict__ nu,
                                                      const SegDesc* __restrict__ desc,
                                                      const uint16_t* __restrict__ loc,
                                                      const uint32_t* __restrict__ roff, uint4* __restrict__ rec) {
  const uint64_t t = blockIdx.x;
  unsigned char* r = reinterpret_cast<unsigned char*>(rec + roff[t]);
  uint32_t nseg[6], nmem[6], k0[6], m0[6];
  uint64_t voff0[6], moff0[6];
  uint32_t K = 0, NM = 0;
  for (int c = 0; c < 6; c++) {
    const uint32_t b = rb[t * 6 + c], e = re[t * 6 + c];
    nseg[c] = e - b;
    k0[c] = K;
    m0[c] = NM;
    nmem[c] = 0;
    voff0[c] = moff0[c] = 0;
    if (b < e) {
      voff0[c] = desc[b].voff;
      moff0[c] = desc[b].moff;
      nmem[c] = desc[e - 1].moff + desc[e - 1].m - desc[b].moff;
      K += e - b;
      NM += nmem[c];
    }
  }
  if (threadIdx.x == 0) {
    uint32_t* h32 = reinterpret_cast<uint32_t*>(r);
    uint16_t* h16 = reinterpret_cast<uint16_t*>(r + 8);
    uint64_t* h64 = reinterpret_cast<uint64_t*>(r + 32);
    h32[0] = nu[t];
    h32[1] = u0[t];
    for (int c = 0; c < 6; c++) {
      h16[c] = (uint16_t)nseg[c];
      h16[6 + c] = (uint16_t)nmem[c];
      h64[c] = voff0[c];
    }
  }
  unsigned char* body = r + kTileHead;
  uint32_t* sp = reinterpret_cast<uint32_t*>(body);
  uint8_t* sm = body + 4 * K;
  uint16_t* sl = reinterpret_cast<uint16_t*>(body + ((5 * K + 1) & ~1u));
  for (int c = 0; c < 6; c++) {
    const uint32_t b = rb[t * 6 + c];
    for (uint32_t k = threadIdx.x; k < nseg[c]; k += 256) {
      const SegDesc D = desc[b + k];
      sp[k0[c] + k] = ((uint32_t)(D.voff - voff0[c]) << 11) | (uint32_t)(D.moff - moff0[c]);
      sm[k0[c] + k] = (uint8_t)D.m;
    }
    for (uint32_t i = threadIdx.x; i < nmem[c]; i += 256) sl[m0[c] + i] = loc[moff0[c] + i];
  }
}

// Ordered segment f: sizes, slot counts, run -> segment map.
__global__ void k_seg_sizes(const uint32_t* __restrict__ srun, const uint32_t* __restrict__ rstart,
                            const uint32_t* __restrict__ useinc, uint64_t S, uint64_t* __restrict__ msz,
                            uint64_t* __restrict__ nslot, uint32_t* __restrict__ finv) {
  const uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= S) return;
  const uint32_t r = srun[f];
  uint32_t before;
  const uint64_t m = run_kept(useinc, rstart, r, before);
  msz[f] = m;
  nslot[f] = m * (m - 1) / 2;
  finv[r] = (uint32_t)f;
}

// First segment of each (class, type) range in the sorted order: bnd[r] for
// r = 0..kSegRanges (bnd[kSegRanges] = S).  Each entry is written once.
__global__ void k_seg_bounds(const uint64_t* __restrict__ skey, uint64_t S, uint64_t* __restrict__ bnd) {
  const uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= S) return;
  auto range = [](uint64_t key) { return (int)(key >> 60) * 3 + (int)((key >> 58) & 3); };
  const int rf = range(skey[f]);
  const int rp = f == 0 ? -1 : range(skey[f - 1]);
  for (int r = rp + 1; r <= rf; r++) bnd[r] = f;
  if (f + 1 == S)
    for (int r = rf + 1; r <= kSegRanges; r++) bnd[r] = S;
}

__global__ void k_seg_desc(SegKeyParams K, const uint32_t* __restrict__ srun, const uint32_t* __restrict__ rstart,
                           const uint32_t* __restrict__ useinc, const uint32_t* __restrict__ vals_s,
                           const uint64_t* __restrict__ moff,
                           const uint64_t* __restrict__ voff, uint64_t S, SegDesc* __restrict__ desc,
                           uint8_t* __restrict__ type, uint64_t* __restrict__ core) {
  const uint64_t f = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= S) return;
  const uint32_t r = srun[f];
  SegDesc D;
  uint32_t before;
  D.moff = (uint32_t)moff[f];
  D.m = run_kept(useinc, rstart, r, before);
  D.voff = voff[f];
  desc[f] = D;
  uint32_t ty, k;
  uint64_t i;
  seg_decode(K, vals_s[rstart[r]], ty, i, k);
  type[f] = (uint8_t)ty;
  core[f] = ty == kSegC ? 0ull : ty == kSegA ? drop_electrons(K.up[i], k, K.ra) : drop_electrons(K.dn[i], k, K.rb);
}

__global__ void k_seg_members(SegKeyParams K, const uint32_t* __restrict__ vals_s, const uint32_t* __restrict__ incl,
                              const uint32_t* __restrict__ rstart, const uint32_t* __restrict__ use,
                              const uint32_t* __restrict__ useinc, const uint32_t* __restrict__ finv,
                              const SegDesc* __restrict__ desc, uint64_t M, uint32_t* __restrict__ mem) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= M || !use[p]) return;
  const uint32_t r = incl[p] - 1;
  const uint32_t f = finv[r];
  if (f == 0xFFFFFFFFu) return;
  uint32_t ty, k;
  uint64_t i;
  seg_decode(K, vals_s[p], ty, i, k);
  const uint32_t a = rstart[r];
  const uint32_t before = a ? useinc[a - 1] : 0u;
  mem[(uint64_t)desc[f].moff + (useinc[p] - 1 - before)] = ty == kSegA ? K.dnpos[i] : (uint32_t)i;
}

// ---- value fill / update ----
struct SegFillParams {
  const SegDesc* __restrict__ desc;
  const uint32_t* __restrict__ mem;
  const uint8_t* __restrict__ type;
  const uint64_t* __restrict__ core;
  const uint64_t* __restrict__ up;
  const uint64_t* __restrict__ dn;
  const uint32_t* __restrict__ dnord;  // beta-order position -> internal det (A members)
  Ints I;
  double* __restrict__ v64;
  float* __restrict__ v32;
  unsigned long long* __restrict__ kept;
  uint32_t ra, rb;
  uint32_t s0, s1;
  int keep_zeros;
  int update;  // recompute the values of the stored pairs only (update_existing_elems)
};

// One slot of the pair (j, l): ownership, element, cut.  Returns 1 if kept.
__device__ __forceinline__ uint32_t seg_slot(const SegFillParams& F, uint32_t ty, uint64_t core, uint64_t slot,
                                             uint64_t uj, uint64_t dj, uint64_t ul, uint64_t dl) {
  double v = 0.0;
  bool keep = false;
  if (seg_owns(ty, core, ty == kSegA ? F.ra : F.rb, uj, dj, ul, dl)) {
    if (F.update) {
      keep = F.keep_zeros || (F.v64 ? F.v64[slot] != 0.0 : F.v32[slot] != 0.0f);
      if (keep) v = elem_any(F.I, uj, dj, ul, dl);
    } else {
      v = elem_any(F.I, uj, dj, ul, dl);
      keep = F.keep_zeros || fabs(v) >= kEps;
      if (!keep) v = 0.0;
    }
  }
  // streaming stores: the slots are not read again by the fill, and keeping
  // them out of L2 leaves it to the integral tables
  if (F.v64) __builtin_nontemporal_store(v, F.v64 + slot);
  if (F.v32) __builtin_nontemporal_store((float)v, F.v32 + slot);
  return keep ? 1u : 0u;
}

// Kept-pair counters: one per 128-byte line, a wavefront adds to the line of
// its workgroup (modulo kKeptLines), so that no single address takes an atomic
// from every wavefront of the fill.
constexpr uint32_t kKeptLines = 1024;

__device__ __forceinline__ void seg_count(const SegFillParams& F, uint32_t kept) {
  for (int o = 32; o > 0; o >>= 1) kept += __shfl_xor(kept, o, kWave);
  if (lane_id() == 0 && kept) atomicAdd(F.kept + (blockIdx.x % kKeptLines) * 16, (unsigned long long)kept);
}

// Lane groups of S lanes, one segment (m <= S) per group.
template <int S>
__global__ __launch_bounds__(256) void k_seg_fill_small(SegFillParams F) {
  const uint64_t s = (uint64_t)F.s0 + (uint64_t)blockIdx.x * (256 / S) + threadIdx.x / S;
  const uint32_t j = threadIdx.x & (S - 1);
  const int gbase = (int)(lane_id() & ~(uint32_t)(S - 1));
  SegDesc D{0u, 0u, 0ull};
  uint32_t ty = 0;
  uint64_t core = 0;
  if (s < F.s1) {
    D = F.desc[s];
    ty = F.type[s];
    core = F.core[s];
  }
  const uint32_t m = D.m, mh = m >> 1;
  const bool valid = j < m;
  uint64_t uj = 0, dj = 0;
  if (valid) {
    const uint32_t mj = F.mem[D.moff + j];
    const uint32_t det = ty == kSegA ? F.dnord[mj] : mj;
    uj = F.up[det];
    dj = F.dn[det];
  }
  uint32_t kept = 0, p = j;
  for (uint32_t d = 1; d <= mh; d++) {
    if (valid) p = p + 1 == m ? 0 : p + 1;
    const uint64_t ul = shfl64(uj, gbase + (int)p), dl = shfl64(dj, gbase + (int)p);
    if (valid && (2 * d != m || j < mh))
      kept += seg_slot(F, ty, core, D.voff + (uint64_t)(d - 1) * m + j, uj, dj, ul, dl);
  }
  seg_count(F, kept);
}

// c2 of orbital indices and of their pair indices (below 2^12): 32-bit form.
__device__ __forceinline__ uint32_t c2s(uint32_t a, uint32_t b) {
  const uint32_t hi = max(a, b), lo = min(a, b);
  return hi * (hi + 1u) / 2u + lo;
}

// A member of a mixed-core (C) segment over the segment's core strings cu, cd
// (alpha and beta, one electron fewer each): the alpha orbital a and beta
// orbital b it adds, and the parity r of the core orbitals below a plus those
// below b, packed as a | b << 7 | r << 14.  For two members j, l with a_j !=
// a_l and b_j != b_l, double_mixed's excitation is a_j -> a_l, b_j -> b_l;
// its below() counts are the core counts below those orbitals, so its sign is
// (-1)^(r_j + r_l) and its integral (a_j a_l | b_j b_l).  When every member
// adds the same alpha (or beta) orbital, the AND of the members' strings
// holds it: a (or b) is then 64 for all members and no pair is owned.
__device__ __forceinline__ uint32_t c_member_word(uint64_t u, uint64_t d, uint64_t cu, uint64_t cd) {
  const uint64_t xu = u & ~cu, xd = d & ~cd;
  const uint32_t a = xu ? (uint32_t)ctz64(xu) : 64u, b = xd ? (uint32_t)ctz64(xd) : 64u;
  const uint32_t r = (uint32_t)(__popcll(cu & ((1ull << (a & 63u)) - 1ull)) +
                                __popcll(cd & ((1ull << (b & 63u)) - 1ull))) & 1u;
  return a | b << 7 | r << 14;
}

// The value index and sign of the C pair (j, l) from their member words, and
// whether the segment owns it (both spins change).
__device__ __forceinline__ bool c_pair(uint32_t wj, uint32_t wl, uint32_t& idx, double& sgn) {
  const uint32_t x = wj ^ wl;
  idx = c2s(c2s(wj & 0x7Fu, wl & 0x7Fu), c2s((wj >> 7) & 0x7Fu, (wl >> 7) & 0x7Fu));
  sgn = (x >> 14) & 1u ? -1.0 : 1.0;
  return (x & 0x7Fu) != 0u && (x & (0x7Fu << 7)) != 0u;
}

// Mixed-core (C) segments at build time: every owned pair is an opposite-spin
// double whose value is sign * (pu qu | pd qd) (chem_system.cc:599-607, as in
// double_mixed), computed from the two members' words (c_member_word; one
// 32-bit lane exchange per slot).  Four steps per iteration: the partner
// words, ownership tests, signs and integral indices first, then four
// unconditional integral loads.
template <int S>
__global__ __launch_bounds__(256) void k_seg_fill_c(SegFillParams F) {
  constexpr int U = 4;
  const uint64_t s = (uint64_t)F.s0 + (uint64_t)blockIdx.x * (256 / S) + threadIdx.x / S;
  const uint32_t j = threadIdx.x & (S - 1);
  const int gbase = (int)(lane_id() & ~(uint32_t)(S - 1));
  SegDesc D{0u, 0u, 0ull};
  if (s < F.s1) D = F.desc[s];
  const uint32_t m = D.m, mh = m >> 1;
  const bool valid = j < m;
  // the members' strings; the AND over a segment's lanes is its core
  uint64_t uj = ~0ull, dj = ~0ull;
  if (valid) {
    const uint32_t det = F.mem[D.moff + j];
    uj = F.up[det];
    dj = F.dn[det];
  }
  uint64_t cu = uj, cd = dj;
#pragma unroll
  for (int o = 1; o < S; o <<= 1) {
    cu &= shfl64(cu, (int)(lane_id() ^ (uint32_t)o));
    cd &= shfl64(cd, (int)(lane_id() ^ (uint32_t)o));
  }
  const uint32_t wj = c_member_word(uj, dj, cu, cd);
  // Steps of the wave's largest segment (segments of a range are in group order).
  uint32_t steps = mh;
  for (int o = 32; o > 0; o >>= 1) steps = max(steps, (uint32_t)__shfl_xor((int)steps, o, kWave));
  uint32_t kept = 0, p = j;
  for (uint32_t d0 = 1; d0 <= steps; d0 += U) {
    uint32_t idx[U];
    double sgn[U];
    bool live[U], own[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t d = d0 + u;
      if (valid) p = p + 1 == m ? 0 : p + 1;
      const uint32_t wl = (uint32_t)__shfl((int)wj, gbase + (int)p, kWave);
      live[u] = valid && d <= mh && (2 * d != m || j < mh);
      own[u] = c_pair(wj, wl, idx[u], sgn[u]) && live[u];
      if (!own[u]) idx[u] = 0;
    }
    double e[U];
#pragma unroll
    for (int u = 0; u < U; u++) e[u] = F.I.eri[idx[u]];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (!live[u]) continue;
      double v = 0.0;
      bool keep = false;
      if (own[u]) {
        v = e[u];
        v *= sgn[u];
        keep = F.keep_zeros || fabs(v) >= kEps;
        if (!keep) v = 0.0;
      }
      const uint64_t slot = D.voff + (uint64_t)(d0 + u - 1) * m + j;
      if (F.v64) __builtin_nontemporal_store(v, F.v64 + slot);
      if (F.v32) __builtin_nontemporal_store((float)v, F.v32 + slot);
      kept += keep ? 1u : 0u;
    }
  }
  seg_count(F, kept);
}

// One wavefront per segment (m > 64), member dets staged in LDS.
__global__ __launch_bounds__(64) void k_seg_fill_large(SegFillParams F) {
  extern __shared__ __align__(16) unsigned char seg_lds[];
  const uint32_t s = F.s0 + blockIdx.x;
  if (s >= F.s1) return;
  const SegDesc D = F.desc[s];
  const uint32_t ty = F.type[s];
  const uint64_t core = F.core[s];
  const uint32_t m = D.m, mh = m >> 1, t = threadIdx.x;
  uint64_t* su = reinterpret_cast<uint64_t*>(seg_lds);
  uint64_t* sd = su + m;
  for (uint32_t j = t; j < m; j += kWave) {
    const uint32_t mj = F.mem[D.moff + j];
    const uint32_t det = ty == kSegA ? F.dnord[mj] : mj;
    su[j] = F.up[det];
    sd[j] = F.dn[det];
  }
  __syncthreads();
  uint32_t kept = 0;
  for (uint32_t d = 1; d <= mh; d++) {
    const uint32_t lim = 2 * d == m ? mh : m;
    for (uint32_t j = t; j < lim; j += kWave) {
      uint32_t p = j + d;
      if (p >= m) p -= m;
      kept += seg_slot(F, ty, core, D.voff + (uint64_t)(d - 1) * m + j, su[j], sd[j], su[p], sd[p]);
    }
  }
  seg_count(F, kept);
}

// Mixed-core (C) segments of m > 64 at build time: one wavefront per segment,
// member words (c_member_word) in LDS, each lane walking its slots (d, j),
// (d, j + 64), ... four at a time so that four integral loads are in flight.
// Values as in k_seg_fill_c.  Dynamic LDS: 4 bytes per member.
__global__ __launch_bounds__(64) void k_seg_fill_large_c(SegFillParams F) {
  extern __shared__ __align__(16) unsigned char seg_lds[];
  constexpr int U = 4;
  const uint32_t s = F.s0 + blockIdx.x;
  if (s >= F.s1) return;
  const SegDesc D = F.desc[s];
  const uint32_t m = D.m, mh = m >> 1, t = threadIdx.x;
  uint32_t* sw = reinterpret_cast<uint32_t*>(seg_lds);
  // the core: the AND of the members' strings
  uint64_t cu = ~0ull, cd = ~0ull;
  for (uint32_t j = t; j < m; j += kWave) {
    const uint32_t det = F.mem[D.moff + j];
    cu &= F.up[det];
    cd &= F.dn[det];
  }
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    cu &= shfl64(cu, (int)(lane_id() ^ (uint32_t)o));
    cd &= shfl64(cd, (int)(lane_id() ^ (uint32_t)o));
  }
  for (uint32_t j = t; j < m; j += kWave) {
    const uint32_t det = F.mem[D.moff + j];
    sw[j] = c_member_word(F.up[det], F.dn[det], cu, cd);
  }
  __syncthreads();
  uint32_t kept = 0;
  uint32_t d = 1, j = t, lim = m;  // slots of step d (m > 64, so step 1 has m); j < 64 < lim
  while (d <= mh) {
    uint32_t idx[U];
    uint64_t slot[U];
    double sgn[U];
    bool live[U], own[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      live[u] = d <= mh;
      own[u] = false;
      idx[u] = 0;
      sgn[u] = 1.0;
      slot[u] = 0;
      if (live[u]) {
        uint32_t p = j + d;
        if (p >= m) p -= m;
        slot[u] = D.voff + (uint64_t)(d - 1) * m + j;
        own[u] = c_pair(sw[j], sw[p], idx[u], sgn[u]);
        if (!own[u]) idx[u] = 0;
        // next slot of this lane: 64 further in the rotation order
        j += kWave;
        while (d <= mh && j >= lim) {
          j -= lim;
          d++;
          lim = 2 * d == m ? mh : m;
        }
      }
    }
    double e[U];
#pragma unroll
    for (int u = 0; u < U; u++) e[u] = F.I.eri[idx[u]];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (!live[u]) continue;
      double v = 0.0;
      bool keep = false;
      if (own[u]) {
        v = e[u];
        v *= sgn[u];
        keep = F.keep_zeros || fabs(v) >= kEps;
        if (!keep) v = 0.0;
      }
      if (F.v64) __builtin_nontemporal_store(v, F.v64 + slot[u]);
      if (F.v32) __builtin_nontemporal_store((float)v, F.v32 + slot[u]);
      kept += keep ? 1u : 0u;
    }
  }
  seg_count(F, kept);
}

// ---- SpMV ----
// y = (d - shift) x for NV interleaved vectors (the diagonal part).
template <int NV, typename XT>
__global__ void k_diag_apply(const double* __restrict__ diag, double shift, const XT* __restrict__ x,
                             XT* __restrict__ y, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dd = diag[i] - shift;
  for (int k = 0; k < NV; k++) y[i * NV + k] = (XT)(dd * (double)x[i * NV + k]);
}

template <int NV, typename XT>
__device__ __forceinline__ void load_x(const XT* __restrict__ x, uint32_t det, XT (&v)[NV]) {
  if constexpr (NV == 4 && sizeof(XT) == 4) {
    const float4 q = *reinterpret_cast<const float4*>(x + (uint64_t)det * 4);
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
  } else {
    for (int k = 0; k < NV; k++) v[k] = x[(uint64_t)det * NV + k];
  }
}

// w if k, else 0, as a product: the compiler keeps the load of w unconditional
// (a select on k feeding a widening conversion is turned into a branch around
// the load that waits for it).  Stored values are finite.
template <typename XT, typename VT>
__device__ __forceinline__ XT keep_if(VT w, bool k) {
  return (XT)w * (k ? (XT)1 : (XT)0);
}

// Value of the lane whose byte address (4 * lane) is a (ds_bpermute).  All
// lanes of the wave must be active.
__device__ __forceinline__ float lane_read(float v, int a) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(a, __float_as_int(v)));
}
__device__ __forceinline__ double lane_read(double v, int a) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)(uint32_t)u);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)(uint32_t)(u >> 32));
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Member sums of NB segments of m <= S members, one lane group of S lanes per
// segment and lane j holding member j (x_j given; zero past m).  At step d lane
// j loads the value v of the pair (j, j + d mod m), adds v * x_{j+d} to its own
// sum and v * x_j to a travelling sum R.  x_{j+d} and R both move one lane down
// the segment per step (lane j reads lane j + 1 mod m, a fixed address), so
// after T steps lane j's R holds what it owes member j + T mod m; one last
// cross-lane read delivers it.  The last step of an even m (m/2 slots, the pair
// (j, j + m/2) read by both of its lanes) adds the forward term only.
//
// Segment b has mm[b] members and its slots start at base + soff[b] (base
// wave-uniform, offsets 32-bit, so each load is a scalar base plus one vector
// offset that one addition advances per step).  The step loop runs for
// `steps` steps, the most of any segment of the wave (wave uniform); the first
// `steps_all` steps belong to every segment of the wave and run without
// masks; past them a lane whose segment has no such step adds zero.  Loads
// past a segment's last step stay inside the store, which ends with
// kSlotPad zero slots; lanes past m hold x = 0 and add nothing.
// All lanes of the wave must call this.
template <int S, int NB, int CH, typename VT, typename XT>
__device__ __forceinline__ void seg_group_sums(const uint32_t (&mm)[NB], const uint32_t (&soff)[NB], const XT (&xj)[NB],
                                               const VT* __restrict__ base, uint32_t steps, uint32_t steps_all,
                                               XT (&out)[NB]) {
  // CH: steps per batch of value loads
  using gptr = const __attribute__((address_space(1))) VT*;  // global loads, not flat
  using gbyte = const __attribute__((address_space(1))) char*;
  const gptr vb = (gptr)base;
  const gbyte bb = (gbyte)base;
  // a slot at byte offset o (32-bit): one scalar base plus one vector offset
  // ordinary (cached) loads: rows of a few tens of bytes share cache lines
  // with the next step's and the next segment's, which a streaming load would
  // let go before they are read again
  auto at = [&](uint32_t o) { return *((gptr)(bb + o)); };
  const uint32_t j = threadIdx.x & (S - 1);
  const uint32_t gbase = lane_id() & ~(uint32_t)(S - 1);
  uint32_t full[NB], off[NB], stride[NB];
  int nxt[NB];
  XT xp[NB], acc[NB], R[NB];
  VT wh[NB];
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const uint32_t m = mm[b], mh = m >> 1;
    const bool valid = j < m;
    full[b] = valid ? (m - 1) / 2 : 0;  // steps with m slots
    const uint32_t jn = j + 1 == m ? 0 : j + 1;
    nxt[b] = (int)((gbase + (valid ? jn : j)) << 2);
    off[b] = (soff[b] + j) * (uint32_t)sizeof(VT);
    stride[b] = m * (uint32_t)sizeof(VT);
    xp[b] = xj[b];
    acc[b] = (XT)0;
    R[b] = (XT)0;
    // the half-step value of an even m, loaded ahead of the steps
    const bool half = valid && (m & 1u) == 0;
    const uint32_t hj = j < mh ? j : j - mh;
    wh[b] = vb[soff[b] + (half ? (mh - 1) * m + hj : 0u)];
  }
  auto apply = [&](const XT (&v)[NB][CH]) {
#pragma unroll
    for (int i = 0; i < CH; i++) {
#pragma unroll
      for (int b = 0; b < NB; b++) {
        xp[b] = lane_read(xp[b], nxt[b]);
        R[b] = lane_read(R[b], nxt[b]);
        acc[b] = fma(v[b][i], xp[b], acc[b]);
        R[b] = fma(v[b][i], xj[b], R[b]);
      }
    }
  };
  uint32_t T = 0;  // rotations done
  uint32_t c0 = 1;
#pragma unroll 1
  for (; c0 + CH - 1 <= steps_all; c0 += CH, T += CH) {
    XT v[NB][CH];
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int i = 0; i < CH; i++) {
        v[b][i] = (XT)at(off[b]);
        off[b] += stride[b];
      }
    apply(v);
  }
#pragma unroll 1
  for (; c0 <= steps; c0 += CH, T += CH) {
    XT v[NB][CH];
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int i = 0; i < CH; i++) {
        const VT w = at(off[b]);
        off[b] += stride[b];
        if constexpr (sizeof(VT) == sizeof(XT)) {
          v[b][i] = c0 + i <= full[b] ? w : (VT)0;
        } else {
          v[b][i] = keep_if<XT>(w, c0 + i <= full[b]);
        }
      }
    apply(v);
  }
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const uint32_t mb = mm[b], mh = mb >> 1;
    const bool valid = j < mb;
    const bool half = valid && (mb & 1u) == 0;
    const uint32_t pj = j + mh >= mb ? j + mh - mb : j + mh;
    const XT xh = lane_read(xj[b], (int)((gbase + (valid ? pj : j)) << 2));
    acc[b] = fma(keep_if<XT>(wh[b], half), xh, acc[b]);
    // (j - T) mod m; T < 2m for the segments of a lane-group class: T is
    // (S/2 - 1) rounded up to a multiple of CH <= S at most, and m > S/2
    // (or m = S = 2, T = 0)
    uint32_t src = j + 2 * mb - T;
    src = src >= mb ? src - mb : src;
    src = src >= mb ? src - mb : src;
    out[b] = acc[b] + lane_read(R[b], (int)((gbase + (valid ? src : j)) << 2));
  }
}

// A tile in LDS: its x and y, and the body of its record (per segment the
// packed slot and member offsets and m, per membership the index into x and
// y; segments numbered class run by class run).
template <typename XT>
struct TileLds {
  XT xs[kTileMax];
  XT ys[kTileMax];
  uint4 body[kTileBodyMax / 16];
};

// Views of a staged record body with K segments.
struct TileBody {
  const uint32_t* sp;  // (slot offset in the class run << 11) | member offset in the class run
  const uint8_t* sm;   // m
  const uint16_t* sl;  // membership -> index into xs / ys
};

// One class run of a tile: its n segments (LDS entries k0 ..) with m <= S,
// lane groups of S lanes with NB segments each per round; the run's first
// slot voff0 and first membership m0 (LDS).
template <int S, int NB, int CH, typename VT, typename XT>
__device__ __forceinline__ void tile_class_run(uint32_t n, uint32_t k0, uint32_t m0, uint64_t voff0,
                                               const VT* __restrict__ val, TileLds<XT>& L, const TileBody& B) {
  static_assert(CH <= S, "the delivery of the travelling sums assumes fewer than 2m rotations");
  constexpr uint32_t GW = kWave / S;  // lane groups per wave
  constexpr uint32_t PW = GW * NB;    // segments per wave and round
  const uint32_t j = threadIdx.x & (S - 1), gl = (threadIdx.x & (kWave - 1)) / S, wv = threadIdx.x / kWave;
  for (uint32_t r0 = 0; r0 < n; r0 += 4 * PW) {
    // the wave's segments are consecutive, the first the largest (a class
    // run lists its segments largest first); waves past the run idle
    const uint32_t sw = r0 + wv * PW;
    if (sw >= n) break;
    const uint32_t steps = (B.sm[k0 + sw] - 1u) / 2;
    // the wave's smallest segment (its last) bounds the steps all of its segments have
    const uint32_t slast = sw + PW - 1;
    const uint32_t steps_all = slast < n ? (B.sm[k0 + slast] - 1u) / 2 : 0u;
    uint32_t mm[NB], so[NB];
    XT xj[NB], out[NB];
    uint32_t l[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) {
      const uint32_t sk = sw + k * GW + gl;
      mm[k] = so[k] = 0;
      uint32_t moff = 0;
      if (sk < n) {
        const uint32_t pk = B.sp[k0 + sk];
        mm[k] = B.sm[k0 + sk];
        moff = m0 + (pk & 0x7FFu);
        so[k] = pk >> 11;
      }
      const bool valid = j < mm[k];
      l[k] = valid ? B.sl[moff + j] : 0u;
      xj[k] = valid ? L.xs[l[k]] : (XT)0;
    }
    seg_group_sums<S, NB, CH, VT, XT>(mm, so, xj, val + voff0, steps, steps_all, out);
#pragma unroll
    for (int k = 0; k < NB; k++)
      if (j < mm[k]) atomicAdd(&L.ys[l[k]], out[k]);
  }
}

// Pair sums of one segment of M members (M <= 8) held by one lane: slot
// (d - 1) * M + j of the segment's rotation rows is the pair (j, j + d mod M)
// (d <= (M - 1) / 2), and for an even M the half row's slot
// ((M - 1) / 2) * M + j the pair (j, j + M / 2), j < M / 2.
template <int M, typename XT>
__device__ __forceinline__ void small_pairs(const XT (&v)[28], const XT (&x)[8], XT (&y)[8]) {
#pragma unroll
  for (int d = 1; d <= (M - 1) / 2; d++)
#pragma unroll
    for (int j = 0; j < M; j++) {
      const int k = (j + d) % M;
      const XT w = v[(d - 1) * M + j];
      y[j] = fma(w, x[k], y[j]);
      y[k] = fma(w, x[j], y[k]);
    }
  if constexpr (M % 2 == 0) {
#pragma unroll
    for (int j = 0; j < M / 2; j++) {
      const XT w = v[((M - 1) / 2) * M + j];
      y[j] = fma(w, x[j + M / 2], y[j]);
      y[j + M / 2] = fma(w, x[j], y[j + M / 2]);
    }
  }
}

// The class runs of 5-8, 3-4 and 2 members of a tile (LDS entries k3 ..
// k3 + n, three runs in order, segments largest first within a run): one
// lane per segment, its slots loaded at once, its members' x from the
// tile's LDS x, its sums added to the LDS y.
template <typename VT, typename XT>
__device__ __forceinline__ void tile_small_runs(const uint32_t (&nseg)[6], const uint32_t (&k0)[6],
                                                const uint32_t (&m0)[6], const uint64_t (&voff0)[6],
                                                const VT* __restrict__ val, TileLds<XT>& L, const TileBody& B) {
  const uint32_t kb = k0[3], ke = k0[5] + nseg[5];
  for (uint32_t e = kb + threadIdx.x; e < ke; e += 256) {
    const int c = e >= k0[5] ? 5 : e >= k0[4] ? 4 : 3;
    const uint32_t pk = B.sp[e], m = B.sm[e];
    const uint32_t moff = m0[c] + (pk & 0x7FFu);
    const VT* __restrict__ base = val + voff0[c] + (pk >> 11);
    const uint32_t ns = m * (m - 1) / 2;
    XT v[28];
#pragma unroll
    for (int t = 0; t < 28; t++) v[t] = (uint32_t)t < ns ? (XT)base[t] : (XT)0;
    uint32_t l[8];
    XT x[8], y[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      l[j] = (uint32_t)j < m ? B.sl[moff + j] : 0u;
      x[j] = (uint32_t)j < m ? L.xs[l[j]] : (XT)0;
      y[j] = (XT)0;
    }
    switch (m) {
      case 8: small_pairs<8, XT>(v, x, y); break;
      case 7: small_pairs<7, XT>(v, x, y); break;
      case 6: small_pairs<6, XT>(v, x, y); break;
      case 5: small_pairs<5, XT>(v, x, y); break;
      case 4: small_pairs<4, XT>(v, x, y); break;
      case 3: small_pairs<3, XT>(v, x, y); break;
      default: small_pairs<2, XT>(v, x, y); break;
    }
#pragma unroll
    for (int j = 0; j < 8; j++)
      if ((uint32_t)j < m) atomicAdd(&L.ys[l[j]], y[j]);
  }
}

// Off-diagonal part of the segments with m <= 64, one workgroup per tile: the
// tile's record body copied into LDS with 16-byte loads (its header read
// through scalar loads), its x gathered (one load per distinct det), its class
// runs of 33-64, 17-32 and 9-16 members with lane groups of 64, 32 and 16
// lanes and those of at most 8 members with one lane per segment, member sums
// into an LDS y, then one global atomic per distinct det.  VT stored value type, XT
// vector type.
template <typename VT, typename XT>
__device__ __forceinline__ void seg_spmv_tile(const uint32_t* __restrict__ roff, const uint4* __restrict__ rec,
                                              const uint32_t* __restrict__ tab, const VT* __restrict__ val,
                                              const XT* __restrict__ xi, XT* __restrict__ yi,
                                              const XT* __restrict__ xb, XT* __restrict__ yb) {
  __shared__ TileLds<XT> L;
  const uint4* r = rec + roff[blockIdx.x];
  const uint32_t* h32 = reinterpret_cast<const uint32_t*>(r);
  const uint16_t* h16 = reinterpret_cast<const uint16_t*>(h32 + 2);
  const uint64_t* h64 = reinterpret_cast<const uint64_t*>(h32 + 8);
  const uint32_t nut = h32[0], U = nut & 0x7FFFFFFFu;
  const XT* __restrict__ x = (nut >> 31) ? xb : xi;  // A tiles: beta-order copies
  XT* __restrict__ y = (nut >> 31) ? yb : yi;
  const uint32_t* __restrict__ tb = tab + h32[1];
  uint32_t nseg[6], k0[6], m0[6];
  uint64_t voff0[6];
  uint32_t K = 0, NM = 0;
#pragma unroll
  for (int c = 0; c < 6; c++) {
    nseg[c] = h16[c];
    k0[c] = K;
    m0[c] = NM;
    voff0[c] = h64[c];
    K += nseg[c];
    NM += h16[6 + c];
  }
  const uint4* body = r + kTileHead / 16;
  const uint32_t nv = tile_body_bytes(K, NM) / 16;
  for (uint32_t v = threadIdx.x; v < nv; v += 256) L.body[v] = body[v];
  for (uint32_t u = threadIdx.x; u < U; u += 256) {
    L.xs[u] = x[tb[u]];
    L.ys[u] = (XT)0;
  }
  const unsigned char* lb = reinterpret_cast<const unsigned char*>(L.body);
  const TileBody B{reinterpret_cast<const uint32_t*>(lb), lb + 4 * K,
                   reinterpret_cast<const uint16_t*>(lb + ((5 * K + 1) & ~1u))};
  __syncthreads();
  // (lanes per segment, segments per lane group and round, steps per load
  // batch); the widening product (FP32 values, FP64 vectors) runs the wide
  // classes one segment per lane group with longer batches
  constexpr bool MX = sizeof(VT) != sizeof(XT);
  tile_class_run<64, MX ? 1 : 2, MX ? 16 : 8, VT, XT>(nseg[0], k0[0], m0[0], voff0[0], val, L, B);
  tile_class_run<32, MX ? 1 : 2, MX ? 16 : 8, VT, XT>(nseg[1], k0[1], m0[1], voff0[1], val, L, B);
  tile_class_run<16, MX ? 1 : 2, 8, VT, XT>(nseg[2], k0[2], m0[2], voff0[2], val, L, B);
  tile_small_runs<VT, XT>(nseg, k0, m0, voff0, val, L, B);
  __syncthreads();
  for (uint32_t u = threadIdx.x; u < U; u += 256) atomicAdd(&y[tb[u]], L.ys[u]);
}

template <typename VT, typename XT>
__global__ __launch_bounds__(256) void k_seg_spmv_tiles(const uint32_t* __restrict__ roff, const uint4* __restrict__ rec,
                                                        const uint32_t* __restrict__ tab, const VT* __restrict__ val,
                                                        const XT* __restrict__ xi, XT* __restrict__ yi,
                                                        const XT* __restrict__ xb, XT* __restrict__ yb) {
  seg_spmv_tile<VT, XT>(roff, rec, tab, val, xi, yi, xb, yb);
}

// The FP32 product's tile kernel: its LDS leaves room for six workgroups per
// CU, so it is compiled for six waves per SIMD (at most 80 VGPRs).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6)))
void k_seg_spmv_tiles_f32(const uint32_t* __restrict__ roff, const uint4* __restrict__ rec,
                          const uint32_t* __restrict__ tab, const float* __restrict__ val,
                          const float* __restrict__ xi, float* __restrict__ yi, const float* __restrict__ xb,
                          float* __restrict__ yb) {
  seg_spmv_tile<float, float>(roff, rec, tab, val, xi, yi, xb, yb);
}


// Off-diagonal part, one wavefront per (segment, group of 64 T members) for
// m > 64 (one vector).  The member x sit twice in LDS (sx[k] = sx[k + m] =
// x_k), so the partners j + d and j - d mod m are read without wrapping; each
// lane owns the members j0 + lane + 64 c (c < T, and only the chunks that hold
// members) and keeps their sums in registers while the steps d run outermost,
// so the forward slot (j, j + d) and the backward slot (j - d, j) of a step
// are read from the same row of the segment within a short window (cache hits
// for the second read).  Value loads are unconditional at in-bounds (clamped)
// slots and the unused half-step term is masked.
constexpr int kSegChunks = 4;
template <int NC, typename VT, typename XT>
__device__ __forceinline__ void seg_large_rows(const SegDesc& D, uint32_t j0, const VT* __restrict__ val,
                                               const XT* sx, const uint32_t* __restrict__ mem, XT* __restrict__ y) {
  constexpr uint32_t U = sizeof(XT) == 8 ? 1 : 4;  // steps per batch of loads (FP64: register budget)
  const uint32_t m = D.m, mh = m >> 1, t = threadIdx.x;
  uint32_t d_hi = (m & 1) ? mh : mh - 1;  // steps with m slots
  uint32_t d_lo = 1;
  bool half_step = (m & 1) == 0;
  if constexpr (sizeof(XT) == 8) {
    // FP64: blockIdx.z's share of the steps (the last share also the half
    // step of an even m); the shares' sums meet in the atomics on y
    const uint32_t per = (d_hi + gridDim.z - 1) / gridDim.z;
    d_lo = 1 + blockIdx.z * per;
    d_hi = min(d_hi, (blockIdx.z + 1) * per);
    half_step = half_step && blockIdx.z + 1 == gridDim.z;
  }
  XT acc[NC];
  uint32_t jj[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const uint32_t j = j0 + kWave * c + t;
    jj[c] = j < m ? j : m - 1;
    acc[c] = (XT)0;
  }
  const VT* __restrict__ vs = val + D.voff;
  uint32_t d = d_lo;
  for (; d + U - 1 <= d_hi; d += U) {
    XT vf[U][NC], vb[U][NC];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const VT* __restrict__ vd = vs + (uint64_t)(d + u - 1) * m;
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const uint32_t j = jj[c], du = d + u;
        vf[u][c] = (XT)vd[j];
        vb[u][c] = (XT)vd[j >= du ? j - du : j + m - du];
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < U; u++)
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const uint32_t j = jj[c], du = d + u;
        acc[c] += vf[u][c] * sx[j + du] + vb[u][c] * sx[j + m - du];
      }
  }
  for (; d <= d_hi; d++) {
    const VT* __restrict__ vd = vs + (uint64_t)(d - 1) * m;
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const uint32_t j = jj[c];
      const XT f = (XT)vd[j], bk = (XT)vd[j >= d ? j - d : j + m - d];
      acc[c] += f * sx[j + d] + bk * sx[j + m - d];
    }
  }
  if (half_step) {
    // last step of an even m: mh slots, the pair (i, i + mh) at slot i, read
    // by both of its members (forward term only)
    const VT* __restrict__ vd = vs + (uint64_t)(mh - 1) * m;
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const uint32_t j = jj[c];
      acc[c] += (XT)vd[j >= mh ? j - mh : j] * sx[j + mh];
    }
  }
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const uint32_t j = j0 + kWave * c + t;
    if (j < m) atomicAdd(&y[mem[D.moff + j]], acc[c]);
  }
}

// Value of lane l + 1 (lane 63: lane 0), a wave-wide DPP rotate.
// Every lane is written (full row and bank masks), so the "old" operand is
// the value itself and the rotate needs no separate initialisation.
__device__ __forceinline__ float wave_next(float v) {
  const int i = __float_as_int(v);
  return __int_as_float(__builtin_amdgcn_update_dpp(i, i, 0x134, 0xF, 0xF, false));
}
__device__ __forceinline__ double wave_next(double v) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const int l = (int)(uint32_t)u, h = (int)(uint32_t)(u >> 32);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(l, l, 0x134, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(h, h, 0x134, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ float lane63(float v) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ double lane63(double v) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 63);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// One-member rotation of a segment held in NC chunks (member j = 64 c + lane):
// member j takes member j + 1's value, member m - 1 member 0's.  A wave-wide
// rotate, then lane 63 of chunk c takes chunk c + 1's first member and the last
// member takes the first.
template <int NC, typename XT>
__device__ __forceinline__ void seg_rotate(XT (&v)[NC], uint32_t lane, uint32_t cl, uint32_t ll) {
  XT r[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) r[c] = wave_next(v[c]);
  const XT first = lane63(r[0]);  // member 0
#pragma unroll
  for (int c = 0; c < NC; c++) {
    XT t = r[c];
    if (c + 1 < NC) t = lane == 63 ? r[c + 1] : t;
    if ((uint32_t)c == cl) t = lane == ll ? first : t;
    v[c] = t;
  }
}

// Segments of 64 < m <= 64 NC members, one wavefront each (m <= 256): the
// travelling-sum scheme of seg_group_sums with one member per lane and chunk.
// At step d member j loads the value of the pair (j, j + d mod m) (a row of the
// segment read once, coalesced), adds v x_{j+d} to its sum and v x_j to its
// travelling sum; x_{j+d} and the travelling sums rotate one member per step.
// The half step of an even m reads x_{j+m/2} from LDS; the travelling sums are
// delivered through LDS.
template <int NC, typename VT, typename XT>
__device__ __forceinline__ void seg_wave_rows(const SegDesc& D, const VT* __restrict__ val,
                                              const uint32_t* __restrict__ mem, const XT* __restrict__ x,
                                              XT* __restrict__ y, XT* sx, XT* sr) {
  constexpr uint32_t U = sizeof(XT) == 8 ? 2 : 4;  // steps per batch of loads
  using gptr = const __attribute__((address_space(1))) VT*;
  const uint32_t m = D.m, lane = threadIdx.x;
  const uint32_t fs = (m - 1) / 2;  // steps with m slots
  const uint32_t cl = (m - 1) >> 6, ll = (m - 1) & 63;
  XT xj[NC], xp[NC], R[NC], acc[NC];
  uint32_t det[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const uint32_t j = 64 * c + lane;
    det[c] = mem[D.moff + (j < m ? j : 0)];
    const XT xv = x[det[c]];
    xj[c] = j < m ? xv : (XT)0;
    if (j < m) sx[j] = xv;
    xp[c] = xj[c];
    R[c] = acc[c] = (XT)0;
  }
  gptr vs = (gptr)(val + D.voff);
  // FP32 rows are read through the caches (the line two rows share stays for
  // the second), FP64 rows stream (measured faster for each)
  auto row_load = [](gptr p) {
    if constexpr (sizeof(VT) == 4) {
      return *p;
    } else {
      return __builtin_nontemporal_load(p);
    }
  };
  uint32_t d = 1;
  for (; d + U - 1 <= fs; d += U) {
    XT v[U][NC];
#pragma unroll
    for (uint32_t u = 0; u < U; u++)
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const uint32_t j = 64 * c + lane;
        v[u][c] = (XT)row_load(vs + (uint64_t)(d + u - 1) * m + (j < m ? j : 0));
      }
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      seg_rotate<NC, XT>(xp, lane, cl, ll);
      seg_rotate<NC, XT>(R, lane, cl, ll);
#pragma unroll
      for (int c = 0; c < NC; c++) {
        acc[c] = fma(v[u][c], xp[c], acc[c]);
        R[c] = fma(v[u][c], xj[c], R[c]);
      }
    }
  }
  for (; d <= fs; d++) {
    XT v[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const uint32_t j = 64 * c + lane;
      v[c] = (XT)row_load(vs + (uint64_t)(d - 1) * m + (j < m ? j : 0));
    }
    seg_rotate<NC, XT>(xp, lane, cl, ll);
    seg_rotate<NC, XT>(R, lane, cl, ll);
#pragma unroll
    for (int c = 0; c < NC; c++) {
      acc[c] = fma(v[c], xp[c], acc[c]);
      R[c] = fma(v[c], xj[c], R[c]);
    }
  }
  // the travelling sums: member j's now belongs to member j + fs mod m
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const uint32_t j = 64 * c + lane;
    const uint32_t to = j + fs >= m ? j + fs - m : j + fs;
    if (j < m) sr[to] = R[c];
  }
  __syncthreads();
  const uint32_t mh = m >> 1;
  const bool even = (m & 1) == 0;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const uint32_t j = 64 * c + lane;
    if (j >= m) continue;
    XT sum = acc[c] + sr[j];
    if (even) {
      // last step: mh slots, the pair (i, i + mh) at slot i, read by both members
      const VT w = val[D.voff + (uint64_t)(mh - 1) * m + (j >= mh ? j - mh : j)];
      sum += (XT)w * sx[j + mh >= m ? j + mh - m : j + mh];
    }
    atomicAdd(&y[det[c]], sum);
  }
}

// 64 < m <= 256: one wavefront per segment (seg_wave_rows), the chunk count
// dispatched per segment.
template <typename VT, typename XT>
__global__ __launch_bounds__(64) void k_seg_spmv_wave(const SegDesc* __restrict__ desc, uint32_t s0, uint32_t s1,
                                                      const uint32_t* __restrict__ mem, const VT* __restrict__ val,
                                                      const XT* __restrict__ x, XT* __restrict__ y) {
  __shared__ XT sx[256];
  __shared__ XT sr[256];
  const uint32_t s = s0 + blockIdx.x;
  if (s >= s1) return;
  const SegDesc D = desc[s];
  const uint32_t nc = (D.m + 63) / 64;
  switch (nc) {
    case 2: seg_wave_rows<2, VT, XT>(D, val, mem, x, y, sx, sr); break;
    case 3: seg_wave_rows<3, VT, XT>(D, val, mem, x, y, sx, sr); break;
    default: seg_wave_rows<4, VT, XT>(D, val, mem, x, y, sx, sr); break;
  }
}

template <int T, typename VT, typename XT>
__global__ __launch_bounds__(64) void k_seg_spmv_large(const SegDesc* __restrict__ desc, uint32_t s0, uint32_t s1,
                                                       const uint32_t* __restrict__ mem, const VT* __restrict__ val,
                                                       const XT* __restrict__ x, XT* __restrict__ y) {
  extern __shared__ __align__(16) unsigned char seg_lds[];
  const uint32_t s = s0 + blockIdx.x;
  if (s >= s1) return;
  const SegDesc D = desc[s];
  const uint32_t m = D.m, mh = m >> 1, t = threadIdx.x;
  const uint32_t j0 = blockIdx.y * (kWave * T);
  if (j0 >= m) return;
  if constexpr (sizeof(XT) == 8) {
    // an FP64 share with no steps (short segments of a split launch) is done
    const uint32_t fs = (m & 1) ? mh : mh - 1;
    const uint32_t per = (fs + gridDim.z - 1) / gridDim.z;
    if (1 + blockIdx.z * per > fs && !((m & 1) == 0 && blockIdx.z + 1 == gridDim.z)) return;
  }
  XT* sx = reinterpret_cast<XT*>(seg_lds);
  for (uint32_t j = t; j < m; j += kWave) {
    const XT v = x[mem[D.moff + j]];
    sx[j] = v;
    sx[j + m] = v;
  }
  __syncthreads();
  const uint32_t nc = min((uint32_t)T, (m - j0 + kWave - 1) / kWave);  // chunks holding members
  static_assert(T == 4, "chunk dispatch below");
  switch (nc) {
    case 1: seg_large_rows<1>(D, j0, val, sx, mem, y); break;
    case 2: seg_large_rows<2>(D, j0, val, sx, mem, y); break;
    case 3: seg_large_rows<3>(D, j0, val, sx, mem, y); break;
    default: seg_large_rows<4>(D, j0, val, sx, mem, y); break;
  }
}

// ---- Block product: 4 interleaved FP32 vectors (x[det * 4 + v]) ----
// Segments of m <= S members (S <= 16): a group of 4S lanes per segment, lane
// 4j + v holding vector v of member j.  The four lanes of a member read the
// same value slot, and their four sums leave in one atomic instruction that
// touches one cache line.
template <int S>
__global__ __launch_bounds__(256) void k_seg_spmm_small(const SegDesc* __restrict__ desc, uint32_t s0, uint32_t s1,
                                                        const uint32_t* __restrict__ mem, const float* __restrict__ val,
                                                        const float* __restrict__ x, float* __restrict__ y) {
  constexpr int L = 4 * S;
  constexpr int MH = S / 2;
  const uint64_t s = (uint64_t)s0 + (uint64_t)blockIdx.x * (256 / L) + threadIdx.x / L;
  const uint32_t l = threadIdx.x & (L - 1), j = l >> 2, v = l & 3;
  const int gbase = (int)(lane_id() & ~(uint32_t)(L - 1));
  SegDesc D{0u, 0u, 0ull};
  if (s < s1) D = desc[s];
  const uint32_t m = D.m, mh = m >> 1;
  const bool valid = j < m;
  const uint32_t det = mem[valid ? D.moff + j : 0];
  const float xj = x[(uint64_t)det * 4 + v] * (valid ? 1.f : 0.f);
  float acc = 0.f;
  uint32_t p = j, q = j;
#pragma unroll
  for (int d = 1; d <= MH; d++) {
    const bool live = valid && (uint32_t)d <= mh;
    const uint32_t vj = (2 * (uint32_t)d == m && j >= mh) ? j - mh : j;
    const uint64_t off = live ? (uint64_t)(d - 1) * m + vj : 0;
    const float w = val[D.voff + off] * (live ? 1.f : 0.f);
    const uint32_t pn = p + 1 == m ? 0 : p + 1, qn = q == 0 ? m - 1 : q - 1;
    p = valid ? pn : p;
    q = valid ? qn : q;
    const float wbr = __shfl(w, gbase + (int)(4 * q + v), kWave);
    const float wb = 2 * (uint32_t)d == m ? 0.f : wbr;
    const float xp = __shfl(xj, gbase + (int)(4 * p + v), kWave);
    const float xq = __shfl(xj, gbase + (int)(4 * q + v), kWave);
    acc += w * xp + wb * xq;
  }
  if (valid) atomicAdd(&y[(uint64_t)det * 4 + v], acc);
}

// Segments of m > 16 members: one wavefront per (segment, 64 members), x of
// all members and vectors in LDS, lane 4 jl + v holding vector v of the
// members j0 + jl + 16 c (c < 4); steps outermost as in k_seg_spmv_large.
__global__ __launch_bounds__(64) void k_seg_spmm_wave(const SegDesc* __restrict__ desc, uint32_t s0, uint32_t s1,
                                                      const uint32_t* __restrict__ mem, const float* __restrict__ val,
                                                      const float* __restrict__ x, float* __restrict__ y) {
  constexpr int T = 4;
  extern __shared__ __align__(16) unsigned char seg_lds[];
  const uint32_t s = s0 + blockIdx.x;
  if (s >= s1) return;
  const SegDesc D = desc[s];
  const uint32_t m = D.m, mh = m >> 1, t = threadIdx.x;
  const uint32_t j0 = blockIdx.y * 16 * T;
  if (j0 >= m) return;
  float* sx = reinterpret_cast<float*>(seg_lds);
  for (uint32_t e = t; e < 4 * m; e += kWave) sx[e] = x[(uint64_t)mem[D.moff + (e >> 2)] * 4 + (e & 3)];
  __syncthreads();
  const uint32_t v = t & 3, jl = t >> 2;
  const float* __restrict__ vs = val + D.voff;
  const bool even = (m & 1) == 0;
  float acc[T];
  uint32_t jj[T], p[T], q[T];
#pragma unroll
  for (int c = 0; c < T; c++) {
    const uint32_t j = j0 + jl + 16 * c;
    jj[c] = p[c] = q[c] = j < m ? j : 0;
    acc[c] = 0.f;
  }
  const float* __restrict__ vd = vs;
  for (uint32_t d = 1; d <= mh; d++, vd += m) {
    const bool half = even && d == mh;
    float vf[T], vb[T];
#pragma unroll
    for (int c = 0; c < T; c++) {
      p[c] = p[c] + 1 == m ? 0 : p[c] + 1;
      q[c] = q[c] == 0 ? m - 1 : q[c] - 1;
      const uint32_t fi = half && jj[c] >= mh ? jj[c] - mh : jj[c];
      vf[c] = vd[fi];
      vb[c] = vd[half ? 0 : q[c]];
    }
#pragma unroll
    for (int c = 0; c < T; c++) {
      const float wb = half ? 0.f : vb[c];
      acc[c] += vf[c] * sx[p[c] * 4 + v] + wb * sx[q[c] * 4 + v];
    }
  }
#pragma unroll
  for (int c = 0; c < T; c++) {
    const uint32_t j = j0 + jl + 16 * c;
    if (j < m) atomicAdd(&y[(uint64_t)mem[D.moff + j] * 4 + v], acc[c]);
  }
}

// ---- RDM over segments: every stored pair once (one wavefront per segment) ----
__global__ __launch_bounds__(256) void k_seg_rdm(RdmParams R, const SegDesc* __restrict__ desc,
                                                 const uint32_t* __restrict__ mem, const uint8_t* __restrict__ type,
                                                 const uint64_t* __restrict__ core, const double* __restrict__ v64,
                                                 const float* __restrict__ v32, const uint32_t* __restrict__ dnord,
                                                 uint32_t ra, uint32_t rb, int keep_zeros, uint64_t n_seg) {
  const uint64_t s = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= n_seg) return;
  const SegDesc D = desc[s];
  const uint32_t ty = type[s];
  const uint64_t cs = core[s];
  const uint64_t slots = (uint64_t)D.m * (D.m - 1) / 2;
  for (uint64_t k = lane_id(); k < slots; k += kWave) {
    const uint32_t d = (uint32_t)(k / D.m) + 1, j = (uint32_t)(k % D.m);
    uint32_t p = j + d;
    if (p >= D.m) p -= D.m;
    uint32_t a = mem[D.moff + j], b = mem[D.moff + p];
    if (ty == kSegA) {
      a = dnord[a];
      b = dnord[b];
    }
    const uint64_t ua = R.up[a], da = R.dn[a], ub = R.up[b], db = R.dn[b];
    if (!seg_owns(ty, cs, ty == kSegA ? ra : rb, ua, da, ub, db)) continue;
    if (!keep_zeros && !(v64 ? v64[D.voff + k] != 0.0 : v32[D.voff + k] != 0.0f)) {
      // a cut pair (|H| below 1e-12): not a connection, but a single of
      // the det list still enters the 1-RDM
      rdm1_single(R, ua, da, a, ub, db, b);
      continue;
    }
    rdm_pair(R, ua, da, a, ub, db, b);
  }
}

// The diagonal pairs (i = j) of the RDMs, one thread per det: what
// rdm_pair's i = j branch writes, summed per workgroup in LDS and then added
// to the results once.  Its 2-RDM elements are of two kinds, (ss, rr) and
// (sr, rs) for occupied s, r (packed a = s n + s, b = r n + r, or a = s n +
// r, b = r n + s, stored when a >= b), kept in two triangular arrays C and X
// indexed by (hi, lo), hi >= lo; the 1-RDM diagonal when the pairs write the
// 1-RDM.  Dynamic LDS: (n (n + 1) + n) doubles.
__global__ __launch_bounds__(256) void k_rdm_diag(RdmParams R) {
  extern __shared__ __align__(16) unsigned char rdiag_lds[];
  const int n = R.n_orb;
  const uint32_t nt = (uint32_t)(n * (n + 1) / 2);
  double* C = reinterpret_cast<double*>(rdiag_lds);
  double* X = C + nt;
  double* D1 = X + nt;
  for (uint32_t k = threadIdx.x; k < 2 * nt + (uint32_t)n; k += blockDim.x) C[k] = 0.0;
  __syncthreads();
  auto tri = [](int hi, int lo) { return (uint32_t)(hi * (hi + 1) / 2 + lo); };
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < R.n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t u = R.up[i], d = R.dn[i];
    const double cii = cc(R, i, i);
    if (R.pair1) {
      for (uint64_t m = u; m; m &= m - 1) atomicAdd(&D1[ctz64(m)], cii);
      for (uint64_t m = d; m; m &= m - 1) atomicAdd(&D1[ctz64(m)], cii);
    }
    // same-spin pairs s < r: w2(r, s, s, r, +1) and w2(r, s, r, s, -1) are
    // the stored ones
    for (int spin = 0; spin < 2; spin++) {
      const uint64_t sp = spin == 0 ? u : d;
      for (uint64_t a = sp; a; a &= a - 1)
        for (uint64_t b = a & (a - 1); b; b &= b - 1) {
          const int lo = ctz64(a), hi = ctz64(b);
          atomicAdd(&C[tri(hi, lo)], cii);
          atomicAdd(&X[tri(hi, lo)], -cii);
        }
    }
    // opposite-spin pairs (s alpha, r beta): w2(s, r, r, s) stored when s >=
    // r, w2(r, s, s, r) when r >= s
    for (uint64_t a = u; a; a &= a - 1)
      for (uint64_t b = d; b; b &= b - 1) {
        const int sa = ctz64(a), rb = ctz64(b);
        if (sa >= rb) atomicAdd(&C[tri(sa, rb)], cii);
        if (rb >= sa) atomicAdd(&C[tri(rb, sa)], cii);
      }
  }
  __syncthreads();
  const uint64_t nn = (uint64_t)n;
  for (uint32_t k = threadIdx.x; k < nt; k += blockDim.x) {
    int hi = 0;
    while ((uint32_t)((hi + 1) * (hi + 2) / 2) <= k) hi++;
    const int lo = (int)(k - (uint32_t)(hi * (hi + 1) / 2));
    if (C[k] != 0.0) {
      const uint64_t a = (uint64_t)hi * nn + hi, b = (uint64_t)lo * nn + lo;
      atomicAdd(&R.rdm2[(a * (a + 1)) / 2 + b], C[k]);
    }
    if (X[k] != 0.0) {
      const uint64_t a = (uint64_t)hi * nn + lo, b = (uint64_t)lo * nn + hi;
      atomicAdd(&R.rdm2[(a * (a + 1)) / 2 + b], X[k]);
    }
  }
  if (R.pair1)
    for (uint32_t k = threadIdx.x; k < (uint32_t)n; k += blockDim.x)
      if (D1[k] != 0.0) atomicAdd(&R.rdm1[k * nn + k], D1[k]);
}


// ---------------------------------------------------------------------------
// Host-side helpers
// ---------------------------------------------------------------------------
template <typename KT>
int radix_sort_pairs(shci_ctx* ctx, const KT* kin, KT* kout, const uint32_t* vin, uint32_t* vout, uint64_t n,
                     int end_bit) {
  size_t tb = 0;
  HIPC(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, kout, vin, vout, (int)n, 0, end_bit, ctx->stream));
  CHECK_RC(ensure(ctx, ctx->tmp, tb));
  HIPC(hipcub::DeviceRadixSort::SortPairs(ctx->tmp.p, tb, kin, kout, vin, vout, (int)n, 0, end_bit, ctx->stream));
  return SHCI_OK;
}

// Prefix sums: rocPRIM's lookback scan with the block shapes of its gfx942
// tuning for 4- and 8-byte sums (256 threads x 24 or 15 items, reduce-then-
// scan blocks); on the MI355X they beat the shapes it picks for gfx950 (the
// two scans over a 1e7-det build's 1.9e9 memberships: 5.6 -> 4.5 ms each).
template <typename T>
using ScanCfg = rocprim::scan_config<256, sizeof(T) == 4 ? 24 : 15, rocprim::block_load_method::block_load_transpose,
                                     rocprim::block_store_method::block_store_transpose,
                                     rocprim::block_scan_algorithm::reduce_then_scan>;

template <typename T>
int incl_scan(shci_ctx* ctx, const T* in, T* out, uint64_t n) {
  size_t tb = 0;
  HIPC(rocprim::inclusive_scan<ScanCfg<T>>(nullptr, tb, in, out, n, rocprim::plus<T>(), ctx->stream));
  CHECK_RC(ensure(ctx, ctx->tmp, tb));
  HIPC(rocprim::inclusive_scan<ScanCfg<T>>(ctx->tmp.p, tb, in, out, n, rocprim::plus<T>(), ctx->stream));
  return SHCI_OK;
}

template <typename TI, typename TO>
int excl_scan(shci_ctx* ctx, const TI* in, TO* out, uint64_t n) {
  size_t tb = 0;
  HIPC(rocprim::exclusive_scan<ScanCfg<TO>>(nullptr, tb, in, out, (TO)0, n, rocprim::plus<TO>(), ctx->stream));
  CHECK_RC(ensure(ctx, ctx->tmp, tb));
  HIPC(rocprim::exclusive_scan<ScanCfg<TO>>(ctx->tmp.p, tb, in, out, (TO)0, n, rocprim::plus<TO>(), ctx->stream));
  return SHCI_OK;
}

int bits_for(uint64_t v) {
  int b = 1;
  while (b < 64 && (v >> b) != 0) b++;
  return b;
}

Ints make_ints(shci_ctx* ctx) {
  Ints I;
  I.h1 = P<double>(ctx->h1);
  I.eri = P<double>(ctx->eri);
  I.tD = P<double>(ctx->tD);
  I.tX = P<double>(ctx->tX);
  I.n = (int)ctx->n_orb;
  I.core = ctx->core;
  return I;
}

// Group the determinants: (alpha, beta) sort, alpha groups, beta order + groups.
int build_groups(shci_ctx* ctx) {
  const uint64_t n = ctx->n;
  const int nb = (int)ctx->n_orb;
  const hipStream_t st = ctx->stream;
  const uint32_t g256 = div_up(n, 256);
  CHECK_RC(ensure(ctx, ctx->ka, n * 8));
  CHECK_RC(ensure(ctx, ctx->kb, n * 8));
  CHECK_RC(ensure(ctx, ctx->va, n * 4));
  CHECK_RC(ensure(ctx, ctx->vb, n * 4));
  CHECK_RC(ensure(ctx, ctx->perm, n * 4));
  CHECK_RC(ensure(ctx, ctx->iperm, n * 4));
  CHECK_RC(ensure(ctx, ctx->up_s, n * 8));
  CHECK_RC(ensure(ctx, ctx->dn_s, n * 8));
  // LSD: stable sort by beta, then by alpha -> (alpha, beta) order.
  k_iota<<<g256, 256, 0, st>>>(P<uint32_t>(ctx->va), n);
  CHECK_RC(radix_sort_pairs(ctx, P<uint64_t>(ctx->dn_c), P<uint64_t>(ctx->kb), P<uint32_t>(ctx->va),
                            P<uint32_t>(ctx->vb), n, nb));
  k_gather_u64<<<g256, 256, 0, st>>>(P<uint64_t>(ctx->up_c), P<uint32_t>(ctx->vb), P<uint64_t>(ctx->ka), n);
  CHECK_RC(radix_sort_pairs(ctx, P<uint64_t>(ctx->ka), P<uint64_t>(ctx->up_s), P<uint32_t>(ctx->vb),
                            P<uint32_t>(ctx->perm), n, nb));
  k_gather_u64<<<g256, 256, 0, st>>>(P<uint64_t>(ctx->dn_c), P<uint32_t>(ctx->perm), P<uint64_t>(ctx->dn_s), n);
  k_inverse_perm<<<g256, 256, 0, st>>>(P<uint32_t>(ctx->perm), P<uint32_t>(ctx->iperm), n);
  {
    CHECK_RC(ensure(ctx, ctx->scal, 64 * 8));
    unsigned int* bad = reinterpret_cast<unsigned int*>(ctx->scal.p);
    unsigned int hb[2] = {0, 0};
    HIPC(hipMemsetAsync(bad, 0, 8, st));
    k_duplicates<<<g256, 256, 0, st>>>(P<uint64_t>(ctx->up_s), P<uint64_t>(ctx->dn_s), n, bad);
    HIPC(hipMemcpyAsync(hb, bad, 8, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (hb[1]) return fail(ctx, SHCI_ERR_ARG, std::to_string(hb[1]) + " duplicate determinants");
  }
  // Alpha groups.
  CHECK_RC(ensure(ctx, ctx->up_gid, n * 4));
  CHECK_RC(ensure(ctx, ctx->up_gstart, (n + 1) * 4));
  k_head_flags<<<g256, 256, 0, st>>>(P<uint64_t>(ctx->up_s), P<uint32_t>(ctx->va), n);
  CHECK_RC(incl_scan(ctx, P<uint32_t>(ctx->va), P<uint32_t>(ctx->vb), n));
  k_group_starts<<<g256, 256, 0, st>>>(P<uint32_t>(ctx->va), P<uint32_t>(ctx->vb), P<uint32_t>(ctx->up_gid),
                                       P<uint32_t>(ctx->up_gstart), n);
  HIPC(hipMemcpyAsync(&ctx->n_ugroups, P<uint32_t>(ctx->vb) + (n - 1), 4, hipMemcpyDeviceToHost, st));
  // Beta order (stable sort of internal indices by beta) and beta groups.
  CHECK_RC(ensure(ctx, ctx->dnord_idx, n * 4));
  CHECK_RC(ensure(ctx, ctx->dnord_up, n * 8));
  CHECK_RC(ensure(ctx, ctx->dn_gid, n * 4));
  CHECK_RC(ensure(ctx, ctx->dn_gstart, (n + 1) * 4));
  k_iota<<<g256, 256, 0, st>>>(P<uint32_t>(ctx->va), n);
  CHECK_RC(radix_sort_pairs(ctx, P<uint64_t>(ctx->dn_s), P<uint64_t>(ctx->kb), P<uint32_t>(ctx->va),
                            P<uint32_t>(ctx->dnord_idx), n, nb));
  k_head_flags<<<g256, 256, 0, st>>>(P<uint64_t>(ctx->kb), P<uint32_t>(ctx->va), n);
  CHECK_RC(incl_scan(ctx, P<uint32_t>(ctx->va), P<uint32_t>(ctx->vb), n));
  // group id in beta order -> reuse ka as a u32 scratch for kgid
  uint32_t* kgid = reinterpret_cast<uint32_t*>(P<uint64_t>(ctx->ka));
  k_group_starts<<<g256, 256, 0, st>>>(P<uint32_t>(ctx->va), P<uint32_t>(ctx->vb), kgid, P<uint32_t>(ctx->dn_gstart), n);
  HIPC(hipMemcpyAsync(&ctx->n_dgroups, P<uint32_t>(ctx->vb) + (n - 1), 4, hipMemcpyDeviceToHost, st));
  k_dn_order<<<g256, 256, 0, st>>>(P<uint32_t>(ctx->dnord_idx), P<uint64_t>(ctx->up_s), kgid,
                                   P<uint64_t>(ctx->dnord_up), P<uint32_t>(ctx->dn_gid), n);
  HIPC(hipStreamSynchronize(st));
  return SHCI_OK;
}

// Minus-one (rm = 1) or minus-two (rm = 2) table of G unique strings with
// n_e electrons each: id[g*P+k] = index of the distinct string "g without its
// k-th electron (pair)", P = n_e or C(n_e, 2); the number of distinct strings
// goes to *R.
int build_m1(shci_ctx* ctx, const uint64_t* uniq, uint32_t G, uint32_t ne, DevBuf& id, uint32_t* R, uint32_t rm = 1) {
  const hipStream_t st = ctx->stream;
  const uint64_t m = (uint64_t)G * (rm == 1 ? ne : (uint64_t)ne * (ne - 1) / 2);
  CHECK_RC(ensure(ctx, id, std::max<uint64_t>(m, 1) * 4));
  if (m == 0) {
    *R = 0;
    return SHCI_OK;
  }
  DevBuf keys, keys_s, vals, vals_s, flag, incl;
  auto cleanup = [&]() {
    release(ctx, keys);
    release(ctx, keys_s);
    release(ctx, vals);
    release(ctx, vals_s);
    release(ctx, flag);
    release(ctx, incl);
  };
  int rc = SHCI_OK;
  if ((rc = ensure(ctx, keys, m * 8)) || (rc = ensure(ctx, keys_s, m * 8)) || (rc = ensure(ctx, vals, m * 4)) ||
      (rc = ensure(ctx, vals_s, m * 4)) || (rc = ensure(ctx, flag, m * 4)) || (rc = ensure(ctx, incl, m * 4))) {
    cleanup();
    return rc;
  }
  if (rm == 1)
    k_minus_one<<<div_up(G, 256), 256, 0, st>>>(uniq, G, ne, P<uint64_t>(keys), P<uint32_t>(vals));
  else
    k_minus_two<<<div_up(G, 256), 256, 0, st>>>(uniq, G, ne, P<uint64_t>(keys), P<uint32_t>(vals));
  rc = radix_sort_pairs(ctx, P<uint64_t>(keys), P<uint64_t>(keys_s), P<uint32_t>(vals), P<uint32_t>(vals_s), m,
                        (int)ctx->n_orb);
  if (rc == SHCI_OK) {
    k_head_flags<<<div_up(m, 256), 256, 0, st>>>(P<uint64_t>(keys_s), P<uint32_t>(flag), m);
    rc = incl_scan(ctx, P<uint32_t>(flag), P<uint32_t>(incl), m);
  }
  if (rc == SHCI_OK) {
    k_scatter_run_id<<<div_up(m, 256), 256, 0, st>>>(P<uint32_t>(vals_s), P<uint32_t>(incl), m, P<uint32_t>(id));
    if (hipMemcpyAsync(R, P<uint32_t>(incl) + (m - 1), 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = fail(ctx, SHCI_ERR_HIP, "minus-one table");
  }
  cleanup();
  return rc;
}

// Minus-one tables of the unique alpha and beta strings (and, for the segment
// layout, the minus-two tables), then for the CSR search the sorted
// mixed-double segments.
int build_segments(shci_ctx* ctx, bool segl) {
  const hipStream_t st = ctx->stream;
  const uint64_t n = ctx->n;
  const uint32_t G = ctx->n_ugroups, H = ctx->n_dgroups, nu = ctx->n_up, nd = ctx->n_dn;
  {
    DevBuf uniq;
    CHECK_RC(ensure(ctx, uniq, (uint64_t)std::max(G, H) * 8));
    k_unique_strings<<<div_up(G, 256), 256, 0, st>>>(P<uint64_t>(ctx->up_s), P<uint32_t>(ctx->up_gstart), G,
                                                     P<uint64_t>(uniq));
    int rc = build_m1(ctx, P<uint64_t>(uniq), G, nu, ctx->m1a_id, &ctx->r_alpha);
    if (rc == SHCI_OK && segl && nu >= 2) rc = build_m1(ctx, P<uint64_t>(uniq), G, nu, ctx->mA_id, &ctx->r2_alpha, 2);
    if (rc == SHCI_OK) {
      // beta strings in beta order: dn_s[dnord_idx[dn_gstart[h]]] -- dnord_up holds alpha, so
      // gather the sorted beta keys through the beta-order index.
      k_gather_u64<<<div_up(n, 256), 256, 0, st>>>(P<uint64_t>(ctx->dn_s), P<uint32_t>(ctx->dnord_idx),
                                                   P<uint64_t>(ctx->ka), n);
      k_unique_strings<<<div_up(H, 256), 256, 0, st>>>(P<uint64_t>(ctx->ka), P<uint32_t>(ctx->dn_gstart), H,
                                                       P<uint64_t>(uniq));
      rc = build_m1(ctx, P<uint64_t>(uniq), H, nd, ctx->m1b_id, &ctx->r_beta);
      if (rc == SHCI_OK && segl && nd >= 2) rc = build_m1(ctx, P<uint64_t>(uniq), H, nd, ctx->mB_id, &ctx->r2_beta, 2);
    }
    release(ctx, uniq);
    CHECK_RC(rc);
  }
  if (segl) {
    // The segment layout sorts its own keys; the CSR search tables go.
    DevBuf* csr[] = {&ctx->dm1_vals, &ctx->dm1_rstart, &ctx->dm1_seg};
    for (DevBuf* b : csr) release(ctx, *b);
    ctx->n_dm1 = ctx->n_dm1_runs = 0;
    return SHCI_OK;
  }
  const uint32_t K = nu * nd;
  const uint64_t m = n * K;
  ctx->n_dm1 = m;
  if (m == 0 || m >= (1ull << 31)) {
    if (m >= (1ull << 31)) return fail(ctx, SHCI_ERR_UNSUPPORTED, "n_dets * n_up * n_dn >= 2^31");
    ctx->n_dm1_runs = 0;
    return SHCI_OK;
  }
  const int bb = bits_for(ctx->r_beta), ba = bits_for(ctx->r_alpha);
  if (ba + bb > 64) return fail(ctx, SHCI_ERR_UNSUPPORTED, "segment keys exceed 64 bits");
  DevBuf keys, keys_s, vals, flag, incl;
  auto cleanup = [&]() {
    release(ctx, keys);
    release(ctx, keys_s);
    release(ctx, vals);
    release(ctx, flag);
    release(ctx, incl);
  };
  int rc = SHCI_OK;
  if ((rc = ensure(ctx, keys, m * 8)) || (rc = ensure(ctx, keys_s, m * 8)) || (rc = ensure(ctx, vals, m * 4)) ||
      (rc = ensure(ctx, ctx->dm1_vals, m * 4)) || (rc = ensure(ctx, ctx->dm1_seg, m * 4))) {
    cleanup();
    return rc;
  }
  k_dm1_keys<<<div_up(m, 256), 256, 0, st>>>(P<uint32_t>(ctx->up_gid), P<uint32_t>(ctx->dn_gid),
                                             P<uint32_t>(ctx->m1a_id), P<uint32_t>(ctx->m1b_id), n, nu, nd, bb,
                                             P<uint64_t>(keys), P<uint32_t>(vals));
  rc = radix_sort_pairs(ctx, P<uint64_t>(keys), P<uint64_t>(keys_s), P<uint32_t>(vals), P<uint32_t>(ctx->dm1_vals),
                        m, ba + bb);
  release(ctx, keys);
  release(ctx, vals);
  if (rc == SHCI_OK) rc = ensure(ctx, flag, m * 4);
  if (rc == SHCI_OK) rc = ensure(ctx, incl, m * 4);
  if (rc == SHCI_OK) {
    k_head_flags<<<div_up(m, 256), 256, 0, st>>>(P<uint64_t>(keys_s), P<uint32_t>(flag), m);
    rc = incl_scan(ctx, P<uint32_t>(flag), P<uint32_t>(incl), m);
  }
  release(ctx, keys_s);
  if (rc == SHCI_OK) rc = ensure(ctx, ctx->dm1_rstart, (m + 1) * 4);
  if (rc == SHCI_OK) {
    k_run_starts<<<div_up(m, 256), 256, 0, st>>>(P<uint32_t>(flag), P<uint32_t>(incl), P<uint32_t>(ctx->dm1_rstart),
                                                 m);
    k_scatter_run_id<<<div_up(m, 256), 256, 0, st>>>(P<uint32_t>(ctx->dm1_vals), P<uint32_t>(incl), m,
                                                     P<uint32_t>(ctx->dm1_seg));
    uint32_t runs = 0;
    if (hipMemcpyAsync(&runs, P<uint32_t>(incl) + (m - 1), 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = fail(ctx, SHCI_ERR_HIP, "segments");
    ctx->n_dm1_runs = runs;
  }
  cleanup();
  return rc;
}


// ---------------------------------------------------------------------------
// Segment layout build (SHCI_LAYOUT_SEGMENTS).
// ---------------------------------------------------------------------------
// Device buffer released when it leaves scope (build temporaries).
struct TmpBuf {
  shci_ctx* ctx;
  DevBuf b;
  explicit TmpBuf(shci_ctx* c) : ctx(c) {}
  ~TmpBuf() { release(ctx, b); }
  TmpBuf(const TmpBuf&) = delete;
  TmpBuf& operator=(const TmpBuf&) = delete;
  template <class T>
  T* p() const {
    return static_cast<T*>(b.p);
  }
};

template <typename T>
int copy_back(shci_ctx* ctx, T* host, const T* dev, size_t count = 1) {
  HIPC(hipMemcpyAsync(host, dev, count * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return SHCI_OK;
}

SegKeyParams seg_key_params(shci_ctx* ctx) {
  SegKeyParams K;
  const uint32_t nu = ctx->n_up, nd = ctx->n_dn;
  K.n = ctx->n;
  K.nu = nu;
  K.nd = nd;
  K.ra = ctx->ra;
  K.rb = ctx->rb;
  K.PC = nu * nd;
  K.PA = nu >= 2 ? nu * (nu - 1) / 2 : nu;
  K.PB = nd >= 2 ? nd * (nd - 1) / 2 : nd;
  K.up_gid = P<uint32_t>(ctx->up_gid);
  K.dn_gid = P<uint32_t>(ctx->dn_gid);
  K.m1a = P<uint32_t>(ctx->m1a_id);
  K.m1b = P<uint32_t>(ctx->m1b_id);
  K.mA = nu >= 2 ? P<uint32_t>(ctx->mA_id) : P<uint32_t>(ctx->m1a_id);
  K.mB = nd >= 2 ? P<uint32_t>(ctx->mB_id) : P<uint32_t>(ctx->m1b_id);
  K.up = P<uint64_t>(ctx->up_s);
  K.dn = P<uint64_t>(ctx->dn_s);
  K.dnpos = P<uint32_t>(ctx->dnpos);
  const uint32_t rA = nu >= 2 ? ctx->r2_alpha : ctx->r_alpha, rB = nd >= 2 ? ctx->r2_beta : ctx->r_beta;
  K.bC = bits_for(ctx->r_beta);
  K.bA = bits_for(rA);
  K.bB = bits_for(rB);
  K.tag_shift = std::max({bits_for(ctx->r_alpha) + K.bC, bits_for(ctx->n_dgroups) + K.bA,
                          bits_for(ctx->n_ugroups) + K.bB});
  K.eC = K.n * K.PC;
  K.eA = K.eC + K.n * K.PA;
  K.dC = div_mul(K.PC);
  K.dA = div_mul(K.PA);
  K.dB = div_mul(K.PB);
  return K;
}

// Segments (sorted by size, largest first), their members and slot offsets.
int build_seg_layout(shci_ctx* ctx) {
  const hipStream_t st = ctx->stream;
  CHECK_RC(ensure(ctx, ctx->dnpos, ctx->n * 4));
  k_inverse_perm<<<div_up(ctx->n, 256), 256, 0, st>>>(P<uint32_t>(ctx->dnord_idx), P<uint32_t>(ctx->dnpos), ctx->n);
  const SegKeyParams K = seg_key_params(ctx);
  const uint64_t M = K.n * ((uint64_t)K.PC + K.PA + K.PB);
  ctx->n_seg = ctx->n_mem = ctx->n_slots = ctx->n_tiles = 0;
  for (auto& c : ctx->seg_cs) c = 0;
  for (auto& c : ctx->seg_maxm) c = 0;
  if (M == 0) return SHCI_OK;
  if (M >= (1ull << 31)) return fail(ctx, SHCI_ERR_UNSUPPORTED, "segment layout: n_dets * keys per det >= 2^31");
  if (K.tag_shift + 2 > 64) return fail(ctx, SHCI_ERR_UNSUPPORTED, "segment keys exceed 64 bits");
  // The sort needs ~36 B per key; value stores of a previous build are
  // reallocated after it anyway, so give them up first when memory is short
  // (their chunks then serve the sort's buffers).
  if ((double)available_bytes(ctx) < 40.0 * (double)M) {
    DevBuf* big[] = {&ctx->sv64, &ctx->sv32, &ctx->seg_mem, &ctx->seg_desc, &ctx->seg_core, &ctx->seg_type,
                     &ctx->seg_loc, &ctx->tile_tab, &ctx->tile_rec, &ctx->dav, &ctx->tmp};
    for (DevBuf* b : big) release(ctx, *b);
  }
  // Keys of every (det, core) membership, radix-sorted: one 64-bit sort with
  // the region tag on top, or, when each region's key fits 32 bits, three
  // 32-bit sorts of the regions (the same order, a third less traffic).
  const uint64_t eC = K.n * K.PC, eA = eC + K.n * K.PA;
  const int wC = bits_for(ctx->r_alpha) + K.bC;
  const int wA = bits_for(ctx->n_dgroups) + K.bA, wB = bits_for(ctx->n_ugroups) + K.bB;
  const bool k32 = std::max({wC, wA, wB}) <= 32;
  TmpBuf keys(ctx), keys_s(ctx), vals(ctx), vals_s(ctx);
  CHECK_RC(ensure(ctx, keys.b, M * 8));  // later holds the head flags and their scan
  CHECK_RC(ensure(ctx, vals.b, M * 4));
  CHECK_RC(ensure(ctx, keys_s.b, M * (k32 ? 4 : 8)));
  CHECK_RC(ensure(ctx, vals_s.b, M * 4));
  uint32_t* flag = reinterpret_cast<uint32_t*>(keys.b.p);
  uint32_t* incl = flag + M;
  if (k32) {
    uint32_t* k32in = keys.p<uint32_t>();
    k_seg_keys32<<<div_up(M, 256), 256, 0, st>>>(K, M, k32in, vals.p<uint32_t>());
    const uint64_t rb[4] = {0, eC, eA, M};
    const int w[3] = {wC, wA, wB};
    for (int r = 0; r < 3; r++)
      if (rb[r + 1] > rb[r])
        CHECK_RC(radix_sort_pairs(ctx, k32in + rb[r], keys_s.p<uint32_t>() + rb[r], vals.p<uint32_t>() + rb[r],
                                  vals_s.p<uint32_t>() + rb[r], rb[r + 1] - rb[r], w[r]));
    release(ctx, vals.b);
    release(ctx, ctx->tmp);
    k_head_flags32<<<div_up(M, 256), 256, 0, st>>>(keys_s.p<uint32_t>(), flag, M, eC, eA);
  } else {
    k_seg_keys<<<div_up(M, 256), 256, 0, st>>>(K, M, keys.p<uint64_t>(), vals.p<uint32_t>());
    CHECK_RC(radix_sort_pairs(ctx, keys.p<uint64_t>(), keys_s.p<uint64_t>(), vals.p<uint32_t>(),
                              vals_s.p<uint32_t>(), M, K.tag_shift + 2));
    release(ctx, vals.b);
    release(ctx, ctx->tmp);
    // Runs of equal keys: head flags and their inclusive scan share the old key buffer.
    k_head_flags<<<div_up(M, 256), 256, 0, st>>>(keys_s.p<uint64_t>(), flag, M);
  }
  HIPC(hipGetLastError());
  release(ctx, keys_s.b);
  CHECK_RC(incl_scan(ctx, flag, incl, M));
  uint32_t R = 0;
  CHECK_RC(copy_back(ctx, &R, incl + (M - 1)));
  TmpBuf rstart(ctx), keep(ctx), sidx(ctx), useinc(ctx);
  CHECK_RC(ensure(ctx, rstart.b, ((uint64_t)R + 1) * 4));
  CHECK_RC(ensure(ctx, keep.b, ((uint64_t)R + 1) * 4));
  CHECK_RC(ensure(ctx, sidx.b, ((uint64_t)R + 1) * 4));
  CHECK_RC(ensure(ctx, useinc.b, M * 4));
  k_run_starts<<<div_up(M, 256), 256, 0, st>>>(flag, incl, rstart.p<uint32_t>(), M);
  // Member flags (reusing the head-flag array) and their inclusive scan.
  uint32_t* use = flag;
  k_seg_useful<<<div_up(M, 256), 256, 0, st>>>(K, vals_s.p<uint32_t>(), incl, rstart.p<uint32_t>(), M, use);
  CHECK_RC(incl_scan(ctx, use, useinc.p<uint32_t>(), M));
  CHECK_RC(ensure(ctx, ctx->scal, 64 * 8));
  unsigned int* bad = reinterpret_cast<unsigned int*>(ctx->scal.p);
  HIPC(hipMemsetAsync(bad, 0, 4, st));
  k_seg_keep<<<div_up((uint64_t)R + 1, 256), 256, 0, st>>>(rstart.p<uint32_t>(), useinc.p<uint32_t>(), R,
                                                          keep.p<uint32_t>(), bad);
  CHECK_RC(excl_scan(ctx, keep.p<uint32_t>(), sidx.p<uint32_t>(), (uint64_t)R + 1));
  uint32_t S = 0, hbad = 0;
  CHECK_RC(copy_back(ctx, &S, sidx.p<uint32_t>() + R));
  CHECK_RC(copy_back(ctx, &hbad, bad));
  if (hbad) return fail(ctx, SHCI_ERR_UNSUPPORTED, "segment with more than 4096 members");
  if (S == 0) return SHCI_OK;
  // Tiles: the kept runs' spin groups cut at kTileCap memberships.
  TmpBuf tileinc(ctx);
  uint32_t n_tiles = 0;
  {
    TmpBuf gkey(ctx), tw(ctx), gflag(ctx), gord(ctx), W(ctx), gW(ctx);
    CHECK_RC(ensure(ctx, gkey.b, (uint64_t)S * 8));
    CHECK_RC(ensure(ctx, tw.b, (uint64_t)S * 4));
    CHECK_RC(ensure(ctx, gflag.b, (uint64_t)S * 4));
    k_seg_groups<<<div_up(R, 256), 256, 0, st>>>(K, vals_s.p<uint32_t>(), rstart.p<uint32_t>(), useinc.p<uint32_t>(),
                                                 keep.p<uint32_t>(), sidx.p<uint32_t>(), R, gkey.p<uint64_t>(),
                                                 tw.p<uint32_t>());
    k_seg_gflags<<<div_up(S, 256), 256, 0, st>>>(gkey.p<uint64_t>(), S, gflag.p<uint32_t>());
    release(ctx, gkey.b);
    CHECK_RC(ensure(ctx, gord.b, (uint64_t)S * 4));
    CHECK_RC(ensure(ctx, W.b, (uint64_t)S * 4));
    CHECK_RC(ensure(ctx, gW.b, (uint64_t)S * 4));
    CHECK_RC(incl_scan(ctx, gflag.p<uint32_t>(), gord.p<uint32_t>(), S));
    CHECK_RC(incl_scan(ctx, tw.p<uint32_t>(), W.p<uint32_t>(), S));
    k_seg_gstart<<<div_up(S, 256), 256, 0, st>>>(gflag.p<uint32_t>(), gord.p<uint32_t>(), W.p<uint32_t>(),
                                                 tw.p<uint32_t>(), S, gW.p<uint32_t>());
    CHECK_RC(ensure(ctx, gkey.b, (uint64_t)S * 4));
    uint32_t* tflag = gkey.p<uint32_t>();
    k_seg_tflags<<<div_up(S, 256), 256, 0, st>>>(gflag.p<uint32_t>(), gord.p<uint32_t>(), W.p<uint32_t>(),
                                                 tw.p<uint32_t>(), gW.p<uint32_t>(), S, tflag);
    HIPC(hipGetLastError());
    CHECK_RC(ensure(ctx, tileinc.b, (uint64_t)S * 4));
    CHECK_RC(incl_scan(ctx, tflag, tileinc.p<uint32_t>(), S));
    CHECK_RC(copy_back(ctx, &n_tiles, tileinc.p<uint32_t>() + (S - 1)));
  }
  // Order the segments: size class, type, tile (k_seg_compact).
  TmpBuf skey(ctx), srun(ctx), skey_s(ctx), srun_s(ctx);
  CHECK_RC(ensure(ctx, skey.b, (uint64_t)S * 8));
  CHECK_RC(ensure(ctx, srun.b, (uint64_t)S * 4));
  CHECK_RC(ensure(ctx, skey_s.b, (uint64_t)S * 8));
  CHECK_RC(ensure(ctx, srun_s.b, (uint64_t)S * 4));
  k_seg_compact<<<div_up(R, 256), 256, 0, st>>>(K, vals_s.p<uint32_t>(), rstart.p<uint32_t>(), useinc.p<uint32_t>(),
                                                keep.p<uint32_t>(), sidx.p<uint32_t>(), tileinc.p<uint32_t>(), R,
                                                skey.p<uint64_t>(), srun.p<uint32_t>());
  release(ctx, keep.b);
  release(ctx, sidx.b);
  release(ctx, tileinc.b);
  CHECK_RC(radix_sort_pairs(ctx, skey.p<uint64_t>(), skey_s.p<uint64_t>(), srun.p<uint32_t>(), srun_s.p<uint32_t>(), S,
                            64));  // class in bits 60-63, type 58-59
  release(ctx, skey.b);
  release(ctx, srun.b);
  TmpBuf msz(ctx), nslot(ctx), finv(ctx), moff(ctx), voff(ctx);
  CHECK_RC(ensure(ctx, msz.b, (uint64_t)S * 8));
  CHECK_RC(ensure(ctx, nslot.b, (uint64_t)S * 8));
  CHECK_RC(ensure(ctx, finv.b, (uint64_t)R * 4));
  CHECK_RC(ensure(ctx, moff.b, (uint64_t)S * 8));
  CHECK_RC(ensure(ctx, voff.b, (uint64_t)S * 8));
  HIPC(hipMemsetAsync(finv.b.p, 0xFF, (uint64_t)R * 4, st));
  uint64_t* bnd = reinterpret_cast<uint64_t*>(ctx->scal.p) + 32;  // bytes 256..479 of the status block
  k_seg_sizes<<<div_up(S, 256), 256, 0, st>>>(srun_s.p<uint32_t>(), rstart.p<uint32_t>(), useinc.p<uint32_t>(), S,
                                             msz.p<uint64_t>(), nslot.p<uint64_t>(), finv.p<uint32_t>());
  k_seg_bounds<<<div_up(S, 256), 256, 0, st>>>(skey_s.p<uint64_t>(), S, bnd);
  CHECK_RC(excl_scan(ctx, msz.p<uint64_t>(), moff.p<uint64_t>(), S));
  CHECK_RC(excl_scan(ctx, nslot.p<uint64_t>(), voff.p<uint64_t>(), S));
  uint64_t last[4];
  CHECK_RC(copy_back(ctx, &last[0], moff.p<uint64_t>() + (S - 1)));
  CHECK_RC(copy_back(ctx, &last[1], msz.p<uint64_t>() + (S - 1)));
  CHECK_RC(copy_back(ctx, &last[2], voff.p<uint64_t>() + (S - 1)));
  CHECK_RC(copy_back(ctx, &last[3], nslot.p<uint64_t>() + (S - 1)));
  const uint64_t n_mem = last[0] + last[1], n_slots = last[2] + last[3];
  if (n_mem >= (1ull << 32)) return fail(ctx, SHCI_ERR_UNSUPPORTED, "segment memberships >= 2^32");
  uint64_t hb[kSegRanges + 1];
  CHECK_RC(copy_back(ctx, hb, bnd, kSegRanges + 1));
  CHECK_RC(ensure(ctx, ctx->seg_desc, (uint64_t)S * sizeof(SegDesc)));
  CHECK_RC(ensure(ctx, ctx->seg_type, (uint64_t)S));
  CHECK_RC(ensure(ctx, ctx->seg_core, (uint64_t)S * 8));
  CHECK_RC(ensure(ctx, ctx->seg_mem, n_mem * 4));
  SegDesc* desc = P<SegDesc>(ctx->seg_desc);
  k_seg_desc<<<div_up(S, 256), 256, 0, st>>>(K, srun_s.p<uint32_t>(), rstart.p<uint32_t>(), useinc.p<uint32_t>(),
                                            vals_s.p<uint32_t>(),
                                            moff.p<uint64_t>(), voff.p<uint64_t>(), S, desc,
                                            P<uint8_t>(ctx->seg_type), P<uint64_t>(ctx->seg_core));
  k_seg_members<<<div_up(M, 256), 256, 0, st>>>(K, vals_s.p<uint32_t>(), incl, rstart.p<uint32_t>(), use,
                                               useinc.p<uint32_t>(), finv.p<uint32_t>(), desc, M,
                                               P<uint32_t>(ctx->seg_mem));
  HIPC(hipGetLastError());
  ctx->n_seg = S;
  ctx->n_mem = n_mem;
  ctx->n_slots = n_slots;
  for (int r = 0; r <= kSegRanges; r++) ctx->seg_cs[r] = hb[r];
  for (int r = 0; r < 9; r++) {  // one-wavefront classes: largest segment first in each range
    if (hb[r + 1] == hb[r]) continue;
    SegDesc d0;
    CHECK_RC(copy_back(ctx, &d0, desc + ctx->seg_cs[r]));
    ctx->seg_maxm[r] = d0.m;
  }
  // Tile class runs, det tables and the memberships' table indices.
  const uint64_t T = n_tiles, f0 = ctx->seg_cs[9];
  CHECK_RC(ensure(ctx, ctx->tile_rb, T * 6 * 4));
  CHECK_RC(ensure(ctx, ctx->tile_re, T * 6 * 4));
  CHECK_RC(ensure(ctx, ctx->tile_u0, T * 4));
  CHECK_RC(ensure(ctx, ctx->tile_nu, T * 4));
  HIPC(hipMemsetAsync(ctx->tile_rb.p, 0, T * 6 * 4, st));
  HIPC(hipMemsetAsync(ctx->tile_re.p, 0, T * 6 * 4, st));
  if (S > f0)
    k_tile_runs<<<div_up(S - f0, 256), 256, 0, st>>>(skey_s.p<uint64_t>(), f0, S, P<uint32_t>(ctx->tile_rb),
                                                     P<uint32_t>(ctx->tile_re));
  release(ctx, skey_s.b);
  {
    TmpBuf tcnt(ctx);
    CHECK_RC(ensure(ctx, tcnt.b, T * 4));
    k_tile_count<<<div_up(T, 256), 256, 0, st>>>(P<uint32_t>(ctx->tile_rb), P<uint32_t>(ctx->tile_re), desc, S,
                                                 n_mem, T, tcnt.p<uint32_t>());
    CHECK_RC(excl_scan(ctx, tcnt.p<uint32_t>(), P<uint32_t>(ctx->tile_u0), T));
    uint32_t lastc = 0, lastu = 0;
    CHECK_RC(copy_back(ctx, &lastc, tcnt.p<uint32_t>() + (T - 1)));
    CHECK_RC(copy_back(ctx, &lastu, P<uint32_t>(ctx->tile_u0) + (T - 1)));
    const uint64_t n_tab = (uint64_t)lastu + lastc;
    CHECK_RC(ensure(ctx, ctx->tile_tab, std::max<uint64_t>(n_tab, 1) * 4));
    CHECK_RC(ensure(ctx, ctx->seg_loc, n_mem * 2));
    HIPC(hipMemsetAsync(bad, 0, 4, st));
    k_tile_check<<<div_up(T, 256), 256, 0, st>>>(tcnt.p<uint32_t>(), T, bad);
    CHECK_RC(copy_back(ctx, &hbad, bad));
    if (hbad) return fail(ctx, SHCI_ERR_STATE, "internal: tile with too many memberships");
  }
  k_tile_tables<<<T, 256, 0, st>>>(P<uint32_t>(ctx->tile_rb), P<uint32_t>(ctx->tile_re), P<uint32_t>(ctx->tile_u0),
                                   desc, S, n_mem, P<uint32_t>(ctx->seg_mem), P<uint8_t>(ctx->seg_type),
                                   P<uint32_t>(ctx->tile_tab), P<uint32_t>(ctx->tile_nu),
                                   P<uint16_t>(ctx->seg_loc));
  HIPC(hipGetLastError());
  // the tile records the product stages
  {
    TmpBuf units(ctx);
    CHECK_RC(ensure(ctx, units.b, T * 4));
    CHECK_RC(ensure(ctx, ctx->tile_roff, T * 4));
    k_tile_rec_size<<<div_up(T, 256), 256, 0, st>>>(P<uint32_t>(ctx->tile_rb), P<uint32_t>(ctx->tile_re), desc, T,
                                                     units.p<uint32_t>());
    CHECK_RC(excl_scan(ctx, units.p<uint32_t>(), P<uint32_t>(ctx->tile_roff), T));
    uint32_t lastr = 0, lastn = 0;
    CHECK_RC(copy_back(ctx, &lastr, P<uint32_t>(ctx->tile_roff) + (T - 1)));
    CHECK_RC(copy_back(ctx, &lastn, units.p<uint32_t>() + (T - 1)));
    CHECK_RC(ensure(ctx, ctx->tile_rec, ((uint64_t)lastr + lastn) * 16));
    k_tile_records<<<T, 256, 0, st>>>(P<uint32_t>(ctx->tile_rb), P<uint32_t>(ctx->tile_re), P<uint32_t>(ctx->tile_u0),
                                      P<uint32_t>(ctx->tile_nu), desc, P<uint16_t>(ctx->seg_loc),
                                      P<uint32_t>(ctx->tile_roff), P<uint4>(ctx->tile_rec));
    HIPC(hipGetLastError());
  }
  release(ctx, ctx->seg_loc);  // copied into the records
  ctx->n_tiles = T;
  return SHCI_OK;
}

SegFillParams seg_fill_params(shci_ctx* ctx, int update) {
  SegFillParams F;
  F.desc = P<SegDesc>(ctx->seg_desc);
  F.mem = P<uint32_t>(ctx->seg_mem);
  F.type = P<uint8_t>(ctx->seg_type);
  F.core = P<uint64_t>(ctx->seg_core);
  F.up = P<uint64_t>(ctx->up_s);
  F.dn = P<uint64_t>(ctx->dn_s);
  F.dnord = P<uint32_t>(ctx->dnord_idx);
  F.I = make_ints(ctx);
  F.v64 = P<double>(ctx->sv64);
  F.v32 = P<float>(ctx->sv32);
  F.kept = P<unsigned long long>(ctx->kept);
  F.ra = ctx->ra;
  F.rb = ctx->rb;
  F.s0 = F.s1 = 0;
  F.keep_zeros = (ctx->build_flags & SHCI_KEEP_ZEROS) ? 1 : 0;
  F.update = update;
  return F;
}

// Value slots of every segment (build) or of the stored pairs (update).
int seg_fill(shci_ctx* ctx, int update) {
  const hipStream_t st = ctx->stream;
  CHECK_RC(ensure(ctx, ctx->scal, 64 * 8));
  // kept-pair counters: kKeptLines lines per range
  CHECK_RC(ensure(ctx, ctx->kept, (size_t)kSegRanges * kKeptLines * 128));
  SegFillParams F = seg_fill_params(ctx, update);
  unsigned long long* const kept0 = F.kept;
  HIPC(hipMemsetAsync(kept0, 0, (size_t)kSegRanges * kKeptLines * 128, st));
  for (int r = 0; r < kSegRanges; r++) {
    const int c = r / 3;
    F.kept = kept0 + (size_t)r * kKeptLines * 16;
    F.s0 = (uint32_t)ctx->seg_cs[r];
    F.s1 = (uint32_t)ctx->seg_cs[r + 1];
    const uint32_t cnt = F.s1 - F.s0;
    if (cnt == 0) continue;
    if (c < 3) {
      if (!update && r % 3 == (int)kSegC)
        k_seg_fill_large_c<<<cnt, kWave, ctx->seg_maxm[r] * 4, st>>>(F);
      else
        k_seg_fill_large<<<cnt, kWave, ctx->seg_maxm[r] * 16, st>>>(F);
      continue;
    }
    const int S = 64 >> (c - 3);
    const uint32_t grid = div_up((uint64_t)cnt * S, 256);
    if (!update && r % 3 == (int)kSegC) {
      switch (S) {
        case 64: k_seg_fill_c<64><<<grid, 256, 0, st>>>(F); break;
        case 32: k_seg_fill_c<32><<<grid, 256, 0, st>>>(F); break;
        case 16: k_seg_fill_c<16><<<grid, 256, 0, st>>>(F); break;
        case 8: k_seg_fill_c<8><<<grid, 256, 0, st>>>(F); break;
        case 4: k_seg_fill_c<4><<<grid, 256, 0, st>>>(F); break;
        default: k_seg_fill_c<2><<<grid, 256, 0, st>>>(F); break;
      }
      continue;
    }
    switch (S) {
      case 64: k_seg_fill_small<64><<<grid, 256, 0, st>>>(F); break;
      case 32: k_seg_fill_small<32><<<grid, 256, 0, st>>>(F); break;
      case 16: k_seg_fill_small<16><<<grid, 256, 0, st>>>(F); break;
      case 8: k_seg_fill_small<8><<<grid, 256, 0, st>>>(F); break;
      case 4: k_seg_fill_small<4><<<grid, 256, 0, st>>>(F); break;
      default: k_seg_fill_small<2><<<grid, 256, 0, st>>>(F); break;
    }
  }
  HIPC(hipGetLastError());
  std::vector<unsigned long long> lines((size_t)kSegRanges * kKeptLines * 16);
  CHECK_RC(copy_back(ctx, lines.data(), kept0, lines.size()));
  unsigned long long kept = 0;
  for (int r = 0; r < kSegRanges; r++) {
    unsigned long long kr = 0;
    for (uint32_t i = 0; i < kKeptLines; i++) kr += lines[((size_t)r * kKeptLines + i) * 16];
    ctx->seg_kept[r] = kr;
    kept += kr;
  }
  if (!update) ctx->n_pairs = kept;
  return SHCI_OK;
}

// y = (H - shift) x on internal-order device vectors, NV interleaved vectors.
// Beta-order copy of x for the A segments (xb[k] = x[dnord[k]]) and a zeroed
// partial product yb; afterwards y[dnord[k]] += yb[k].
template <int NV, typename XT>
__global__ void k_perm_in(const uint32_t* __restrict__ dnord, const XT* __restrict__ x, XT* __restrict__ xb,
                          XT* __restrict__ yb, uint64_t n) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t i = dnord[k];
  for (int v = 0; v < NV; v++) {
    xb[k * NV + v] = x[i * NV + v];
    yb[k * NV + v] = (XT)0;
  }
}

template <int NV, typename XT>
__global__ void k_perm_add(const uint32_t* __restrict__ dnord, const XT* __restrict__ yb, XT* __restrict__ y,
                           uint64_t n) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint64_t i = dnord[k];
  for (int v = 0; v < NV; v++) y[i * NV + v] += yb[k * NV + v];
}

// y = (H - shift) x on internal-order device vectors, NV interleaved vectors.
// A segments (members in beta order) run on xb/yb, C and B segments on x/y.
template <int NV, typename VT, typename XT>
int launch_seg_spmv(shci_ctx* ctx, double shift, const XT* x, XT* y, XT* xb_ext = nullptr, XT* yb_ext = nullptr) {
  const hipStream_t st = ctx->stream;
  const uint64_t n = ctx->n;
  const VT* val = sizeof(VT) == 4 ? reinterpret_cast<const VT*>(ctx->sv32.p) : reinterpret_cast<const VT*>(ctx->sv64.p);
  if (!val && ctx->n_slots) return fail(ctx, SHCI_ERR_STATE, "requested precision not stored");
  k_diag_apply<NV, XT><<<div_up(n, 256), 256, 0, st>>>(P<double>(ctx->diag), shift, x, y, n);
  // A segments (and A tiles) address beta-order copies of x and y.
  bool have_a = false;
  for (int c = 0; c < kSegClasses; c++) have_a |= ctx->seg_cs[3 * c + kSegA + 1] > ctx->seg_cs[3 * c + kSegA];
  XT* xb = nullptr;
  XT* yb = nullptr;
  const uint32_t* dnord = P<uint32_t>(ctx->dnord_idx);
  if (have_a) {
    if (xb_ext && yb_ext) {  // caller-provided (the allocation-free product)
      xb = xb_ext;
      yb = yb_ext;
    } else {
      CHECK_RC(ensure(ctx, ctx->xb, n * NV * sizeof(XT)));
      CHECK_RC(ensure(ctx, ctx->yb, n * NV * sizeof(XT)));
      xb = static_cast<XT*>(ctx->xb.p);
      yb = static_cast<XT*>(ctx->yb.p);
    }
    k_perm_in<NV, XT><<<div_up(n, 256), 256, 0, st>>>(dnord, x, xb, yb, n);
  }
  const SegDesc* desc = P<SegDesc>(ctx->seg_desc);
  const uint32_t* mem = P<uint32_t>(ctx->seg_mem);
  for (int r = 0; r < kSegRanges; r++) {
    const int c = r / 3;
    const uint32_t s0 = (uint32_t)ctx->seg_cs[r], s1 = (uint32_t)ctx->seg_cs[r + 1];
    if (s1 == s0) continue;
    const XT* xr = (r % 3 == (int)kSegA) ? xb : x;
    XT* yr = (r % 3 == (int)kSegA) ? yb : y;
    if constexpr (NV == 4) {
      // block product: one lane per (member, vector)
      const float* vf = reinterpret_cast<const float*>(val);
      if (c <= 4) {  // m > 16: one wavefront per (segment, 64 members)
        const uint32_t mx = c < 3 ? ctx->seg_maxm[r] : (c == 3 ? 64u : 32u);
        const dim3 grid(s1 - s0, div_up(mx, 64));
        k_seg_spmm_wave<<<grid, kWave, (size_t)mx * 4 * sizeof(float), st>>>(desc, s0, s1, mem, vf, xr, yr);
      } else {
        const int S = 64 >> (c - 3);
        const uint32_t grid = div_up((uint64_t)(s1 - s0) * 4 * S, 256);
        switch (S) {
          case 16: k_seg_spmm_small<16><<<grid, 256, 0, st>>>(desc, s0, s1, mem, vf, xr, yr); break;
          case 8: k_seg_spmm_small<8><<<grid, 256, 0, st>>>(desc, s0, s1, mem, vf, xr, yr); break;
          case 4: k_seg_spmm_small<4><<<grid, 256, 0, st>>>(desc, s0, s1, mem, vf, xr, yr); break;
          default: k_seg_spmm_small<2><<<grid, 256, 0, st>>>(desc, s0, s1, mem, vf, xr, yr); break;
        }
      }
    } else if (c < 3) {
      if (c == 2) {
        // 64 < m <= 256: one wavefront per segment, rotating travelling sums
        k_seg_spmv_wave<VT, XT><<<s1 - s0, kWave, 0, st>>>(desc, s0, s1, mem, val, xr, yr);
        continue;
      }
      const size_t lds = (size_t)ctx->seg_maxm[r] * 2 * sizeof(XT);
      if (lds > 65536) return fail(ctx, SHCI_ERR_UNSUPPORTED, "segment too large for the one-wavefront kernel");
      // one wavefront per (segment, 256 members, share of the steps): FP64
      // reads its rows one step at a time, so a long segment is one long chain
      // of loads, cut into shares of about 128 steps; FP32 reads 4 steps per
      // batch and keeps one share (the split measured no gain there)
      const uint32_t nz = sizeof(XT) == 8 ? std::min<uint32_t>(16, std::max<uint32_t>(1, div_up(ctx->seg_maxm[r] / 2, 128)))
                                          : 1u;
      const dim3 grid(s1 - s0, div_up(ctx->seg_maxm[r], kWave * kSegChunks), nz);
      k_seg_spmv_large<kSegChunks, VT, XT><<<grid, kWave, lds, st>>>(desc, s0, s1, mem, val, xr, yr);
    }
  }
  if constexpr (NV == 1) {
    // m <= 64: one workgroup per tile
    if (ctx->n_tiles) {
      if constexpr (std::is_same<VT, float>::value && std::is_same<XT, float>::value)
        k_seg_spmv_tiles_f32<<<ctx->n_tiles, 256, 0, st>>>(P<uint32_t>(ctx->tile_roff), P<uint4>(ctx->tile_rec),
                                                            P<uint32_t>(ctx->tile_tab), val, x, y, xb, yb);
      else
        k_seg_spmv_tiles<VT, XT><<<ctx->n_tiles, 256, 0, st>>>(P<uint32_t>(ctx->tile_roff), P<uint4>(ctx->tile_rec),
                                                               P<uint32_t>(ctx->tile_tab), val, x, y, xb, yb);
    }
  }
  if (have_a) k_perm_add<NV, XT><<<div_up(n, 256), 256, 0, st>>>(dnord, yb, y, n);
  HIPC(hipGetLastError());
  return SHCI_OK;
}

bool has_fp32(const shci_ctx* ctx) { return ctx->segl ? (ctx->sv32.p || !ctx->n_slots) : ctx->v32.p != nullptr; }
bool has_fp64(const shci_ctx* ctx) { return ctx->segl ? (ctx->sv64.p || !ctx->n_slots) : ctx->v64.p != nullptr; }

// Lanes per row for the SpMV: each lane consumes 8 entries per loop trip
// (two 4-entry steps).  SHCI_SPMV_LANES overrides (8, 16, 32 or 64).
int choose_lanes(double mean) {
  if (const char* env = getenv("SHCI_SPMV_LANES")) {
    const int v = atoi(env);
    if (v == 8 || v == 16 || v == 32 || v == 64) return v;
  }
  if (mean >= 384) return 64;
  if (mean >= 160) return 32;
  if (mean >= 64) return 16;
  return 8;
}

RowParams row_params(shci_ctx* ctx) {
  RowParams Pm;
  Pm.n = ctx->n;
  Pm.up = P<uint64_t>(ctx->up_s);
  Pm.dn = P<uint64_t>(ctx->dn_s);
  Pm.up_gid = P<uint32_t>(ctx->up_gid);
  Pm.up_gstart = P<uint32_t>(ctx->up_gstart);
  Pm.dn_gid = P<uint32_t>(ctx->dn_gid);
  Pm.dn_gstart = P<uint32_t>(ctx->dn_gstart);
  Pm.dnord_idx = P<uint32_t>(ctx->dnord_idx);
  Pm.dnord_up = P<uint64_t>(ctx->dnord_up);
  Pm.dm1_vals = P<uint32_t>(ctx->dm1_vals);
  Pm.dm1_rstart = P<uint32_t>(ctx->dm1_rstart);
  Pm.dm1_seg = P<uint32_t>(ctx->dm1_seg);
  Pm.K = ctx->n_up * ctx->n_dn;
  Pm.overflow = reinterpret_cast<unsigned int*>(ctx->flag.p);
  Pm.I = make_ints(ctx);
  Pm.cnt = P<uint32_t>(ctx->cnt);
  Pm.row_ptr = P<int64_t>(ctx->row_ptr);
  Pm.row_end = P<int64_t>(ctx->row_end);
  Pm.col = P<uint32_t>(ctx->col);
  Pm.v64 = P<double>(ctx->v64);
  Pm.v32 = P<float>(ctx->v32);
  Pm.len = P<uint32_t>(ctx->len);
  Pm.keep_zeros = (ctx->build_flags & SHCI_KEEP_ZEROS) ? 1 : 0;
  return Pm;
}

int do_build(shci_ctx* ctx, uint32_t flags) {
  if (!ctx->have_ints) return fail(ctx, SHCI_ERR_STATE, "integrals not set");
  if (!(flags & (SHCI_STORE_FP64 | SHCI_STORE_FP32))) flags |= SHCI_STORE_FP64;
  ctx->build_flags = flags;
  ctx->built = false;
  const uint64_t n = ctx->n;
  const hipStream_t st = ctx->stream;
  HIPC(hipEventRecord(ctx->ev[0], st));
  CHECK_RC(build_groups(ctx));
  HIPC(hipEventRecord(ctx->ev[1], st));
  const bool segl = (flags & SHCI_LAYOUT_SEGMENTS) != 0;
  CHECK_RC(build_segments(ctx, segl));
  HIPC(hipEventRecord(ctx->ev[2], st));
  CHECK_RC(ensure(ctx, ctx->diag, n * 8));
  k_diag<<<div_up(n, 256), 256, 0, st>>>(make_ints(ctx), P<uint64_t>(ctx->up_s), P<uint64_t>(ctx->dn_s),
                                         P<double>(ctx->diag), n);
  HIPC(hipEventRecord(ctx->ev[3], st));
  CHECK_RC(ensure(ctx, ctx->flag, 256));
  HIPC(hipMemsetAsync(ctx->flag.p, 0, 256, st));
  ctx->segl = segl;
  if (segl) {
    // The segment layout replaces the CSR: release it.
    DevBuf* csr[] = {&ctx->row_ptr, &ctx->row_end, &ctx->col, &ctx->v64, &ctx->v32, &ctx->len, &ctx->cnt};
    for (DevBuf* b : csr) release(ctx, *b);
    ctx->ra = std::min(ctx->n_up, 2u);
    ctx->rb = std::min(ctx->n_dn, 2u);
    CHECK_RC(build_seg_layout(ctx));
    HIPC(hipEventRecord(ctx->ev[4], st));
    // drop the store that is not asked for before allocating the other
    if (!(flags & SHCI_STORE_FP64)) release(ctx, ctx->sv64);
    if (!(flags & SHCI_STORE_FP32)) release(ctx, ctx->sv32);
    if (flags & SHCI_STORE_FP64) {
      CHECK_RC(ensure(ctx, ctx->sv64, (ctx->n_slots + kSlotPad) * 8));
      HIPC(hipMemsetAsync(P<double>(ctx->sv64) + ctx->n_slots, 0, kSlotPad * 8, st));
    }
    if (flags & SHCI_STORE_FP32) {
      CHECK_RC(ensure(ctx, ctx->sv32, (ctx->n_slots + kSlotPad) * 4));
      HIPC(hipMemsetAsync(P<float>(ctx->sv32) + ctx->n_slots, 0, kSlotPad * 4, st));
    }
    HIPC(hipEventRecord(ctx->ev[5], st));
    CHECK_RC(seg_fill(ctx, 0));
    HIPC(hipEventRecord(ctx->ev[6], st));
    HIPC(hipEventRecord(ctx->ev[7], st));
    HIPC(hipEventSynchronize(ctx->ev[7]));
    ctx->nnz = 2 * ctx->n_pairs;
    ctx->nnz_alloc = 2 * ctx->n_slots;
    ctx->n_dropped = 0;
    for (int k = 0; k < 7; k++) HIPC(hipEventElapsedTime(&ctx->phase_ms[k], ctx->ev[k], ctx->ev[k + 1]));
    HIPC(hipEventElapsedTime(&ctx->phase_ms[7], ctx->ev[0], ctx->ev[7]));
    ctx->built = true;
    return SHCI_OK;
  }
  {
    DevBuf* seg[] = {&ctx->seg_desc, &ctx->seg_type, &ctx->seg_core, &ctx->seg_mem, &ctx->sv32, &ctx->sv64,
                     &ctx->mA_id, &ctx->mB_id, &ctx->dnpos, &ctx->xb, &ctx->yb, &ctx->tile_rb, &ctx->tile_re,
                     &ctx->tile_u0, &ctx->tile_nu, &ctx->tile_tab, &ctx->seg_loc, &ctx->tile_roff, &ctx->tile_rec};
    for (DevBuf* b : seg) release(ctx, *b);
    ctx->n_seg = ctx->n_mem = ctx->n_slots = ctx->n_pairs = ctx->n_tiles = 0;
  }
  // Pass 1: count candidates per row.
  CHECK_RC(ensure(ctx, ctx->cnt, n * 4));
  CHECK_RC(ensure(ctx, ctx->len, n * 4));
  CHECK_RC(ensure(ctx, ctx->row_ptr, (n + 1) * 8));
  {
    RowParams Pm = row_params(ctx);
    k_rows<false><<<div_up(n, 4), 256, 0, st>>>(Pm);
    HIPC(hipGetLastError());
  }
  HIPC(hipEventRecord(ctx->ev[4], st));
  // Scan -> row pointers; allocate the CSR.
  {
    DevBuf cnt64;
    CHECK_RC(ensure(ctx, cnt64, (n + 1) * 8));
    // widen counts to int64 with a trailing zero, then exclusive-scan
    HIPC(hipMemsetAsync(cnt64.p, 0, (n + 1) * 8, st));
    k_widen<<<div_up(n, 256), 256, 0, st>>>(P<uint32_t>(ctx->cnt), P<int64_t>(cnt64), n);
    int rc = excl_scan(ctx, P<int64_t>(cnt64), P<int64_t>(ctx->row_ptr), n + 1);
    release(ctx, cnt64);
    CHECK_RC(rc);
  }
  int64_t total = 0;
  HIPC(hipMemcpyAsync(&total, P<int64_t>(ctx->row_ptr) + n, 8, hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  // Grow-only CSR storage sized by the candidate count; rows whose |H| cut
  // drops entries keep a short row_end (no compaction, no reallocation).
  const uint64_t cap = std::max<int64_t>(total, 1);
  ctx->nnz_alloc = (uint64_t)total;
  if (!(flags & SHCI_STORE_FP64)) release(ctx, ctx->v64);
  if (!(flags & SHCI_STORE_FP32)) release(ctx, ctx->v32);
  CHECK_RC(ensure(ctx, ctx->row_end, n * 8));
  CHECK_RC(ensure(ctx, ctx->col, cap * 4));
  if (flags & SHCI_STORE_FP64) CHECK_RC(ensure(ctx, ctx->v64, cap * 8));
  if (flags & SHCI_STORE_FP32) CHECK_RC(ensure(ctx, ctx->v32, cap * 4));
  HIPC(hipEventRecord(ctx->ev[5], st));
  // Pass 2: fill.
  {
    RowParams Pm = row_params(ctx);
    k_rows<true><<<div_up(n, 4), 256, 0, st>>>(Pm);
    HIPC(hipGetLastError());
  }
  HIPC(hipEventRecord(ctx->ev[6], st));
  {
    unsigned int ovf = 0;
    HIPC(hipMemcpyAsync(&ovf, ctx->flag.p, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (ovf) return fail(ctx, SHCI_ERR_STATE, "internal error: fill pass found more connections than the count pass");
  }
  // Stored off-diagonal count (sum of row lengths after the |H| cut).
  {
    CHECK_RC(ensure(ctx, ctx->scal, 64 * 8));
    unsigned long long* sum = reinterpret_cast<unsigned long long*>(ctx->scal.p);
    HIPC(hipMemsetAsync(sum, 0, 8, st));
    k_sum_u32<<<std::min<uint32_t>(1024, std::max<uint32_t>(1, div_up(n, 256))), 256, 0, st>>>(P<uint32_t>(ctx->len), n,
                                                                                               sum);
    unsigned long long kept = 0;
    HIPC(hipMemcpyAsync(&kept, sum, 8, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    ctx->n_dropped = (uint64_t)total - kept;
    total = (int64_t)kept;
  }
  HIPC(hipEventRecord(ctx->ev[7], st));
  HIPC(hipEventSynchronize(ctx->ev[7]));
  ctx->nnz = (uint64_t)total;
  ctx->lanes = choose_lanes(n ? (double)ctx->nnz_alloc / (double)n : 0.0);
  for (int k = 0; k < 7; k++) HIPC(hipEventElapsedTime(&ctx->phase_ms[k], ctx->ev[k], ctx->ev[k + 1]));
  HIPC(hipEventElapsedTime(&ctx->phase_ms[7], ctx->ev[0], ctx->ev[7]));
  ctx->built = true;
  return SHCI_OK;
}

template <typename VT, typename XT, typename AT>
int launch_csr_spmv(shci_ctx* ctx, const VT* val, double shift, const XT* x, XT* y);

// SpMV launch on internal-order device vectors.
template <typename VT, typename XT, typename AT>
int launch_spmv(shci_ctx* ctx, const VT* val, double shift, const XT* x, XT* y, XT* xb_ext = nullptr,
                XT* yb_ext = nullptr) {
  if (ctx->segl) {
    CHECK_RC(launch_seg_spmv<1, VT, XT>(ctx, shift, x, y, xb_ext, yb_ext));
  } else {
    CHECK_RC(launch_csr_spmv<VT, XT, AT>(ctx, val, shift, x, y));
  }
  if (ctx->ts) {
    // keep the products in the symmetric subspace (the Davidson's vectors
    // start there; the projection stops rounding from drifting out of it)
    k_ts_sym<XT><<<div_up(ctx->n, 256), 256, 0, ctx->stream>>>(y, P<uint32_t>(ctx->ts_rev), ctx->n);
    HIPC(hipGetLastError());
  }
  return SHCI_OK;
}

template <typename VT, typename XT, typename AT>
int launch_csr_spmv(shci_ctx* ctx, const VT* val, double shift, const XT* x, XT* y) {
  const uint64_t n = ctx->n;
  const int L = ctx->lanes;
  const uint32_t grid = div_up(n, 256 / L);
  const hipStream_t st = ctx->stream;
  const int64_t* rp = P<int64_t>(ctx->row_ptr);
  const int64_t* re = P<int64_t>(ctx->row_end);
  const uint32_t* col = P<uint32_t>(ctx->col);
  const double* d = P<double>(ctx->diag);
  switch (L) {
    case 64: k_spmv<64, VT, XT, AT><<<grid, 256, 0, st>>>(n, rp, re, col, val, d, shift, x, y); break;
    case 32: k_spmv<32, VT, XT, AT><<<grid, 256, 0, st>>>(n, rp, re, col, val, d, shift, x, y); break;
    case 16: k_spmv<16, VT, XT, AT><<<grid, 256, 0, st>>>(n, rp, re, col, val, d, shift, x, y); break;
    default: k_spmv<8, VT, XT, AT><<<grid, 256, 0, st>>>(n, rp, re, col, val, d, shift, x, y); break;
  }
  HIPC(hipGetLastError());
  return SHCI_OK;
}

int launch_spmm4(shci_ctx* ctx, double shift, const float* x, float* y) {
  const uint64_t n = ctx->n;
  const int L = ctx->lanes;
  const uint32_t grid = div_up(n, 256 / L);
  const hipStream_t st = ctx->stream;
  const int64_t* rp = P<int64_t>(ctx->row_ptr);
  const int64_t* re = P<int64_t>(ctx->row_end);
  const uint32_t* col = P<uint32_t>(ctx->col);
  const float* v = P<float>(ctx->v32);
  const double* d = P<double>(ctx->diag);
  switch (L) {
    case 64: k_spmm_f32<64, 4><<<grid, 256, 0, st>>>(n, rp, re, col, v, d, shift, x, y); break;
    case 32: k_spmm_f32<32, 4><<<grid, 256, 0, st>>>(n, rp, re, col, v, d, shift, x, y); break;
    case 16: k_spmm_f32<16, 4><<<grid, 256, 0, st>>>(n, rp, re, col, v, d, shift, x, y); break;
    default: k_spmm_f32<8, 4><<<grid, 256, 0, st>>>(n, rp, re, col, v, d, shift, x, y); break;
  }
  HIPC(hipGetLastError());
  return SHCI_OK;
}

// FP64 results of up to kMaxDots dot products of a with b[0..m) (device vectors).
template <typename T>
int dots(shci_ctx* ctx, const T* a, const T* const* b, int m, double* host_out) {
  CHECK_RC(ensure(ctx, ctx->part, (size_t)kMaxDots * kDotBlocks * 8));
  CHECK_RC(ensure(ctx, ctx->scal, 64 * 8));
  DotArgs B;
  for (int j = 0; j < m; j++) B.b[j] = b[j];
  const uint32_t nblk = std::min<uint32_t>(kDotBlocks, std::max<uint32_t>(1, div_up(ctx->n, 256)));
  k_dots_partial<T><<<nblk, 256, 0, ctx->stream>>>(a, B, m, ctx->n, P<double>(ctx->part));
  k_dots_final<<<1, 256, 0, ctx->stream>>>(P<double>(ctx->part), m, (int)nblk, P<double>(ctx->scal));
  if (host_out) {
    HIPC(hipMemcpyAsync(host_out, ctx->scal.p, m * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
  }
  return SHCI_OK;
}

// Dense symmetric eigen-decomposition (cyclic Jacobi) of the lower triangle of
// an n x n row-major matrix; ascending eigenvalues, eigenvectors as columns.
void jacobi_eigh(int n, const std::vector<double>& a_in, std::vector<double>& w, std::vector<double>& v) {
  std::vector<double> a = a_in;
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) v[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) off += a[(size_t)p * n + q] * a[(size_t)p * n + q];
    if (off < 1e-300) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = a[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double theta = (a[(size_t)q * n + q] - a[(size_t)p * n + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = c * akp - s * akq;
          a[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = c * apk - s * aqk;
          a[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
          v[(size_t)k * n + p] = c * vkp - s * vkq;
          v[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int x, int y) { return a[(size_t)x * n + x] < a[(size_t)y * n + y]; });
  w.resize(n);
  std::vector<double> vs((size_t)n * n);
  for (int j = 0; j < n; j++) {
    w[j] = a[(size_t)ord[j] * n + ord[j]];
    for (int i = 0; i < n; i++) vs[(size_t)i * n + j] = v[(size_t)i * n + ord[j]];
  }
  v.swap(vs);
}

// One reference Davidson call (davidson.cc:8-196) on internal-order device
// vectors of type T.  For T = float the matrix is applied shifted by `shift`
// (FP32 store, FP32 row accumulation, FP64 diagonal) and eigenvalues are
// reported in the shifted frame.
template <typename T>
struct Davidson {
  shci_ctx* ctx = nullptr;
  uint32_t ns;    // n_states
  size_t K;       // ns * n_store
  uint64_t n;
  double shift;
  std::vector<T*> v, hv, w, hw;
  T* tmp;
  std::vector<double> lowest, lowest_prev;
  bool converged = false;
  std::chrono::steady_clock::time_point t0;
  double* traj_e;
  double* traj_t;
  int traj_cap;
  int* n_traj;
  double report_shift = 0.0;

  int spmv(const T* x, T* y);
  // H v[k] -> hv[k] for the listed slots (one block product where possible).
  int products(const std::vector<size_t>& slots);
  DevBuf bx, by;  // interleaved block of the FP32 block product
  ~Davidson() {
    if (ctx) {
      release(ctx, bx);
      release(ctx, by);
    }
  }

  void record(double e) {
    if (traj_e && *n_traj < traj_cap) {
      traj_e[*n_traj] = e;
      if (traj_t) traj_t[*n_traj] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      (*n_traj)++;
    }
  }
  // A Rayleigh-Ritz step's estimate: recorded, and its ground-state Ritz
  // vector captured when capture is on (after the timestamp; the widening
  // copy is queued on the stream).
  void record_ritz(double e) {
    record(e);
    if (ctx->cap_n < ctx->cap_max) {
      double* dst = P<double>(ctx->cap) + (size_t)ctx->cap_n * ((n + 63) & ~(uint64_t)63);
      k_copy_scale<T, double><<<div_up(n, 256), 256, 0, ctx->stream>>>(w[0], dst, 1.0, n);
      ctx->cap_n++;
    }
  }

  // init: ns device vectors (internal order, already normalised/orthogonal).
  int run(double tol) {
    const uint32_t g = div_up(n, 256);
    const hipStream_t st = ctx->stream;
    const size_t n_store = K / ns;
    std::vector<double> hk(K * K, 0.0);
    auto H = [&](size_t i, size_t j) -> double& { return hk[i * K + j]; };
    lowest.assign(ns, 0.0);
    size_t n_converged = 0;
    converged = false;
    {
      std::vector<size_t> first(ns);
      for (uint32_t s = 0; s < ns; s++) first[s] = s;
      CHECK_RC(products(first));
    }
    for (uint32_t s = 0; s < ns; s++) {
      const T* b[1] = {hv[s]};
      double e;
      CHECK_RC(dots<T>(ctx, v[s], b, 1, &e));
      lowest[s] = e;
      H(s, s) = e;
      HIPC(hipMemcpyAsync(w[s], v[s], n * sizeof(T), hipMemcpyDeviceToDevice, st));
      HIPC(hipMemcpyAsync(hw[s], hv[s], n * sizeof(T), hipMemcpyDeviceToDevice, st));
    }
    record_ritz(lowest[0] + report_shift);
    lowest_prev = lowest;
    size_t i_state_precond = 0;
    // Products of a round's corrections wait for its Rayleigh-Ritz step: a
    // correction depends on the last Ritz pairs only, and its product only
    // enters that step (the same arithmetic as multiplying each at once).
    std::vector<size_t> pending;
    for (size_t it = ns; it < n_store * ns * 2; it++) {
      const size_t ic = it % K;
      if (it >= K) {
        if (ic < ns - 1) continue;
        if (ic == ns - 1) {
          pending.clear();  // the restart drops this round's corrections so far
          for (uint32_t s = 0; s < ns; s++) {
            HIPC(hipMemcpyAsync(v[s], w[s], n * sizeof(T), hipMemcpyDeviceToDevice, st));
            HIPC(hipMemcpyAsync(hv[s], hw[s], n * sizeof(T), hipMemcpyDeviceToDevice, st));
          }
          for (uint32_t s = 0; s < ns; s++) {
            std::vector<const T*> b;
            for (uint32_t k = s; k < ns; k++) b.push_back(hv[k]);
            std::vector<double> d(b.size());
            CHECK_RC(dots<T>(ctx, v[s], b.data(), (int)b.size(), d.data()));
            lowest[s] = d[0];
            H(s, s) = d[0];
            for (uint32_t k = s + 1; k < ns; k++) H(s, k) = H(k, s) = d[k - s];
          }
          continue;
        }
      }
      const double lam = lowest[i_state_precond];
      k_precond<T><<<g, 256, 0, st>>>(v[ic], hw[i_state_precond], w[i_state_precond], P<double>(ctx->diag), lam,
                                      shift, n);
      // time_sym: the correction back onto the spin-flip-even vectors.  The
      // diagonal of a det and of its reverse can differ in the last bit, and
      // an odd component would meet the projected products' null space:
      // eigenvalue 0, below every root of an all-positive spectrum.
      if (ctx->ts) k_ts_sym<T><<<g, 256, 0, st>>>(v[ic], P<uint32_t>(ctx->ts_rev), n);
      // Modified Gram-Schmidt with device-resident scalars (davidson.cc:116-123).
      for (size_t i = 0; i < ic; i++) {
        const T* b[1] = {v[i]};
        CHECK_RC(dots<T>(ctx, v[ic], b, 1, nullptr));
        k_axpy_dev<T><<<g, 256, 0, st>>>(v[ic], v[i], P<double>(ctx->scal), n);
      }
      double nrm2;
      {
        const T* b[1] = {v[ic]};
        CHECK_RC(dots<T>(ctx, v[ic], b, 1, &nrm2));
      }
      const double norm = std::sqrt(nrm2);
      if (norm < 1e-12) {
        converged = true;
        break;
      }
      k_scale<T><<<g, 256, 0, st>>>(v[ic], 1.0 / norm, n);
      pending.push_back(ic);
      if (i_state_precond + 1 == ns) {
        CHECK_RC(products(pending));
        for (size_t pc : pending) {
          std::vector<const T*> b(v.begin(), v.begin() + pc + 1);
          std::vector<double> d(pc + 1);
          CHECK_RC(dots<T>(ctx, hv[pc], b.data(), (int)pc + 1, d.data()));
          for (size_t i = 0; i <= pc; i++) H(pc, i) = d[i];
        }
        pending.clear();
        const int ks = (int)ic + 1;
        std::vector<double> sub((size_t)ks * ks);
        for (int i = 0; i < ks; i++)
          for (int j = 0; j < ks; j++) sub[(size_t)i * ks + j] = i >= j ? H(i, j) : H(j, i);
        std::vector<double> ev, evec;
        jacobi_eigh(ks, sub, ev, evec);
        for (uint32_t s = 0; s < ns; s++) {
          lowest[s] = ev[s];
          const double factor = evec[s] < 0 ? -1.0 : 1.0;
          RitzArgs R;
          for (int i = 0; i < ks; i++) {
            R.v[i] = v[i];
            R.hv[i] = hv[i];
            R.c[i] = evec[(size_t)i * ks + s] * factor;
          }
          k_ritz<T, T><<<g, 256, 0, st>>>(R, ks, n, tmp, hw[s]);
          HIPC(hipMemcpyAsync(w[s], tmp, n * sizeof(T), hipMemcpyDeviceToDevice, st));
        }
        HIPC(hipStreamSynchronize(st));
        record_ritz(lowest[0] + report_shift);
        for (uint32_t s = (uint32_t)n_converged; s < ns; s++) {
          if (std::fabs(lowest[s] - lowest_prev[s]) > tol) break;
          n_converged++;
        }
        if (n_converged == ns) converged = true;
        if (!converged) lowest_prev = lowest;
        if (converged) break;
      }
      i_state_precond++;
      if (i_state_precond >= ns) i_state_precond = n_converged;
    }
    return SHCI_OK;
  }
};

template <>
int Davidson<double>::spmv(const double* x, double* y) {
  return launch_spmv<double, double, double>(ctx, P<double>(ctx->v64), 0.0, x, y);
}
template <>
int Davidson<float>::spmv(const float* x, float* y) {
  return launch_spmv<float, float, float>(ctx, P<float>(ctx->v32), shift, x, y);
}

template <>
int Davidson<double>::products(const std::vector<size_t>& slots) {
  for (size_t k : slots) CHECK_RC(spmv(v[k], hv[k]));
  return SHCI_OK;
}
template <>
int Davidson<float>::products(const std::vector<size_t>& slots) {
  // The block product (values read once for up to 4 vectors) when it applies:
  // two or more vectors, segment x staging in LDS.  Under time_sym each
  // product is then projected, as launch_spmv projects a single one.
  const uint32_t seg_big = *std::max_element(ctx->seg_maxm, ctx->seg_maxm + 9);
  const bool block_ok = ctx->segl ? (size_t)seg_big * 4 * sizeof(float) <= 65536 : true;
  if (slots.size() < 2 || !block_ok) {
    for (size_t k : slots) CHECK_RC(spmv(v[k], hv[k]));
    return SHCI_OK;
  }
  const hipStream_t st = ctx->stream;
  CHECK_RC(ensure(ctx, bx, 4 * n * sizeof(float)));
  CHECK_RC(ensure(ctx, by, 4 * n * sizeof(float)));
  for (size_t s0 = 0; s0 < slots.size(); s0 += 4) {
    const uint32_t k = (uint32_t)std::min<size_t>(4, slots.size() - s0);
    Vec4 src{}, dst{};
    for (uint32_t j = 0; j < k; j++) {
      src.p[j] = v[slots[s0 + j]];
      dst.p[j] = hv[slots[s0 + j]];
    }
    k_pack4<<<div_up(n, 256), 256, 0, st>>>(src, k, n, P<float>(bx));
    CHECK_RC(ctx->segl ? launch_seg_spmv<4, float, float>(ctx, shift, P<float>(bx), P<float>(by))
                       : launch_spmm4(ctx, shift, P<float>(bx), P<float>(by)));
    k_unpack4<<<div_up(n, 256), 256, 0, st>>>(P<float>(by), k, n, dst);
    if (ctx->ts)
      for (uint32_t j = 0; j < k; j++)
        k_ts_sym<float><<<div_up(n, 256), 256, 0, st>>>(dst.p[j], P<uint32_t>(ctx->ts_rev), n);
    HIPC(hipGetLastError());
  }
  return SHCI_OK;
}

// Normalise + orthogonalise the initial vectors (davidson.cc:44-59) into v[0..ns).
template <typename T>
int init_vectors(shci_ctx* ctx, Davidson<T>& D, const double* init_dev_internal) {
  const uint64_t n = ctx->n;
  const uint32_t g = div_up(n, 256);
  for (uint32_t s = 0; s < D.ns; s++) {
    const double* x = init_dev_internal + (size_t)s * n;
    double nrm2;
    const double* b[1] = {x};
    CHECK_RC(dots<double>(ctx, x, b, 1, &nrm2));
    k_copy_scale<double, T><<<g, 256, 0, ctx->stream>>>(x, D.v[s], 1.0 / std::sqrt(nrm2), n);
    if (s > 0) {
      for (uint32_t k = 0; k < s; k++) {
        const T* bb[1] = {D.v[k]};
        CHECK_RC(dots<T>(ctx, D.v[s], bb, 1, nullptr));
        k_axpy_dev<T><<<g, 256, 0, ctx->stream>>>(D.v[s], D.v[k], P<double>(ctx->scal), n);
      }
      const T* bb[1] = {D.v[s]};
      CHECK_RC(dots<T>(ctx, D.v[s], bb, 1, &nrm2));
      k_scale<T><<<g, 256, 0, ctx->stream>>>(D.v[s], 1.0 / std::sqrt(nrm2), n);
    }
  }
  return SHCI_OK;
}

// ---------------------------------------------------------------------------
// Heat-bath selection: one sweep of the variational loop
// (src/solver/solver.h:435-495) with ChemSystem::find_connected_dets
// (src/chem/chem_system.cc:289-402; second rejection and active space off,
// all orbitals in one irrep).
//
// Queues (chem_system.cc:115-287): singles list p holds (S(p, r), r) with
// S > 0, S descending; double list L of an occupied spin-orbital pair holds
// (|H|, r, s) with |H| > 0, descending.  A det with eps_min = eps_var / max
// state |c| below its eps_tried_prev walks, for every occupied orbital, the
// singles list while S >= eps_min and, for every occupied pair, the double
// list over eps_min <= |H| < eps_tried_prev.  A candidate is kept when it is
// not in the variational space and, for a single, when |H_ij| >= eps_min.
// The kept candidates are radix-sorted on the packed (up, dn) key and made
// unique: the new dets come out (up, dn) ascending.
//
// Double lists: same-spin pair p < q < n is list q(q-1)/2 + p with r < s,
// r, s not in {p, q}, |(pr|qs) - (ps|qr)|; opposite-spin pair (a up, b down),
// a <= b, is list n(n-1)/2 + b(b+1)/2 + a with r != a, s != b, |(ar|bs)|
// (the up electron goes a -> r, the down one b -> s).  An up/down pair with
// the up orbital above the down one uses the swapped list (chem_system.cc:
// 357-375).
// ---------------------------------------------------------------------------
struct DetKey {
  uint64_t hi, lo;
};

struct DetKeyDecomposer {
  __host__ __device__ ::rocprim::tuple<uint64_t&, uint64_t&> operator()(DetKey& k) const {
    return ::rocprim::tuple<uint64_t&, uint64_t&>(k.hi, k.lo);
  }
};

// (up, dn) as the 2n-bit number up * 2^n + dn, so that the key order is the
// (up, dn) order.
__device__ __forceinline__ DetKey pack_det(uint64_t u, uint64_t d, uint32_t n) {
  return n < 64 ? DetKey{u >> (64 - n), (u << n) | d} : DetKey{u, d};
}

__device__ __forceinline__ void unpack_det(const DetKey& k, uint32_t n, uint64_t& u, uint64_t& d) {
  if (n < 64) {
    d = k.lo & ((1ull << n) - 1ull);
    u = (k.hi << (64 - n)) | (k.lo >> n);
  } else {
    u = k.hi;
    d = k.lo;
  }
}

__device__ __forceinline__ uint64_t det_hash(uint64_t u, uint64_t d) {
  uint64_t x = u * 0x9E3779B97F4A7C15ull ^ (d + 0x632BE59BD9B4E019ull + (u >> 17));
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  x ^= x >> 31;
  return x;
}

// The dets as an open-addressing set (at most a quarter full) on det_hash:
// a slot holds the high 32 bits of the det's hash above its index, ~0 when
// free, so that a probe loads the det's strings only when the hashes agree.
constexpr uint64_t kDetNone = ~0ull;

__global__ void k_det_index_insert(const uint64_t* __restrict__ up, const uint64_t* __restrict__ dn, uint64_t n,
                                   unsigned long long* __restrict__ tab, uint64_t mask) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t hv = det_hash(up[i], dn[i]);
  const unsigned long long e = (hv & 0xFFFFFFFF00000000ull) | i;
  uint64_t h = hv & mask;
  while (atomicCAS(&tab[h], (unsigned long long)kDetNone, e) != (unsigned long long)kDetNone) h = (h + 1) & mask;
}

__device__ __forceinline__ uint32_t det_index_find(const unsigned long long* __restrict__ tab, uint64_t mask,
                                                   const uint64_t* __restrict__ up, const uint64_t* __restrict__ dn,
                                                   uint64_t u, uint64_t d) {
  const uint64_t hv = det_hash(u, d);
  uint64_t h = hv & mask;
  while (true) {
    const uint64_t e = tab[h];
    if (e == kDetNone) return 0xFFFFFFFFu;
    const uint32_t k = (uint32_t)e;
    if ((e >> 32) == (hv >> 32) && up[k] == u && dn[k] == d) return k;
    h = (h + 1) & mask;
  }
}

// 1-RDM (rdm.cc:12-131, get_1rdm) of a CSR build that drops cut pairs: for
// each det i and each electron p of either spin, every orbital r of its spin
// (r = p, or r empty in det i) whose det i with p -> r is in the list, as
// det j, adds diff(i, j)'s permutation factor times sum_s c_s[i] c_s[j] to
// element (p, r).  The singles are looked up in the det list, as the
// reference does, whether or not the Hamiltonian stores the pair (|H_ij|
// below 1e-12 is cut there).  The reference's orbital-symmetry filter holds
// by itself for a selected list: its dets share one spatial symmetry, which
// a single p -> r keeps only when p and r share an irrep.  One thread per
// (det, electron); each workgroup sums into an LDS copy of the 1-RDM
// (n_orb^2 doubles of dynamic LDS) and adds it to the result at the end.
__global__ __launch_bounds__(256) void k_rdm1(RdmParams R, const unsigned long long* __restrict__ tab, uint64_t mask,
                                              uint32_t ne) {
  extern __shared__ __align__(16) unsigned char rdm1_lds[];
  double* acc = reinterpret_cast<double*>(rdm1_lds);
  const int n = R.n_orb;
  const uint32_t nn = (uint32_t)(n * n);
  for (uint32_t k = threadIdx.x; k < nn; k += blockDim.x) acc[k] = 0.0;
  __syncthreads();
  const uint64_t total = R.n * ne;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = t / ne;
    const uint32_t e = (uint32_t)(t - i * ne);
    const uint64_t u = R.up[i], d = R.dn[i];
    const uint32_t nu = (uint32_t)__popcll(u);
    const bool alpha = e < nu;
    const uint64_t sp = alpha ? u : d;
    // the electron's orbital: the (e or e - nu)-th set bit of its string
    uint64_t m = sp;
    for (uint32_t q = alpha ? e : e - nu; q > 0; q--) m &= m - 1;
    const int p = ctz64(m);
    atomicAdd(&acc[p * n + p], cc(R, i, i));
    for (int r = 0; r < n; r++) {
      if ((sp >> r) & 1ull) continue;
      const uint64_t ns = (sp & ~(1ull << p)) | (1ull << r);
      const uint32_t j = alpha ? det_index_find(tab, mask, R.up, R.dn, ns, d) : det_index_find(tab, mask, R.up, R.dn, u, ns);
      if (j == 0xFFFFFFFFu) continue;
      atomicAdd(&acc[p * n + r], pf_of(sp, ns) * cc(R, i, j));
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < nn; k += blockDim.x)
    if (acc[k] != 0.0) atomicAdd(&R.rdm1[k], acc[k]);
}

// S(p, r) of the singles queue (chem_system.cc:237-259): h_pr plus the
// n_e - 1 smallest (largest) of the 2n - 2 one-orbital terms (pr|oo) and
// (pr|oo) - (po|or), summed smallest (largest) first; S = max(|min|, |max|).
__global__ void k_sel_singles_S(Ints I, uint32_t n_elecs, double* __restrict__ S) {
  const uint32_t n = (uint32_t)I.n;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const uint32_t p = t / n, r = t % n;
  if (p == r) {
    S[t] = 0.0;
    return;
  }
  double el[126];
  uint32_t m = 0;
  for (uint32_t o = 0; o < n; o++) {
    const double di = eri4(I, p, r, o, o);
    el[m++] = di;
    if (o != p && o != r) el[m++] = di - eri4(I, p, o, o, r);
  }
  for (uint32_t i = 1; i < m; i++) {
    const double v = el[i];
    int j = (int)i - 1;
    while (j >= 0 && el[j] > v) {
      el[j + 1] = el[j];
      j--;
    }
    el[j + 1] = v;
  }
  double smin = I.h1[p * n + r], smax = smin;
  const uint32_t k = min(n_elecs - 1u, 2 * n - 2);
  for (uint32_t i = 0; i < k; i++) {
    smin += el[i];
    smax += el[2 * n - 3 - i];
  }
  S[t] = fmax(fabs(smin), fabs(smax));
}

// Singles list p: the r with S(p, r) > 0, S descending (one wavefront per p).
__global__ __launch_bounds__(64) void k_sel_singles_sort(const double* __restrict__ S, uint32_t n,
                                                          double* __restrict__ sqS, uint8_t* __restrict__ sqR,
                                                          uint32_t* __restrict__ sqLen, double* __restrict__ sqMax) {
  __shared__ double sv[64];
  const uint32_t p = blockIdx.x, r = threadIdx.x;
  const double v = r < n ? S[p * n + r] : 0.0;
  sv[r] = v;
  __syncthreads();
  const bool live = r < n && v != 0.0;
  if (live) {
    uint32_t rank = 0;
    for (uint32_t q = 0; q < n; q++) {
      const double w = sv[q];
      rank += (w != 0.0 && (w > v || (w == v && q < r))) ? 1u : 0u;
    }
    sqS[p * n + rank] = v;
    sqR[p * n + rank] = (uint8_t)r;
    if (rank == 0) sqMax[p] = v;
  }
  const uint64_t b = __ballot(live);
  if (r == 0) {
    sqLen[p] = (uint32_t)__popcll(b);
    if (b == 0) sqMax[p] = 0.0;
  }
}

// Double list L (one workgroup per list): every (r, s) of the list's pair,
// |H| as the bits of a non-negative double, bitonic-sorted descending in LDS.
__global__ __launch_bounds__(256) void k_sel_hci_lists(Ints I, double* __restrict__ hqH,
                                                        uint16_t* __restrict__ hqRS, uint32_t* __restrict__ hqLen,
                                                        double* __restrict__ hqMax) {
  __shared__ uint64_t key[4096];
  __shared__ uint16_t pay[4096];
  const uint32_t n = (uint32_t)I.n, nn = n * n, ns = n * (n - 1) / 2, L = blockIdx.x, tid = threadIdx.x;
  const bool same = L < ns;
  uint32_t a, b;
  if (same) {
    b = 1;
    while ((b + 1) * b / 2 <= L) b++;
    a = L - b * (b - 1) / 2;
  } else {
    const uint32_t l = L - ns;
    b = 0;
    while ((b + 1) * (b + 2) / 2 <= l) b++;
    a = l - b * (b + 1) / 2;
  }
  uint32_t P = 64;
  while (P < nn) P <<= 1;
  for (uint32_t e = tid; e < P; e += 256) {
    double H = 0.0;
    uint16_t rs = 0;
    if (e < nn) {
      const uint32_t r = e / n, s = e % n;
      rs = (uint16_t)(r | (s << 8));
      if (same) {
        if (r < s && a != r && b != s && a != s && b != r) H = fabs(eri4(I, a, r, b, s) - eri4(I, a, s, b, r));
      } else if (a != r && b != s) {
        H = fabs(eri4(I, a, r, b, s));
      }
    }
    key[e] = (uint64_t)__double_as_longlong(H);
    pay[e] = rs;
  }
  __syncthreads();
  for (uint32_t k = 2; k <= P; k <<= 1) {
    for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
      for (uint32_t i = tid; i < P; i += 256) {
        const uint32_t ixj = i ^ jj;
        if (ixj > i) {
          const uint64_t ka = key[i], kb = key[ixj];
          if ((ka < kb) == ((i & k) == 0)) {
            key[i] = kb;
            key[ixj] = ka;
            const uint16_t pa = pay[i];
            pay[i] = pay[ixj];
            pay[ixj] = pa;
          }
        }
      }
      __syncthreads();
    }
  }
  const uint64_t base = (uint64_t)L * nn;
  for (uint32_t e = tid; e < nn; e += 256) {
    hqH[base + e] = __longlong_as_double((long long)key[e]);
    hqRS[base + e] = pay[e];
    if (key[e] != 0 && (e + 1 == nn || key[e + 1] == 0)) hqLen[L] = e + 1;
  }
  if (tid == 0) {
    hqMax[L] = __longlong_as_double((long long)key[0]);
    if (key[0] == 0) hqLen[L] = 0;
  }
}

// Open-addressing set of the variational dets: slot (up, dn), up = ~0 empty
// (no det has every one of 64 orbitals occupied in one spin: rejected by the
// checks).  A slot is claimed by a CAS on its up word; the dets are distinct,
// so a claimed slot never needs its dn compared during the build.
constexpr uint64_t kSelEmpty = ~0ull;

__global__ void k_sel_hash_insert(const uint64_t* __restrict__ up, const uint64_t* __restrict__ dn, uint64_t n,
                                  uint64_t* __restrict__ tab, uint64_t mask) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t u = up[i];
  uint64_t h = det_hash(u, dn[i]) & mask;
  while (atomicCAS((unsigned long long*)&tab[2 * h], (unsigned long long)kSelEmpty, (unsigned long long)u) !=
         (unsigned long long)kSelEmpty)
    h = (h + 1) & mask;
  tab[2 * h + 1] = dn[i];
}

struct SelParams {
  const uint64_t* __restrict__ up;
  const uint64_t* __restrict__ dn;
  uint64_t n;
  const double* __restrict__ coefs;  // [state * n + det]
  uint32_t n_states;
  double eps_var;
  double* __restrict__ eps_tried;
  const uint64_t* __restrict__ htab;  // the dets' open-addressing set: (up, dn) slots
  uint64_t hmask;
  const double* __restrict__ sqS;
  const uint8_t* __restrict__ sqR;
  const uint32_t* __restrict__ sqLen;
  const double* __restrict__ sqMax;   // first (largest) S of each singles list, 0 if empty
  const double* __restrict__ hqH;
  const uint16_t* __restrict__ hqRS;
  const uint32_t* __restrict__ hqLen;
  const double* __restrict__ hqMax;   // first (largest) |H| of each double list, 0 if empty
  Ints I;
  uint64_t* __restrict__ cnt;        // count pass: candidates per det
  const uint64_t* __restrict__ off;  // fill pass: exclusive offsets
  uint64_t base;                     // fill pass: offset of the chunk's first candidate
  DetKey* __restrict__ out;
  uint32_t* __restrict__ bad;        // fill pass: set when a det's count differs from the count pass
  unsigned long long* __restrict__ total;  // append pass: candidates found (also past cap)
  uint64_t cap;                      // append pass: capacity of out
  double* __restrict__ outv;         // PT fill pass: H_ai c_i of each written candidate
  uint32_t n_batches, batch;         // PT: only candidates whose hash falls in this batch
  uint32_t ts;                       // time_sym rules of the selection (solver.h:447-454)
  uint64_t i0, i1;                   // det range
};

__device__ __forceinline__ bool sel_in_var(const SelParams& S, uint64_t u, uint64_t d) {
  uint64_t h = det_hash(u, d) & S.hmask;
  while (true) {
    const uint64_t tu = S.htab[2 * h], td = S.htab[2 * h + 1];
    if (tu == kSelEmpty) return false;
    if (tu == u && td == d) return true;
    h = (h + 1) & S.hmask;
  }
}

__device__ __forceinline__ uint32_t rdl(uint32_t v, int l) { return __builtin_amdgcn_readlane(v, l); }

// One wavefront (a workgroup of its own) per det.  Lanes find the non-empty
// walks (one per occupied orbital and per occupied pair); the wavefront then
// runs each of them 64 list entries at a time.  Modes:
//   kSelCount:  candidates kept per det (no writes);
//   kSelFill:   writes them at the det's offset from the count pass;
//   kSelAppend: one pass, candidates staged in LDS and appended to out with
//               one global atomic per flush (writes past cap dropped; the
//               total counts them all).
// Fill and append record eps_tried_prev = 1.000000001 eps_min (solver.h:466).
constexpr int kSelCount = 0, kSelFill = 1, kSelAppend = 2;
constexpr uint32_t kSelStage = 256;  // LDS stage of the append pass (flushed above 192)

// Appends the wavefront's staged candidates: one atomic for the block of
// slots, then coalesced copies (slots at or past cap are dropped).
__device__ __forceinline__ void sel_flush(const SelParams& S, const DetKey* stage, uint32_t& scnt) {
  if (!scnt) return;
  unsigned long long b0 = 0;
  if (threadIdx.x == 0) b0 = atomicAdd(S.total, (unsigned long long)scnt);
  const uint64_t b = (uint64_t)__shfl((long long)b0, 0, kWave);
  __syncthreads();  // staged entries visible to every lane
  for (uint32_t k = threadIdx.x; k < scnt; k += kWave)
    if (b + k < S.cap) S.out[b + k] = stage[k];
  __syncthreads();  // stage reads done before it is refilled
  scnt = 0;
}

template <int MODE, bool PT = false>
__global__ __launch_bounds__(64) void k_sel_walk(SelParams S) {
  constexpr bool FILL = MODE != kSelCount;
  __shared__ uint8_t occ[128];
  __shared__ DetKey stage[MODE == kSelAppend ? kSelStage : 1];
  uint32_t scnt = 0;  // staged entries (append)
  const uint32_t lane = threadIdx.x;
  const uint32_t n = (uint32_t)S.I.n, nn = n * n, ns = n * (n - 1) / 2;
  for (uint64_t i = S.i0 + blockIdx.x; i < S.i1; i += gridDim.x) {
    const uint64_t u = S.up[i], d = S.dn[i];
    double mc = S.coefs[i];
    for (uint32_t s = 1; s < S.n_states; s++) {
      const double c = S.coefs[(uint64_t)s * S.n + i];
      if (fabs(c) > fabs(mc)) mc = c;
    }
    double emin = S.eps_var / fabs(mc);
    if (S.ts && u != d) emin *= 1.4142135623730951;  // Util::SQRT2 (solver.h:447)
    const double emax = PT ? __builtin_inf() : S.eps_tried[i];  // PT: no upper cut (solver.h:639-642)
    uint64_t cur = MODE == kSelFill ? S.off[i] - S.base : 0;
    const uint64_t lim = MODE == kSelFill ? S.off[i + 1] - S.base : 0;  // the det's range from the count pass
    if (emin >= emax) {
      if (!FILL && lane == 0) S.cnt[i] = 0;
      continue;
    }
    const uint32_t nu = (uint32_t)__popcll(u), ne = nu + (uint32_t)__popcll(d);
    __syncthreads();  // the previous det's occ reads are done
    const uint64_t below = (1ull << lane) - 1ull;
    if ((u >> lane) & 1ull) occ[__popcll(u & below)] = (uint8_t)lane;
    if ((d >> lane) & 1ull) occ[nu + __popcll(d & below)] = (uint8_t)(lane + n);
    __syncthreads();
    const uint32_t nw = ne + ne * (ne - 1) / 2;
    for (uint32_t w0 = 0; w0 < nw; w0 += kWave) {
      const uint32_t w = w0 + lane;
      uint32_t kind = 0, p = 0, q = 0, ks = 0, len = 0, L = 0;
      if (w < ne) {
        p = occ[w];
        L = p < n ? p : p - n;
        if (S.sqMax[L] >= emin) {
          kind = 1;
          len = S.sqLen[L];
        }
      } else if (w < nw) {
        const uint32_t k = w - ne;
        uint32_t b = (uint32_t)((1.0f + sqrtf(1.0f + 8.0f * (float)k)) * 0.5f);
        while (b > 1 && b * (b - 1) / 2 > k) b--;
        while ((b + 1) * b / 2 <= k) b++;
        p = occ[k - b * (b - 1) / 2];
        q = occ[b];
        if (q < n) {
          L = q * (q - 1) / 2 + p;
        } else if (p >= n) {
          L = (q - n) * (q - n - 1) / 2 + (p - n);
        } else if (p <= q - n) {
          L = ns + (q - n) * (q - n + 1) / 2 + p;
        } else {
          L = ns + p * (p + 1) / 2 + (q - n);
        }
        const double* H = S.hqH + (uint64_t)L * nn;
        if (S.hqMax[L] >= emin) {
          len = S.hqLen[L];
          // entries at or above eps_tried_prev were walked before: skip them
          uint32_t hi = 0;
          if (H[0] >= emax) {
            uint32_t lo = 0;
            hi = len;
            while (hi - lo > 1) {
              const uint32_t mid = (lo + hi) / 2;
              if (H[mid] >= emax) lo = mid; else hi = mid;
            }
          }
          ks = hi;
          if (ks < len && H[ks] >= emin) kind = 2;
        }
      }
      uint64_t todo = __ballot(kind != 0);
      while (todo) {
        const int bl = ctz64(todo);
        todo &= todo - 1;
        const uint32_t bk = rdl(kind, bl), bp = rdl(p, bl), bq = rdl(q, bl), bks = rdl(ks, bl), blen = rdl(len, bl);
        const uint64_t lb = (uint64_t)rdl(L, bl) * (bk == 1 ? n : nn);
        for (uint32_t k0 = bks; k0 < blen; k0 += kWave) {
          const uint32_t k = k0 + lane;
          uint64_t cu = u, cd = d;
          const double hv = k < blen ? (bk == 1 ? S.sqS[lb + k] : S.hqH[lb + k]) : -1.0;
          bool ok = hv >= emin;
          const bool more = __ballot(ok) == ~0ull;  // every lane in range: the list may go on
          if (ok && bk == 1) {
            const uint32_t r = S.sqR[lb + k];
            if (bp < n) {
              ok = !((u >> r) & 1ull);
              cu = (u & ~(1ull << bp)) | (1ull << r);
            } else {
              ok = !((d >> r) & 1ull);
              cd = (d & ~(1ull << (bp - n))) | (1ull << r);
            }
          } else if (ok) {
            const uint32_t rs = S.hqRS[lb + k];
            uint32_t r = rs & 0xFFu, s = rs >> 8;
            if (bq < n) {
            } else if (bp >= n) {
              r += n;
              s += n;
            } else if (bp <= bq - n) {
              s += n;
            } else {
              const uint32_t t = r;
              r = s;
              s = t + n;
            }
            const bool occ_r = r < n ? ((u >> r) & 1ull) : ((d >> (r - n)) & 1ull);
            const bool occ_s = s < n ? ((u >> s) & 1ull) : ((d >> (s - n)) & 1ull);
            ok = !occ_r && !occ_s;
            if (ok) {
              if (bp < n) cu &= ~(1ull << bp); else cd &= ~(1ull << (bp - n));
              if (bq < n) cu &= ~(1ull << bq); else cd &= ~(1ull << (bq - n));
              if (r < n) cu |= 1ull << r; else cd |= 1ull << (r - n);
              if (s < n) cu |= 1ull << s; else cd |= 1ull << (s - n);
            }
          }
          // time_sym: the candidate is tested and kept in its canonical form
          // (up <= dn, solver.h:452-455); the single's element uses it as found
          const uint64_t fu = cu, fd = cd;
          if (S.ts && cu > cd) {
            cu = fd;
            cd = fu;
          }
          if (ok) ok = !sel_in_var(S, cu, cd);
          double hc = 0.0;
          if (!PT) {
            if (ok && bk == 1) {
              const double h = bp < n ? single_elem(S.I, u, fu, d) : single_elem(S.I, d, fd, u);
              ok = fabs(h) >= emin;
            }
          } else if (ok) {
            // PT: H_ai c_i with the build's elements (det i on the left), kept
            // when |H_ai c_i| >= eps_pt_dtm (solver.h:710-717)
            double h;
            if (bk == 1) h = bp < n ? single_elem(S.I, u, fu, d) : single_elem(S.I, d, fd, u);
            else if (bq < n) h = double_same(S.I, u, fu);
            else if (bp >= n) h = double_same(S.I, d, fd);
            else h = double_mixed(S.I, u, fu, d, fd);
            hc = h * mc;
            ok = fabs(hc) >= S.eps_var;
            if (ok && S.n_batches > 1) ok = (uint32_t)((det_hash(cu, cd) >> 20) % S.n_batches) == S.batch;
          }
          const uint64_t em = __ballot(ok);
          if (MODE == kSelFill && ok && cur + mbcnt(em) < lim) {
            S.out[cur + mbcnt(em)] = pack_det(cu, cd, n);
            if (PT) S.outv[cur + mbcnt(em)] = hc;
          }
          if (MODE == kSelAppend && em) {
            if (ok) stage[scnt + mbcnt(em)] = pack_det(cu, cd, n);
            scnt += (uint32_t)__popcll(em);
            if (scnt > kSelStage - kWave) sel_flush(S, stage, scnt);
          }
          cur += (uint64_t)__popcll(em);
          if (!more) break;
        }
      }
    }
    if (lane == 0) {
      if (FILL && !PT) S.eps_tried[i] = 1.000000001 * emin;
      if (MODE == kSelFill && cur != lim) atomicOr(S.bad, 1u);
      if (MODE == kSelCount) S.cnt[i] = cur;
    }
  }
  if (MODE == kSelAppend) sel_flush(S, stage, scnt);
}

__global__ void k_sel_first_flags(const DetKey* __restrict__ k, uint64_t m, uint64_t* __restrict__ flag) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  flag[i] = (i == 0 || k[i].hi != k[i - 1].hi || k[i].lo != k[i - 1].lo) ? 1ull : 0ull;
}

__global__ void k_sel_compact(const DetKey* __restrict__ k, const uint64_t* __restrict__ flag,
                              const uint64_t* __restrict__ pos, uint64_t m, DetKey* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && flag[i]) out[pos[i]] = k[i];
}

__global__ void k_sel_unpack(const DetKey* __restrict__ k, uint64_t m, uint32_t n, uint64_t* __restrict__ up,
                             uint64_t* __restrict__ dn) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) unpack_det(k[i], n, up[i], dn[i]);
}

// Builds the queues for the context's integrals and n_elecs electrons.
int sel_queues(shci_ctx* ctx, uint32_t n_elecs) {
  if (ctx->sel_ne == n_elecs) return SHCI_OK;
  const uint32_t n = ctx->n_orb, nn = n * n;
  const Ints I = make_ints(ctx);
  const hipStream_t st = ctx->stream;
  DevBuf Sd;
  CHECK_RC(ensure(ctx, Sd, (size_t)nn * 8));
  CHECK_RC(ensure(ctx, ctx->sq_S, (size_t)nn * 8));
  CHECK_RC(ensure(ctx, ctx->sq_r, (size_t)nn));
  CHECK_RC(ensure(ctx, ctx->sq_len, (size_t)n * 4));
  CHECK_RC(ensure(ctx, ctx->hq_H, (size_t)nn * nn * 8));
  CHECK_RC(ensure(ctx, ctx->hq_rs, (size_t)nn * nn * 2));
  CHECK_RC(ensure(ctx, ctx->hq_len, (size_t)nn * 4));
  CHECK_RC(ensure(ctx, ctx->sq_max, (size_t)n * 8));
  CHECK_RC(ensure(ctx, ctx->hq_max, (size_t)nn * 8));
  k_sel_singles_S<<<div_up(nn, 64), 64, 0, st>>>(I, n_elecs, P<double>(Sd));
  k_sel_singles_sort<<<n, 64, 0, st>>>(P<double>(Sd), n, P<double>(ctx->sq_S), P<uint8_t>(ctx->sq_r),
                                       P<uint32_t>(ctx->sq_len), P<double>(ctx->sq_max));
  k_sel_hci_lists<<<nn, 256, 0, st>>>(I, P<double>(ctx->hq_H), P<uint16_t>(ctx->hq_rs), P<uint32_t>(ctx->hq_len),
                                      P<double>(ctx->hq_max));
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(st));
  release(ctx, Sd);
  ctx->sel_ne = n_elecs;
  return SHCI_OK;
}

int sort_det_keys(shci_ctx* ctx, const DetKey* kin, DetKey* kout, uint64_t m, int end_bit) {
  size_t tb = 0;
  HIPC(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, kin, kout, m, DetKeyDecomposer{}, 0, end_bit, ctx->stream));
  CHECK_RC(ensure(ctx, ctx->tmp, tb));
  HIPC(hipcub::DeviceRadixSort::SortKeys(ctx->tmp.p, tb, kin, kout, m, DetKeyDecomposer{}, 0, end_bit,
                                         ctx->stream));
  return SHCI_OK;
}

// Sorts keys[0..m) (into srt) and writes the distinct ones to out; returns
// their number in *mu.
int unique_det_keys(shci_ctx* ctx, const DetKey* keys, DetKey* srt, uint64_t m, DetKey* out, uint64_t* mu) {
  *mu = 0;
  if (m == 0) return SHCI_OK;
  CHECK_RC(sort_det_keys(ctx, keys, srt, m, 2 * (int)ctx->n_orb));
  DevBuf fl, pos;
  int rc = SHCI_OK;
  if ((rc = ensure(ctx, fl, (m + 1) * 8)) || (rc = ensure(ctx, pos, (m + 1) * 8))) {
    release(ctx, fl);
    release(ctx, pos);
    return rc;
  }
  const hipStream_t st = ctx->stream;
  k_sel_first_flags<<<div_up(m, 256), 256, 0, st>>>(srt, m, P<uint64_t>(fl));
  if (hipMemsetAsync(P<uint64_t>(fl) + m, 0, 8, st) != hipSuccess) rc = fail(ctx, SHCI_ERR_HIP, "memset");
  if (rc == SHCI_OK) rc = excl_scan(ctx, P<uint64_t>(fl), P<uint64_t>(pos), m + 1);
  if (rc == SHCI_OK) {
    k_sel_compact<<<div_up(m, 256), 256, 0, st>>>(srt, P<uint64_t>(fl), P<uint64_t>(pos), m, out);
    if (hipMemcpyAsync(mu, P<uint64_t>(pos) + m, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = fail(ctx, SHCI_ERR_HIP, "unique count");
  }
  release(ctx, fl);
  release(ctx, pos);
  return rc;
}

__global__ void k_sel_check(const uint64_t* __restrict__ up, const uint64_t* __restrict__ dn, uint64_t n,
                            uint64_t omask, uint32_t ne, uint32_t* __restrict__ bad) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t u = up[i], d = dn[i];
  if ((u & ~omask) || (d & ~omask) || u == kSelEmpty || (uint32_t)(__popcll(u) + __popcll(d)) != ne)
    atomicOr(bad, 1u);
}

// Uploads the dets, coefficients and (if given) eps_tried_prev, checks the
// dets, builds the queues and the set of the dets; fills S for the walks.
int sel_prepare(shci_ctx* ctx, const uint64_t* up, const uint64_t* dn, uint64_t n, const double* coefs,
                uint32_t n_states, const double* eps_tried, DevBuf* bufs, SelParams& S) {
  DevBuf &dU = bufs[0], &dD = bufs[1], &dC = bufs[2], &dE = bufs[3], &tab = bufs[4];
  const hipStream_t st = ctx->stream;
  const uint32_t n_orb = ctx->n_orb;
  const uint64_t omask = n_orb < 64 ? (1ull << n_orb) - 1ull : ~0ull;
  const uint32_t ne = (uint32_t)(__builtin_popcountll(up[0]) + __builtin_popcountll(dn[0]));
  CHECK_RC(ensure(ctx, dU, n * 8));
  CHECK_RC(ensure(ctx, dD, n * 8));
  CHECK_RC(ensure(ctx, dC, (size_t)n_states * n * 8));
  if (eps_tried) CHECK_RC(ensure(ctx, dE, n * 8));
  CHECK_RC(ensure(ctx, ctx->scal, 64));
  HIPC(hipEventRecord(ctx->ev[0], st));
  HIPC(hipMemcpyAsync(dU.p, up, n * 8, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(dD.p, dn, n * 8, hipMemcpyHostToDevice, st));
  HIPC(hipMemcpyAsync(dC.p, coefs, (size_t)n_states * n * 8, hipMemcpyHostToDevice, st));
  if (eps_tried) HIPC(hipMemcpyAsync(dE.p, eps_tried, n * 8, hipMemcpyHostToDevice, st));
  HIPC(hipMemsetAsync(ctx->scal.p, 0, 8, st));
  k_sel_check<<<div_up(n, 256), 256, 0, st>>>(P<uint64_t>(dU), P<uint64_t>(dD), n, omask, ne, P<uint32_t>(ctx->scal));
  uint32_t bad = 0;
  HIPC(hipMemcpyAsync(&bad, ctx->scal.p, 4, hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  if (bad) return fail(ctx, SHCI_ERR_ARG, "selection: dets need orbitals < n_orb and equal electron counts");
  CHECK_RC(sel_queues(ctx, ne));
  // the variational space as an open-addressing set of (up, dn) slots (load <= 1/2)
  uint64_t H = 1024;
  while (H < 2 * n) H <<= 1;
  CHECK_RC(ensure(ctx, tab, H * 16));
  HIPC(hipMemsetAsync(tab.p, 0xFF, H * 16, st));
  k_sel_hash_insert<<<div_up(n, 256), 256, 0, st>>>(P<uint64_t>(dU), P<uint64_t>(dD), n, P<uint64_t>(tab), H - 1);
  S = SelParams{};
  S.up = P<uint64_t>(dU);
  S.dn = P<uint64_t>(dD);
  S.n = n;
  S.coefs = P<double>(dC);
  S.n_states = n_states;
  S.eps_tried = eps_tried ? P<double>(dE) : nullptr;
  S.htab = P<uint64_t>(tab);
  S.hmask = H - 1;
  S.sqS = P<double>(ctx->sq_S);
  S.sqR = P<uint8_t>(ctx->sq_r);
  S.sqLen = P<uint32_t>(ctx->sq_len);
  S.sqMax = P<double>(ctx->sq_max);
  S.hqH = P<double>(ctx->hq_H);
  S.hqRS = P<uint16_t>(ctx->hq_rs);
  S.hqLen = P<uint32_t>(ctx->hq_len);
  S.hqMax = P<double>(ctx->hq_max);
  S.I = make_ints(ctx);
  S.bad = P<uint32_t>(ctx->scal) + 1;
  S.total = reinterpret_cast<unsigned long long*>(P<uint64_t>(ctx->scal) + 1);
  S.n_batches = 1;
  S.batch = 0;
  S.ts = 0;
  S.i0 = 0;
  S.i1 = n;
  return SHCI_OK;
}

int select_sweep(shci_ctx* ctx, const uint64_t* up, const uint64_t* dn, uint64_t n, const double* coefs,
                 uint32_t n_states, double eps_var, double* eps_tried, uint64_t* n_new, bool ts, DevBuf* bufs) {
  DevBuf &dE = bufs[3], &cnt = bufs[5], &off = bufs[6], &cand = bufs[7], &srt = bufs[8], &acc = bufs[9];
  const hipStream_t st = ctx->stream;
  const uint32_t n_orb = ctx->n_orb;
  SelParams S;
  CHECK_RC(sel_prepare(ctx, up, dn, n, coefs, n_states, eps_tried, bufs, S));
  S.eps_var = eps_var;
  S.ts = ts ? 1u : 0u;
  uint32_t bad = 0;
  // Candidate budget: ~64 B each (keys, sorted keys, flags, offsets, sort
  // storage) within 3/4 of the memory the context can hand out.
  const size_t free_b = available_bytes(ctx);
  const uint64_t mem_cap = std::max<uint64_t>(1ull << 20, (uint64_t)(0.75 * (double)free_b / 64.0));
  uint64_t cap = mem_cap;
  if (const char* e = std::getenv("SHCI_SELECT_CAP")) {  // smaller buffers (tests of the fallback paths)
    const unsigned long long v = std::strtoull(e, nullptr, 10);
    if (v > 0) cap = std::min<uint64_t>(cap, v);
  }
  const uint32_t grid = (uint32_t)std::min<uint64_t>(n, 1u << 20);
  uint64_t total = 0, n_acc = 0;
  auto run_append = [&](uint64_t cap_a) -> int {
    CHECK_RC(ensure(ctx, cand, (cap_a + 1) * sizeof(DetKey)));
    S.out = P<DetKey>(cand);
    S.cap = cap_a;
    HIPC(hipMemsetAsync(P<uint64_t>(ctx->scal) + 1, 0, 8, st));
    k_sel_walk<kSelAppend><<<grid, 64, 0, st>>>(S);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(&total, P<uint64_t>(ctx->scal) + 1, 8, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    return SHCI_OK;
  };
  // One pass into a buffer sized for the common case; if the candidates do
  // not fit, the sweep starts again (eps_tried_prev re-uploaded) with an
  // exact buffer, or, beyond the budget, in chunks of dets from a count pass.
  uint64_t cap0 = std::min<uint64_t>(cap, std::max<uint64_t>(1ull << 22, 8 * n));
  if (const char* e = std::getenv("SHCI_SELECT_FIRST")) {  // a smaller first buffer (tests of the repeat)
    const unsigned long long v = std::strtoull(e, nullptr, 10);
    if (v > 0) cap0 = std::min<uint64_t>(cap0, v);
  }
  HIPC(hipEventRecord(ctx->ev[1], st));
  CHECK_RC(run_append(cap0));
  HIPC(hipEventRecord(ctx->ev[2], st));
  bool done = total <= cap0;
  if (!done) {
    HIPC(hipMemcpyAsync(dE.p, eps_tried, n * 8, hipMemcpyHostToDevice, st));
    if (total <= cap) {
      const uint64_t need = total;
      CHECK_RC(run_append(need));
      if (total != need) return fail(ctx, SHCI_ERR_STATE, "selection: the repeated pass found a different count");
      done = true;
    }
  }
  if (done) {
    if (total > 0) {
      CHECK_RC(ensure(ctx, srt, total * sizeof(DetKey)));
      CHECK_RC(ensure(ctx, acc, total * sizeof(DetKey)));
      CHECK_RC(unique_det_keys(ctx, P<DetKey>(cand), P<DetKey>(srt), total, P<DetKey>(acc), &n_acc));
    }
  } else {
    // count pass, offsets, then fill passes over chunks of dets
    CHECK_RC(ensure(ctx, cnt, (n + 1) * 8));
    CHECK_RC(ensure(ctx, off, (n + 1) * 8));
    S.cnt = P<uint64_t>(cnt);
    S.off = P<uint64_t>(off);
    k_sel_walk<kSelCount><<<grid, 64, 0, st>>>(S);
    HIPC(hipMemsetAsync(P<uint64_t>(cnt) + n, 0, 8, st));
    CHECK_RC(excl_scan(ctx, P<uint64_t>(cnt), P<uint64_t>(off), n + 1));
    HIPC(hipMemcpyAsync(&total, P<uint64_t>(off) + n, 8, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    std::vector<uint64_t> hoff(n + 1);
    HIPC(hipMemcpy(hoff.data(), off.p, (n + 1) * 8, hipMemcpyDeviceToHost));
    HIPC(hipMemsetAsync(P<uint32_t>(ctx->scal) + 1, 0, 4, st));
    for (uint64_t a = 0; a < n;) {
      const uint64_t oa = hoff[a];
      uint64_t b = a + 1;
      while (b < n && hoff[b + 1] - oa <= cap / 2) b++;
      const uint64_t m = hoff[b] - oa;
      if (m > mem_cap) return fail(ctx, SHCI_ERR_OOM, "selection: one det's candidates exceed the memory budget");
      CHECK_RC(ensure(ctx, cand, (n_acc + m + 1) * sizeof(DetKey)));
      if (n_acc) HIPC(hipMemcpyAsync(cand.p, acc.p, n_acc * sizeof(DetKey), hipMemcpyDeviceToDevice, st));
      S.out = P<DetKey>(cand) + n_acc;
      S.base = oa;
      S.i0 = a;
      S.i1 = b;
      k_sel_walk<kSelFill><<<(uint32_t)std::min<uint64_t>(b - a, 1u << 20), 64, 0, st>>>(S);
      HIPC(hipGetLastError());
      if (n_acc + m > 0) {
        CHECK_RC(ensure(ctx, srt, (n_acc + m) * sizeof(DetKey)));
        CHECK_RC(ensure(ctx, acc, (n_acc + m) * sizeof(DetKey)));
        CHECK_RC(unique_det_keys(ctx, P<DetKey>(cand), P<DetKey>(srt), n_acc + m, P<DetKey>(acc), &n_acc));
      }
      a = b;
    }
    HIPC(hipMemcpyAsync(&bad, P<uint32_t>(ctx->scal) + 1, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (bad) return fail(ctx, SHCI_ERR_STATE, "selection: fill pass differs from the count pass");
  }
  HIPC(hipEventRecord(ctx->ev[3], st));
  CHECK_RC(ensure(ctx, ctx->sel_up, std::max<uint64_t>(n_acc, 1) * 8));
  CHECK_RC(ensure(ctx, ctx->sel_dn, std::max<uint64_t>(n_acc, 1) * 8));
  if (n_acc)
    k_sel_unpack<<<div_up(n_acc, 256), 256, 0, st>>>(P<DetKey>(acc), n_acc, n_orb, P<uint64_t>(ctx->sel_up),
                                                     P<uint64_t>(ctx->sel_dn));
  HIPC(hipMemcpyAsync(eps_tried, dE.p, n * 8, hipMemcpyDeviceToHost, st));
  HIPC(hipEventRecord(ctx->ev[4], st));
  HIPC(hipStreamSynchronize(st));
  for (int k = 0; k < 4; k++) HIPC(hipEventElapsedTime(&ctx->sel_ms[k], ctx->ev[k], ctx->ev[k + 1]));
  ctx->sel_n = n_acc;
  *n_new = n_acc;
  return SHCI_OK;
}


// Deterministic PT2 reduction: run starts of the sorted (det a, H_ai c_i)
// pairs, then per distinct a the sum of its terms in their sorted order (the
// stable sort keeps the walk's det order i), and its contribution
// hc_sum^2 / (E_var - H_aa) (solver.h:728-733).
__global__ void k_pt_starts(const uint64_t* __restrict__ flag, const uint64_t* __restrict__ pos, uint64_t m,
                            uint64_t* __restrict__ starts) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && flag[i]) starts[pos[i]] = i;
}

__global__ void k_pt_contrib(Ints I, const DetKey* __restrict__ k, const double* __restrict__ v,
                             const uint64_t* __restrict__ starts, uint64_t nu, double energy_var,
                             double* __restrict__ contrib) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nu) return;
  const uint64_t b = starts[r], e = starts[r + 1];
  double sum = 0.0;
  for (uint64_t t = b; t < e; t++) sum += v[t];
  uint64_t au, ad;
  unpack_det(k[b], (uint32_t)I.n, au, ad);
  const double haa = diag_elem(I, au, ad);
  contrib[r] = sum * sum / (energy_var - haa);
}

// Fixed-order sums: block b adds the contiguous chunk [b * chunk, (b + 1) *
// chunk) (strided per thread, then an LDS tree); the host adds the block
// sums in order.  Deterministic for a given length.
__global__ __launch_bounds__(256) void k_sum_chunks(const double* __restrict__ x, uint64_t m, uint64_t chunk,
                                                    double* __restrict__ part) {
  __shared__ double sh[256];
  const uint64_t b0 = (uint64_t)blockIdx.x * chunk, b1 = b0 + chunk < m ? b0 + chunk : m;
  double a = 0.0;
  for (uint64_t t = b0 + threadIdx.x; t < b1; t += 256) a += x[t];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

int sort_det_pairs(shci_ctx* ctx, const DetKey* kin, DetKey* kout, const double* vin, double* vout, uint64_t m,
                   int end_bit) {
  size_t tb = 0;
  HIPC(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kin, kout, vin, vout, m, DetKeyDecomposer{}, 0, end_bit,
                                          ctx->stream));
  CHECK_RC(ensure(ctx, ctx->tmp, tb));
  HIPC(hipcub::DeviceRadixSort::SortPairs(ctx->tmp.p, tb, kin, kout, vin, vout, m, DetKeyDecomposer{}, 0, end_bit,
                                          ctx->stream));
  return SHCI_OK;
}

// One deterministic PT2 pass (all batches) over the dets: bufs as for the
// selection plus [10..14] values, sorted values, run flags, their offsets,
// run starts.
int pt_dtm_sweep(shci_ctx* ctx, const uint64_t* up, const uint64_t* dn, uint64_t n, const double* coefs,
                 double eps_pt, double energy_var, uint32_t n_batches, double* energy_pt, uint64_t* n_pt,
                 DevBuf* bufs) {
  DevBuf &cnt = bufs[5], &off = bufs[6], &keys = bufs[7], &srt = bufs[8], &vals = bufs[10], &vsrt = bufs[11],
         &fl = bufs[12], &pos = bufs[13], &runs = bufs[14], &sums = bufs[9];
  const hipStream_t st = ctx->stream;
  SelParams S;
  CHECK_RC(sel_prepare(ctx, up, dn, n, coefs, 1, nullptr, bufs, S));
  S.eps_var = eps_pt;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(n, 1u << 20);
  CHECK_RC(ensure(ctx, cnt, (n + 1) * 8));
  CHECK_RC(ensure(ctx, off, (n + 1) * 8));
  S.cnt = P<uint64_t>(cnt);
  S.off = P<uint64_t>(off);
  const size_t free_b = available_bytes(ctx);
  // ~104 B per pair: keys and values, their sorted copies, flags, offsets,
  // run starts and contributions, sort storage
  const uint64_t cap = std::max<uint64_t>(1ull << 20, (uint64_t)(0.75 * (double)free_b / 104.0));
  const bool autob = n_batches == 0;
  uint32_t nb = autob ? 1 : n_batches;
  double e_sum = 0.0;
  uint64_t np_sum = 0;
  HIPC(hipEventRecord(ctx->ev[1], st));
  for (uint32_t b = 0; b < nb; b++) {
    S.n_batches = nb;
    S.batch = b;
    k_sel_walk<kSelCount, true><<<grid, 64, 0, st>>>(S);
    HIPC(hipMemsetAsync(P<uint64_t>(cnt) + n, 0, 8, st));
    CHECK_RC(excl_scan(ctx, P<uint64_t>(cnt), P<uint64_t>(off), n + 1));
    uint64_t m = 0;
    HIPC(hipMemcpyAsync(&m, P<uint64_t>(off) + n, 8, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (m > cap) {
      if (!autob || b > 0) return fail(ctx, SHCI_ERR_OOM, "PT: a batch's terms exceed the memory budget");
      nb = (uint32_t)std::min<uint64_t>((m + cap - 1) / cap * 2, 1u << 16);  // hash batches are uneven
      b = (uint32_t)-1;
      continue;
    }
    if (m == 0) continue;
    CHECK_RC(ensure(ctx, keys, m * sizeof(DetKey)));
    CHECK_RC(ensure(ctx, srt, m * sizeof(DetKey)));
    CHECK_RC(ensure(ctx, vals, m * 8));
    CHECK_RC(ensure(ctx, vsrt, m * 8));
    CHECK_RC(ensure(ctx, fl, (m + 1) * 8));
    CHECK_RC(ensure(ctx, pos, (m + 1) * 8));
    S.out = P<DetKey>(keys);
    S.outv = P<double>(vals);
    S.base = 0;
    HIPC(hipMemsetAsync(P<uint32_t>(ctx->scal) + 1, 0, 4, st));
    k_sel_walk<kSelFill, true><<<grid, 64, 0, st>>>(S);
    HIPC(hipGetLastError());
    CHECK_RC(sort_det_pairs(ctx, P<DetKey>(keys), P<DetKey>(srt), P<double>(vals), P<double>(vsrt), m,
                            2 * (int)ctx->n_orb));
    k_sel_first_flags<<<div_up(m, 256), 256, 0, st>>>(P<DetKey>(srt), m, P<uint64_t>(fl));
    HIPC(hipMemsetAsync(P<uint64_t>(fl) + m, 0, 8, st));
    CHECK_RC(excl_scan(ctx, P<uint64_t>(fl), P<uint64_t>(pos), m + 1));
    uint64_t nu = 0;
    uint32_t bad = 0;
    HIPC(hipMemcpyAsync(&nu, P<uint64_t>(pos) + m, 8, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(&bad, P<uint32_t>(ctx->scal) + 1, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    if (bad) return fail(ctx, SHCI_ERR_STATE, "PT: fill pass differs from the count pass");
    // run starts (nu + 1 of them, the last = m: nu may exceed the det count);
    // the contributions (nu <= m) reuse the unsorted values
    CHECK_RC(ensure(ctx, runs, (nu + 1) * 8));
    uint64_t* starts = P<uint64_t>(runs);
    k_pt_starts<<<div_up(m, 256), 256, 0, st>>>(P<uint64_t>(fl), P<uint64_t>(pos), m, starts);
    HIPC(hipMemcpyAsync(starts + nu, &m, 8, hipMemcpyHostToDevice, st));
    double* contrib = P<double>(vals);
    k_pt_contrib<<<div_up(nu, 256), 256, 0, st>>>(S.I, P<DetKey>(srt), P<double>(vsrt), starts, nu, energy_var,
                                                   contrib);
    const uint64_t G = std::min<uint64_t>(1024, (nu + 255) / 256);
    const uint64_t chunk = (nu + G - 1) / G;
    CHECK_RC(ensure(ctx, sums, G * 8));
    k_sum_chunks<<<(uint32_t)G, 256, 0, st>>>(contrib, nu, chunk, P<double>(sums));
    std::vector<double> part(G);
    HIPC(hipMemcpyAsync(part.data(), sums.p, G * 8, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    for (double x : part) e_sum += x;
    np_sum += nu;
  }
  HIPC(hipEventRecord(ctx->ev[2], st));
  HIPC(hipEventSynchronize(ctx->ev[2]));
  HIPC(hipEventElapsedTime(&ctx->sel_ms[1], ctx->ev[1], ctx->ev[2]));
  HIPC(hipEventElapsedTime(&ctx->sel_ms[0], ctx->ev[0], ctx->ev[1]));
  ctx->sel_ms[2] = ctx->sel_ms[3] = 0.0f;
  ctx->pt_batches = nb;
  *energy_pt = e_sum;
  *n_pt = np_sum;
  return SHCI_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of the allocation-free product (shci_spmv_noalloc): the internal-
// order x and y and the beta-order copies of the A segments, four vectors of
// n_dets FP64 slots on 256-byte boundaries, at most kNoAllocBytes.  Each GPU
// holds its own copy for the life of the process.  In host code `scratch` is
// only a stand-in for the symbol: its device address comes from
// hipGetSymbolAddress, per device, and only that address reaches kernels.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = 256ull << 20;  // n_dets <= 8,388,608
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;

// The current device's address of `scratch`, looked up once per device.
int noalloc_scratch_address(shci_ctx* ctx) {
  static std::mutex mu;
  static unsigned char* addr[kMaxDevices] = {};
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return fail(ctx, SHCI_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[ctx->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ctx, SHCI_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    addr[ctx->device] = static_cast<unsigned char*>(p);
  }
  ctx->noalloc_scratch = addr[ctx->device];
  return SHCI_OK;
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

int shci_ctx_create(shci_ctx** out, int device) {
  if (!out) return SHCI_ERR_ARG;
  auto* ctx = new shci_ctx();
  ctx->device = device;
  if (const char* e = getenv("SHCI_CHUNK_MIB")) {
    const unsigned long long mib = strtoull(e, nullptr, 10);
    if (mib < 2 || mib % 2 || mib > (1ull << 20)) {
      *out = ctx;
      return fail(ctx, SHCI_ERR_ARG, "SHCI_CHUNK_MIB must be an even number of MiB in [2, 2^20]");
    }
    ctx->chunk = (size_t)mib << 20;
  }
  *out = ctx;
  HIPC(hipSetDevice(device));
  HIPC(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  for (auto& e : ctx->ev) HIPC(hipEventCreate(&e));
  CHECK_RC(noalloc_scratch_address(ctx));
  return SHCI_OK;
}

int shci_ctx_destroy(shci_ctx* ctx) {
  if (!ctx) return SHCI_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  DevBuf* bufs[] = {&ctx->h1, &ctx->eri, &ctx->tD, &ctx->tX, &ctx->kept, &ctx->up_c, &ctx->dn_c, &ctx->up_s, &ctx->dn_s,
                    &ctx->perm, &ctx->iperm, &ctx->up_gid, &ctx->up_gstart, &ctx->dn_gid, &ctx->dn_gstart,
                    &ctx->dnord_idx, &ctx->dnord_up, &ctx->m1a_id, &ctx->m1b_id, &ctx->dm1_vals, &ctx->dm1_rstart,
                    &ctx->dm1_seg, &ctx->flag, &ctx->mA_id, &ctx->mB_id, &ctx->seg_desc, &ctx->seg_type,
                    &ctx->seg_core, &ctx->seg_mem, &ctx->sv32, &ctx->sv64, &ctx->dnpos, &ctx->xb, &ctx->yb,
                    &ctx->tile_rb, &ctx->tile_re, &ctx->tile_u0, &ctx->tile_nu, &ctx->tile_tab, &ctx->seg_loc, &ctx->tile_roff, &ctx->tile_rec,
                    &ctx->diag,
                    &ctx->cnt,
                    &ctx->row_ptr, &ctx->row_end, &ctx->col, &ctx->v64, &ctx->v32, &ctx->len, &ctx->tmp, &ctx->ka, &ctx->kb,
                    &ctx->va, &ctx->vb, &ctx->scal, &ctx->part, &ctx->xv, &ctx->yv, &ctx->x32, &ctx->y32,
                    &ctx->dav, &ctx->sq_S, &ctx->sq_r, &ctx->sq_len, &ctx->sq_max, &ctx->hq_H, &ctx->hq_rs, &ctx->hq_len, &ctx->hq_max,
                    &ctx->sel_up, &ctx->sel_dn, &ctx->ts_rev, &ctx->ts_diag, &ctx->cap};
  for (DevBuf* b : bufs) release(ctx, *b);
  (void)cache_trim(ctx);
  pool_release(ctx);
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SHCI_OK;
}

const char* shci_last_error(const shci_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int shci_set_integrals(shci_ctx* ctx, uint32_t n_orb, const double* h1, const double* eri_packed, double core) {
  if (!ctx || !h1 || !eri_packed) return SHCI_ERR_ARG;
  if (n_orb == 0 || n_orb > 64) return fail(ctx, SHCI_ERR_UNSUPPORTED, "n_orb must be in [1, 64]");
  HIPC(hipSetDevice(ctx->device));
  const uint64_t n = n_orb;
  const uint64_t np = n * (n + 1) / 2, ne = np * (np + 1) / 2;
  std::vector<double> D(n * n * n), X(n * n * n);
  auto c2h = [](uint64_t a, uint64_t b) { return a > b ? (a * (a + 1)) / 2 + b : (b * (b + 1)) / 2 + a; };
  for (uint64_t p = 0; p < n; p++)
    for (uint64_t q = 0; q < n; q++)
      for (uint64_t k = 0; k < n; k++) {
        D[(p * n + q) * n + k] = eri_packed[c2h(c2h(p, q), c2h(k, k))];
        X[(p * n + q) * n + k] = eri_packed[c2h(c2h(p, k), c2h(k, q))];
      }
  // h1 symmetric lookup semantics: the value for key combine2(p,q) is h1[max][min].
  std::vector<double> hs(n * n);
  for (uint64_t p = 0; p < n; p++)
    for (uint64_t q = 0; q < n; q++) hs[p * n + q] = p >= q ? h1[p * n + q] : h1[q * n + p];
  CHECK_RC(ensure(ctx, ctx->h1, n * n * 8));
  CHECK_RC(ensure(ctx, ctx->eri, ne * 8));
  CHECK_RC(ensure(ctx, ctx->tD, n * n * n * 8));
  CHECK_RC(ensure(ctx, ctx->tX, n * n * n * 8));
  HIPC(hipMemcpy(ctx->h1.p, hs.data(), n * n * 8, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(ctx->eri.p, eri_packed, ne * 8, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(ctx->tD.p, D.data(), n * n * n * 8, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(ctx->tX.p, X.data(), n * n * n * 8, hipMemcpyHostToDevice));
  ctx->n_orb = n_orb;
  ctx->core = core;
  ctx->have_ints = true;
  ctx->sel_ne = 0;  // selection queues follow the integrals
  return SHCI_OK;
}

static int build_full(shci_ctx* ctx, const uint64_t* up, const uint64_t* dn, uint64_t n_dets, uint64_t n_dets_prev,
                      uint32_t flags) {
  if (!ctx) return SHCI_ERR_ARG;
  if (!up || !dn || n_dets == 0) return fail(ctx, SHCI_ERR_ARG, "empty determinant list");
  if (n_dets >= (1ull << 32)) return fail(ctx, SHCI_ERR_UNSUPPORTED, "n_dets >= 2^32");
  HIPC(hipSetDevice(ctx->device));
  if (ctx->built && n_dets == n_dets_prev && n_dets == ctx->n) return SHCI_OK;  // hamiltonian.h:91
  std::vector<uint64_t> hu, hd;
  const uint64_t* su = up;
  const uint64_t* sd = dn;
  if (flags & SHCI_DEVICE_PTRS) {
    hu.resize(1);
    hd.resize(1);
    HIPC(hipMemcpy(hu.data(), up, 8, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(hd.data(), dn, 8, hipMemcpyDeviceToHost));
    su = hu.data();
    sd = hd.data();
  }
  const uint32_t nu = __builtin_popcountll(su[0]), nd = __builtin_popcountll(sd[0]);
  if (!ctx->have_ints) return fail(ctx, SHCI_ERR_STATE, "integrals not set");
  CHECK_RC(ensure(ctx, ctx->up_c, n_dets * 8));
  CHECK_RC(ensure(ctx, ctx->dn_c, n_dets * 8));
  const hipMemcpyKind kind = (flags & SHCI_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  HIPC(hipMemcpyAsync(ctx->up_c.p, up, n_dets * 8, kind, ctx->stream));
  HIPC(hipMemcpyAsync(ctx->dn_c.p, dn, n_dets * 8, kind, ctx->stream));
  {
    const uint64_t mask = ctx->n_orb >= 64 ? 0ull : ~((1ull << ctx->n_orb) - 1ull);
    CHECK_RC(ensure(ctx, ctx->scal, 64 * 8));
    unsigned int* bad = reinterpret_cast<unsigned int*>(ctx->scal.p);
    unsigned int hb[2] = {0, 0};
    HIPC(hipMemsetAsync(bad, 0, 8, ctx->stream));
    k_validate<<<div_up(n_dets, 256), 256, 0, ctx->stream>>>(P<uint64_t>(ctx->up_c), P<uint64_t>(ctx->dn_c), n_dets,
                                                             mask, (int)nu, (int)nd, bad);
    HIPC(hipMemcpyAsync(hb, bad, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    if (hb[0])
      return fail(ctx, SHCI_ERR_ARG,
                  std::to_string(hb[0]) + " determinants outside n_orb or with electron counts differing from det 0");
  }
  ctx->n = n_dets;
  ctx->n_up = nu;
  ctx->n_dn = nd;
  return do_build(ctx, flags & ~SHCI_DEVICE_PTRS);
}

int shci_rebuild_hamiltonian(shci_ctx* ctx, uint32_t flags) {
  if (!ctx) return SHCI_ERR_ARG;
  if (ctx->n == 0) return fail(ctx, SHCI_ERR_STATE, "no determinants resident");
  HIPC(hipSetDevice(ctx->device));
  return do_build(ctx, flags & ~SHCI_DEVICE_PTRS);
}

static int update_full(shci_ctx* ctx);
static int ts_diag_update(shci_ctx* ctx);

int shci_update_existing_elems(shci_ctx* ctx) {
  if (!ctx) return SHCI_ERR_ARG;
  if (!ctx->built) return fail(ctx, SHCI_ERR_STATE, "no Hamiltonian built");
  HIPC(hipSetDevice(ctx->device));
  CHECK_RC(update_full(ctx));
  return ctx->ts ? ts_diag_update(ctx) : SHCI_OK;
}

static int update_full(shci_ctx* ctx) {
  if (ctx->segl) {
    k_diag<<<div_up(ctx->n, 256), 256, 0, ctx->stream>>>(make_ints(ctx), P<uint64_t>(ctx->up_s), P<uint64_t>(ctx->dn_s),
                                                        P<double>(ctx->diag), ctx->n);
    HIPC(hipGetLastError());
    CHECK_RC(seg_fill(ctx, 1));
    HIPC(hipStreamSynchronize(ctx->stream));
    return SHCI_OK;
  }
  k_update_values<<<div_up(ctx->n, 4), 256, 0, ctx->stream>>>(
      ctx->n, make_ints(ctx), P<uint64_t>(ctx->up_s), P<uint64_t>(ctx->dn_s), P<int64_t>(ctx->row_ptr),
      P<int64_t>(ctx->row_end), P<uint32_t>(ctx->col), P<double>(ctx->v64), P<float>(ctx->v32), P<double>(ctx->diag));
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(ctx->stream));
  return SHCI_OK;
}

int shci_get_counts(shci_ctx* ctx, uint64_t* n_dets, uint64_t* n_stored_upper, uint64_t* n_elems) {
  if (!ctx) return SHCI_ERR_ARG;
  if (!ctx->built) return fail(ctx, SHCI_ERR_STATE, "no Hamiltonian built");
  // Under time_sym the stored matrix is the one over the dets and their
  // reverses, not H_ts: its entry counts belong to neither the matrix the
  // caller's vectors live in nor the reference's (which counts H_ts), so only
  // the row count is answered.
  if (ctx->ts && (n_stored_upper || n_elems))
    return fail(ctx, SHCI_ERR_UNSUPPORTED, "time_sym: the element counts of H_ts are not available (n_dets only)");
  const uint64_t up = ctx->nnz / 2 + ctx->n;
  if (n_dets) *n_dets = ctx->ts ? ctx->n_red : ctx->n;
  if (n_stored_upper) *n_stored_upper = up;
  if (n_elems) *n_elems = 2 * up - ctx->n;
  return SHCI_OK;
}

int shci_get_diag(shci_ctx* ctx, double* diag) {
  if (!ctx || !diag) return SHCI_ERR_ARG;
  if (!ctx->built) return fail(ctx, SHCI_ERR_STATE, "no Hamiltonian built");
  HIPC(hipSetDevice(ctx->device));
  if (ctx->ts) {
    HIPC(hipMemcpy(diag, ctx->ts_diag.p, ctx->n_red * 8, hipMemcpyDeviceToHost));
    return SHCI_OK;
  }
  CHECK_RC(ensure(ctx, ctx->xv, ctx->n * 8));
  k_to_caller<double, double><<<div_up(ctx->n, 256), 256, 0, ctx->stream>>>(P<double>(ctx->diag), P<uint32_t>(ctx->perm),
                                                                            P<double>(ctx->xv), ctx->n);
  HIPC(hipMemcpyAsync(diag, ctx->xv.p, ctx->n * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return SHCI_OK;
}

static int spmv_full(shci_ctx* ctx, const double* x, double* y, uint32_t flags) {
  if (!ctx || !x || !y) return SHCI_ERR_ARG;
  if (!ctx->built) return fail(ctx, SHCI_ERR_STATE, "no Hamiltonian built");
  HIPC(hipSetDevice(ctx->device));
  const uint64_t n = ctx->n;
  const uint32_t g = div_up(n, 256);
  const hipStream_t st = ctx->stream;
  const bool fp32 = flags & SHCI_FP32;
  if (fp32 && !has_fp32(ctx)) return fail(ctx, SHCI_ERR_STATE, "FP32 store not built");
  if (!fp32 && !has_fp64(ctx)) return fail(ctx, SHCI_ERR_STATE, "FP64 store not built");
  CHECK_RC(ensure(ctx, ctx->xv, n * 8));
  CHECK_RC(ensure(ctx, ctx->yv, n * 8));
  const double* xd = x;
  if (!(flags & SHCI_DEVICE_PTRS)) {
    CHECK_RC(ensure(ctx, ctx->x32, n * 8));
    HIPC(hipMemcpyAsync(ctx->x32.p, x, n * 8, hipMemcpyHostToDevice, st));
    xd = P<double>(ctx->x32);
  }
  if (fp32) {
    CHECK_RC(ensure(ctx, ctx->y32, n * 8));
    float* xf = reinterpret_cast<float*>(P<double>(ctx->xv));
    float* yf = reinterpret_cast<float*>(P<double>(ctx->yv));
    k_to_internal<double, float><<<g, 256, 0, st>>>(xd, P<uint32_t>(ctx->perm), xf, n);
    CHECK_RC(launch_spmv<float, float, float>(ctx, P<float>(ctx->v32), 0.0, xf, yf));
    double* yo = (flags & SHCI_DEVICE_PTRS) ? y : P<double>(ctx->y32);
    k_to_caller<float, double><<<g, 256, 0, st>>>(yf, P<uint32_t>(ctx->perm), yo, n);
  } else {
    k_to_internal<double, double><<<g, 256, 0, st>>>(xd, P<uint32_t>(ctx->perm), P<double>(ctx->xv), n);
    CHECK_RC(launch_spmv<double, double, double>(ctx, P<double>(ctx->v64), 0.0, P<double>(ctx->xv),
                                                 P<double>(ctx->yv)));
    CHECK_RC(ensure(ctx, ctx->y32, n * 8));
    double* yo = (flags & SHCI_DEVICE_PTRS) ? y : P<double>(ctx->y32);
    k_to_caller<double, double><<<g, 256, 0, st>>>(P<double>(ctx->yv), P<uint32_t>(ctx->perm), yo, n);
  }
  if (!(flags & SHCI_DEVICE_PTRS)) HIPC(hipMemcpyAsync(y, ctx->y32.p, n * 8, hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  return SHCI_OK;
}

int shci_spmv_noalloc(shci_ctx* ctx, const double* x, double* y, uint32_t flags) {
  if (!ctx) return SHCI_ERR_ARG;
  if (!x || !y) return fail(ctx, SHCI_ERR_ARG, "shci_spmv_noalloc: null vector");
  if (!ctx->built) return fail(ctx, SHCI_ERR_STATE, "no Hamiltonian built");
  if (ctx->ts) return fail(ctx, SHCI_ERR_UNSUPPORTED, "shci_spmv_noalloc: time_sym (reduced basis) is not supported");
  const bool fp32 = flags & SHCI_FP32;
  if (fp32 && !has_fp32(ctx)) return fail(ctx, SHCI_ERR_STATE, "FP32 store not built");
  if (!fp32 && !has_fp64(ctx)) return fail(ctx, SHCI_ERR_STATE, "FP64 store not built");
  const uint64_t n = ctx->n;
  const size_t nst = (n + 63) & ~(uint64_t)63;  // vector stride: 512-byte multiples
  if (4 * nst * 8 > kNoAllocBytes)
    return fail(ctx, SHCI_ERR_UNSUPPORTED,
                "shci_spmv_noalloc: " + std::to_string(n) + " dets exceed its scratch (at most " +
                    std::to_string(kNoAllocBytes / 32) + ")");
  if (!ctx->noalloc_scratch) return fail(ctx, SHCI_ERR_STATE, "scratch address not set up");
  HIPC(hipSetDevice(ctx->device));
  const uint32_t g = div_up(n, 256);
  const hipStream_t st = ctx->stream;
  unsigned char* base = ctx->noalloc_scratch;
  if (fp32) {
    float* xv = reinterpret_cast<float*>(base);
    float* yv = reinterpret_cast<float*>(base + nst * 8);
    float* xb = reinterpret_cast<float*>(base + 2 * nst * 8);
    float* yb = reinterpret_cast<float*>(base + 3 * nst * 8);
    k_to_internal<double, float><<<g, 256, 0, st>>>(x, P<uint32_t>(ctx->perm), xv, n);
    CHECK_RC(launch_spmv<float, float, float>(ctx, P<float>(ctx->v32), 0.0, xv, yv, xb, yb));
    k_to_caller<float, double><<<g, 256, 0, st>>>(yv, P<uint32_t>(ctx->perm), y, n);
  } else {
    double* xv = reinterpret_cast<double*>(base);
    double* yv = reinterpret_cast<double*>(base + nst * 8);
    double* xb = reinterpret_cast<double*>(base + 2 * nst * 8);
    double* yb = reinterpret_cast<double*>(base + 3 * nst * 8);
    k_to_internal<double, double><<<g, 256, 0, st>>>(x, P<uint32_t>(ctx->perm), xv, n);
    CHECK_RC(launch_spmv<double, double, double>(ctx, P<double>(ctx->v64), 0.0, xv, yv, xb, yb));
    k_to_caller<double, double><<<g, 256, 0, st>>>(yv, P<uint32_t>(ctx->perm), y, n);
  }
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(st));
  return SHCI_OK;
}

int shci_spmm(shci_ctx* ctx, const double* X, double* Y, uint32_t k, uint32_t flags) {
  if (!ctx || !X || !Y || k == 0 || k > 8) return SHCI_ERR_ARG;
  if (!ctx->built) return fail(ctx, SHCI_ERR_STATE, "no Hamiltonian built");
  HIPC(hipSetDevice(ctx->device));
  if (ctx->ts) {  // reduced basis: one product per vector
    for (uint32_t v = 0; v < k; v++)
      CHECK_RC(shci_spmv(ctx, X + (size_t)v * ctx->n_red, Y + (size_t)v * ctx->n_red, flags));
    return SHCI_OK;
  }
  const uint64_t n = ctx->n;
  const hipStream_t st = ctx->stream;
  // The segment layout's block kernels stage the 4 vectors of x of a
  // one-wavefront segment in LDS (64 KiB: m <= 4096).
  const uint32_t seg_big = *std::max_element(ctx->seg_maxm, ctx->seg_maxm + 9);
  const bool block_ok = ctx->segl ? (size_t)seg_big * 4 * sizeof(float) <= 65536 : true;
  if (!(flags & SHCI_FP32) || !has_fp32(ctx) || !block_ok) {
    // FP64 block product as k SpMVs.
    for (uint32_t v = 0; v < k; v++) CHECK_RC(shci_spmv(ctx, X + (size_t)v * n, Y + (size_t)v * n, flags));
    return SHCI_OK;
  }
  const uint32_t Kp = 4;
  for (uint32_t v0 = 0; v0 < k; v0 += Kp) {
    const uint32_t kk = std::min(Kp, k - v0);
    DevBuf Xd, Yd, xi, yi;
    auto cleanup = [&]() {
      release(ctx, Xd);
      release(ctx, Yd);
      release(ctx, xi);
      release(ctx, yi);
    };
    int rc = SHCI_OK;
    const double* Xs = X + (size_t)v0 * n;
    double* Ys = Y + (size_t)v0 * n;
    if (!(flags & SHCI_DEVICE_PTRS)) {
      if ((rc = ensure(ctx, Xd, (size_t)kk * n * 8)) || (rc = ensure(ctx, Yd, (size_t)kk * n * 8))) {
        cleanup();
        return rc;
      }
      if (hipMemcpyAsync(Xd.p, Xs, (size_t)kk * n * 8, hipMemcpyHostToDevice, st) != hipSuccess) {
        cleanup();
        return fail(ctx, SHCI_ERR_HIP, "spmm copy");
      }
      Xs = P<double>(Xd);
    }
    if ((rc = ensure(ctx, xi, (size_t)Kp * n * 4)) || (rc = ensure(ctx, yi, (size_t)Kp * n * 4))) {
      cleanup();
      return rc;
    }
    k_interleave<<<div_up(n, 256), 256, 0, st>>>(Xs, P<uint32_t>(ctx->perm), P<float>(xi), kk, Kp, n);
    rc = ctx->segl ? launch_seg_spmv<4, float, float>(ctx, 0.0, P<float>(xi), P<float>(yi))
                   : launch_spmm4(ctx, 0.0, P<float>(xi), P<float>(yi));
    double* Yo = (flags & SHCI_DEVICE_PTRS) ? Ys : P<double>(Yd);
    if (rc == SHCI_OK) k_deinterleave<<<div_up(n, 256), 256, 0, st>>>(P<float>(yi), P<uint32_t>(ctx->perm), Yo, kk, Kp, n);
    if (rc == SHCI_OK && !(flags & SHCI_DEVICE_PTRS) &&
        hipMemcpyAsync(Ys, Yd.p, (size_t)kk * n * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
      rc = fail(ctx, SHCI_ERR_HIP, "spmm copy back");
    if (rc == SHCI_OK && hipStreamSynchronize(st) != hipSuccess) rc = fail(ctx, SHCI_ERR_HIP, "spmm sync");
    cleanup();
    CHECK_RC(rc);
  }
  return SHCI_OK;
}

static int davidson_full(shci_ctx* ctx, uint32_t n_states, const double* init, double tol, uint32_t mode,
                         uint32_t max_calls, double* evals, double* evecs, int* converged, double* traj_e,
                         double* traj_t, int traj_cap, int* n_traj) {
  if (!ctx || !init || !evals || n_states == 0 || n_states > 4) return ctx ? fail(ctx, SHCI_ERR_ARG, "bad davidson args") : SHCI_ERR_ARG;
  if (!ctx->built) return fail(ctx, SHCI_ERR_STATE, "no Hamiltonian built");
  const bool mixed = mode == SHCI_DAVIDSON_MIXED;
  if (mixed && !has_fp32(ctx)) return fail(ctx, SHCI_ERR_STATE, "mixed Davidson needs the FP32 store");
  if (!mixed && !has_fp64(ctx)) return fail(ctx, SHCI_ERR_STATE, "FP64 Davidson needs the FP64 store");
  HIPC(hipSetDevice(ctx->device));
  const uint64_t n = ctx->n;
  const uint32_t g = div_up(n, 256);
  const hipStream_t st = ctx->stream;
  int nt_local = 0;
  if (!n_traj) n_traj = &nt_local;
  *n_traj = 0;
  ctx->cap_n = 0;
  if (ctx->cap_max) CHECK_RC(ensure(ctx, ctx->cap, (size_t)ctx->cap_max * ((n + 63) & ~(uint64_t)63) * 8));
  const auto t0 = std::chrono::steady_clock::now();
  if (max_calls == 0) max_calls = 1;
  if (n == 1) {  // davidson.cc:20-26
    double d;
    HIPC(hipMemcpy(&d, ctx->diag.p, 8, hipMemcpyDeviceToHost));
    evals[0] = d;
    if (evecs) evecs[0] = 1.0;
    if (converged) *converged = 1;
    return SHCI_OK;
  }
  const uint32_t ns = (uint32_t)std::min<uint64_t>(n, n_states);
  const size_t n_store = std::min<uint64_t>(n, 5);
  const size_t K = ns * n_store;
  const size_t es = mixed ? 4 : 8;
  // Layout: v[K], hv[K], w[ns], hw[ns], tmp, init64[ns], work64[2].
  const size_t n_vec = 2 * K + 2 * ns + 1;
  // Every vector starts on a 256-byte boundary (stride rounded up to 64
  // elements): FP64 global atomics in the segment SpMV need natural alignment.
  const size_t nst = (n + 63) & ~(size_t)63;
  CHECK_RC(ensure(ctx, ctx->dav, n_vec * nst * es + (ns + 2) * nst * 8));
  char* base = static_cast<char*>(ctx->dav.p);
  double* init_i = reinterpret_cast<double*>(base + n_vec * nst * es);
  double* w64 = init_i + ns * nst;
  double* y64 = w64 + nst;
  // init -> internal order (FP64)
  {
    DevBuf hbuf;
    CHECK_RC(ensure(ctx, hbuf, (size_t)ns * n * 8));
    int rc = SHCI_OK;
    if (hipMemcpyAsync(hbuf.p, init, (size_t)ns * n * 8, hipMemcpyHostToDevice, st) != hipSuccess)
      rc = fail(ctx, SHCI_ERR_HIP, "init copy");
    for (uint32_t s = 0; s < ns && rc == SHCI_OK; s++)
      k_to_internal<double, double><<<g, 256, 0, st>>>(P<double>(hbuf) + (size_t)s * n, P<uint32_t>(ctx->perm),
                                                       init_i + (size_t)s * n, n);
    if (rc == SHCI_OK && hipStreamSynchronize(st) != hipSuccess) rc = fail(ctx, SHCI_ERR_HIP, "init sync");
    release(ctx, hbuf);
    CHECK_RC(rc);
  }
  std::vector<double> result(ns, 0.0);
  int conv = 0;
  // internal index of the only nonzero entry of the ground-state start vector
  // (n if there are several)
  uint64_t single_k = n;
  {
    uint64_t nz = 0, k = 0;
    for (uint64_t i = 0; i < n && nz < 2; i++)
      if (init[i] != 0.0) {
        nz++;
        k = i;
      }
    if (nz == 1) {
      uint32_t ki = 0;
      HIPC(hipMemcpy(&ki, P<uint32_t>(ctx->iperm) + k, 4, hipMemcpyDeviceToHost));
      single_k = ki;
    }
  }
  auto run_mode = [&](auto tag) -> int {
    using T = decltype(tag);
    Davidson<T> D;
    D.ctx = ctx;
    D.ns = ns;
    D.K = K;
    D.n = n;
    D.t0 = t0;
    D.traj_e = traj_e;
    D.traj_t = traj_t;
    D.traj_cap = traj_cap;
    D.n_traj = n_traj;
    T* vb = reinterpret_cast<T*>(base);
    for (size_t i = 0; i < K; i++) D.v.push_back(vb + i * nst);
    for (size_t i = 0; i < K; i++) D.hv.push_back(vb + (K + i) * nst);
    for (size_t i = 0; i < ns; i++) D.w.push_back(vb + (2 * K + i) * nst);
    for (size_t i = 0; i < ns; i++) D.hw.push_back(vb + (2 * K + ns + i) * nst);
    D.tmp = vb + (2 * K + 2 * ns) * nst;
    D.shift = 0.0;
    if (mixed && single_k < n) {
      // A start vector with one nonzero entry k: its Rayleigh quotient is H_kk.
      HIPC(hipMemcpy(&D.shift, P<double>(ctx->diag) + single_k, 8, hipMemcpyDeviceToHost));
      D.report_shift = D.shift;
    } else if (mixed) {
      // Shift = Rayleigh quotient of the initial ground-state vector: keeps
      // FP32 values of H - shift small so FP32 row sums lose no absolute
      // accuracy.  It is subtracted on the FP64 diagonal and added back to
      // the eigenvalues, so any nearby value serves: an FP32 product (vectors
      // staged in w64's storage) with FP64 dots.
      float* x32 = reinterpret_cast<float*>(w64);
      float* y32 = x32 + nst;
      k_copy_scale<double, float><<<g, 256, 0, st>>>(init_i, x32, 1.0, n);
      CHECK_RC(launch_spmv<float, float, float>(ctx, P<float>(ctx->v32), 0.0, x32, y32));
      k_copy_scale<float, double><<<g, 256, 0, st>>>(y32, y64, 1.0, n);
      const double* b1[1] = {y64};
      const double* b2[1] = {init_i};
      double num, den;
      CHECK_RC(dots<double>(ctx, init_i, b1, 1, &num));
      CHECK_RC(dots<double>(ctx, init_i, b2, 1, &den));
      D.shift = num / den;
      D.report_shift = D.shift;
    }
    CHECK_RC(init_vectors<T>(ctx, D, init_i));
    // Mixed mode: the FP64-vector energies of the Ritz vectors, taken when a
    // call reports convergence.  Their residuals |H c - E c| (|c| = 1) guard
    // the eigenvalue-change test, which FP32 products can satisfy by
    // stagnation: an exactly degenerate pair whose partner lies in a cluster
    // the start vectors reach only through couplings near FP32 resolution
    // stopped 3e-4 Ha high.  A residual above the guard makes the call a
    // stall, and the loop restarts from the Ritz vectors.  The guard sits far
    // above the FP32 noise of a converged residual (~1e-6) and above the
    // residual that an eigenvalue error of ~100 tol implies (10 sqrt(tol) for
    // a gap near 1).
    const double guard = std::max(10.0 * std::sqrt(tol), 1e-4);
    std::vector<double> reeval(ns, 0.0);
    bool have_reeval = false;
    auto reevaluate = [&](double& worst) -> int {
      worst = 0.0;
      for (uint32_t s = 0; s < ns; s++) {
        k_copy_scale<T, double><<<g, 256, 0, st>>>(D.w[s], w64, 1.0, n);
        CHECK_RC(launch_spmv<float, double, double>(ctx, P<float>(ctx->v32), 0.0, w64, y64));
        const double* b[2] = {y64, w64};
        double d[2], yy;
        CHECK_RC(dots<double>(ctx, w64, b, 2, d));
        CHECK_RC(dots<double>(ctx, y64, b, 1, &yy));
        reeval[s] = d[0] / d[1];
        const double r2 = yy / d[1] - reeval[s] * reeval[s];
        worst = std::max(worst, std::sqrt(std::max(r2, 0.0)));
      }
      return SHCI_OK;
    };
    for (uint32_t call = 0; call < max_calls; call++) {
      CHECK_RC(D.run(tol));
      if (D.converged && mixed) {
        double worst;
        CHECK_RC(reevaluate(worst));
        have_reeval = true;
        if (worst > guard) {
          D.converged = false;
          have_reeval = call + 1 == max_calls;  // the last call's values stand
        }
      }
      if (D.converged) break;
      if (call + 1 < max_calls) {
        // Warm restart from the Ritz vectors (the reference caller's loop, solver.h:430-530).
        for (uint32_t s = 0; s < ns; s++)
          k_copy_scale<T, double><<<g, 256, 0, st>>>(D.w[s], init_i + (size_t)s * n, 1.0, n);
        CHECK_RC(init_vectors<T>(ctx, D, init_i));
      }
    }
    conv = D.converged ? 1 : 0;
    for (uint32_t s = 0; s < ns; s++) result[s] = D.lowest[s] + D.shift;
    // Eigenvectors (caller order) and, in mixed mode, the FP64 re-evaluated energy.
    for (uint32_t s = 0; s < ns; s++) {
      k_copy_scale<T, double><<<g, 256, 0, st>>>(D.w[s], w64, 1.0, n);
      if (mixed) {
        if (have_reeval) {
          result[s] = reeval[s];
        } else {
          CHECK_RC(launch_spmv<float, double, double>(ctx, P<float>(ctx->v32), 0.0, w64, y64));
          const double* b1[1] = {y64};
          const double* b2[1] = {w64};
          double num, den;
          CHECK_RC(dots<double>(ctx, w64, b1, 1, &num));
          CHECK_RC(dots<double>(ctx, w64, b2, 1, &den));
          result[s] = num / den;
        }
        if (s == 0) D.record(result[0]);
      }
      if (evecs) {
        k_to_caller<double, double><<<g, 256, 0, st>>>(w64, P<uint32_t>(ctx->perm), y64, n);
        HIPC(hipMemcpyAsync(evecs + (size_t)s * n, y64, n * 8, hipMemcpyDeviceToHost, st));
        HIPC(hipStreamSynchronize(st));
      }
    }
    return SHCI_OK;
  };
  if (mixed) {
    CHECK_RC(run_mode(float{}));
  } else {
    CHECK_RC(run_mode(double{}));
  }
  for (uint32_t s = 0; s < ns; s++) evals[s] = result[s];
  if (evecs)
    for (uint32_t s = ns; s < n_states; s++) std::memcpy(evecs + (size_t)s * n, init + (size_t)s * n, n * 8);
  if (converged) *converged = conv;
  return SHCI_OK;
}

static int rdm_full(shci_ctx* ctx, uint32_t n_states, const double* coefs, double* rdm1, double* rdm2) {
  if (!ctx || !coefs || n_states == 0) return SHCI_ERR_ARG;
  if (!ctx->built) return fail(ctx, SHCI_ERR_STATE, "no Hamiltonian built");
  HIPC(hipSetDevice(ctx->device));
  const uint64_t n = ctx->n;
  const uint64_t no = ctx->n_orb;
  const uint64_t n2 = no * no * (no * no + 1) / 2;
  const hipStream_t st = ctx->stream;
  DevBuf cb, ci, r1, r2;
  DevBuf ht;  // the dets' index set (1-RDM lookups)
  auto cleanup = [&]() {
    release(ctx, cb);
    release(ctx, ci);
    release(ctx, r1);
    release(ctx, r2);
    release(ctx, ht);
  };
  int rc = SHCI_OK;
  if ((rc = ensure(ctx, cb, (size_t)n_states * n * 8)) || (rc = ensure(ctx, ci, (size_t)n_states * n * 8)) ||
      (rc = ensure(ctx, r1, no * no * 8)) || (rc = ensure(ctx, r2, n2 * 8))) {
    cleanup();
    return rc;
  }
  if (hipMemcpyAsync(cb.p, coefs, (size_t)n_states * n * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemsetAsync(r1.p, 0, no * no * 8, st) != hipSuccess || hipMemsetAsync(r2.p, 0, n2 * 8, st) != hipSuccess) {
    cleanup();
    return fail(ctx, SHCI_ERR_HIP, "rdm setup");
  }
  for (uint32_t s = 0; s < n_states; s++)
    k_to_internal<double, double><<<div_up(n, 256), 256, 0, st>>>(P<double>(cb) + (size_t)s * n, P<uint32_t>(ctx->perm),
                                                                  P<double>(ci) + (size_t)s * n, n);
  RdmParams R;
  R.n = n;
  R.n_orb = (int)no;
  R.n_states = n_states;
  R.up = P<uint64_t>(ctx->up_s);
  R.dn = P<uint64_t>(ctx->dn_s);
  R.rp = P<int64_t>(ctx->row_ptr);
  R.re = P<int64_t>(ctx->row_end);
  R.col = P<uint32_t>(ctx->col);
  R.c = P<double>(ci);
  R.rdm1 = P<double>(r1);
  R.rdm2 = P<double>(r2);
  // The 2-RDM over the stored pairs.  The 1-RDM over the det list: every
  // single of the list is a stored pair of a CSR build that keeps cut pairs,
  // and an owned slot of the segment layout (cut or not); a CSR build that
  // drops cut pairs looks the singles up instead (k_rdm1).
  const bool keep_zeros = (ctx->build_flags & SHCI_KEEP_ZEROS) != 0;
  R.pair1 = (ctx->segl || keep_zeros) ? 1 : 0;
  if (!R.pair1) {
    uint64_t H = 1024;
    while (H < 4 * n) H <<= 1;
    if ((rc = ensure(ctx, ht, H * 8)) != SHCI_OK) {
      cleanup();
      return rc;
    }
    if (hipMemsetAsync(ht.p, 0xFF, H * 8, st) != hipSuccess) {
      cleanup();
      return fail(ctx, SHCI_ERR_HIP, "rdm index set");
    }
    uint64_t u0 = 0, d0 = 0;
    if (hipMemcpy(&u0, ctx->up_s.p, 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&d0, ctx->dn_s.p, 8, hipMemcpyDeviceToHost) != hipSuccess) {
      cleanup();
      return fail(ctx, SHCI_ERR_HIP, "rdm electron count");
    }
    const uint32_t ne = (uint32_t)(__builtin_popcountll(u0) + __builtin_popcountll(d0));
    k_det_index_insert<<<div_up(n, 256), 256, 0, st>>>(R.up, R.dn, n, P<unsigned long long>(ht), H - 1);
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(div_up(n * ne, 256), 2048));
    k_rdm1<<<blocks, 256, no * no * 8, st>>>(R, P<unsigned long long>(ht), H - 1, ne);
  }
  k_rdm_diag<<<(uint32_t)std::min<uint64_t>(div_up(n, 256), 2048), 256, (no * (no + 1) + no) * 8, st>>>(R);
  if (ctx->segl) {
    if (ctx->n_seg)
      k_seg_rdm<<<div_up(ctx->n_seg, 4), 256, 0, st>>>(
          R, P<SegDesc>(ctx->seg_desc), P<uint32_t>(ctx->seg_mem), P<uint8_t>(ctx->seg_type),
          P<uint64_t>(ctx->seg_core), P<double>(ctx->sv64), P<float>(ctx->sv32), P<uint32_t>(ctx->dnord_idx),
          ctx->ra, ctx->rb,
          (ctx->build_flags & SHCI_KEEP_ZEROS) ? 1 : 0, ctx->n_seg);
  } else {
    k_rdm<<<div_up(n, 4), 256, 0, st>>>(R);
  }
  if (hipGetLastError() != hipSuccess) rc = fail(ctx, SHCI_ERR_HIP, "rdm launch");
  if (rc == SHCI_OK && rdm1 && hipMemcpyAsync(rdm1, r1.p, no * no * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
    rc = fail(ctx, SHCI_ERR_HIP, "rdm1 copy");
  if (rc == SHCI_OK && rdm2 && hipMemcpyAsync(rdm2, r2.p, n2 * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
    rc = fail(ctx, SHCI_ERR_HIP, "rdm2 copy");
  if (rc == SHCI_OK && hipStreamSynchronize(st) != hipSuccess) rc = fail(ctx, SHCI_ERR_HIP, "rdm sync");
  cleanup();
  return rc;
}

int shci_export_upper(shci_ctx* ctx, int64_t* row_ptr, int64_t* cols, double* vals) {
  if (!ctx || !row_ptr) return SHCI_ERR_ARG;
  if (!ctx->built) return fail(ctx, SHCI_ERR_STATE, "no Hamiltonian built");
  if (ctx->segl) return fail(ctx, SHCI_ERR_STATE, "row export needs the CSR layout (build without SHCI_LAYOUT_SEGMENTS)");
  if (ctx->ts) return fail(ctx, SHCI_ERR_UNSUPPORTED, "row export of a time_sym build (its rows are those of the expanded list)");
  HIPC(hipSetDevice(ctx->device));
  const uint64_t n = ctx->n;
  const uint64_t m = ctx->nnz / 2 + n;
  const hipStream_t st = ctx->stream;
  DevBuf cnt, orp, oc, ov;
  auto cleanup = [&]() {
    release(ctx, cnt);
    release(ctx, orp);
    release(ctx, oc);
    release(ctx, ov);
  };
  int rc = SHCI_OK;
  if ((rc = ensure(ctx, cnt, (n + 1) * 8)) || (rc = ensure(ctx, orp, (n + 1) * 8))) {
    cleanup();
    return rc;
  }
  if (hipMemsetAsync(cnt.p, 0, (n + 1) * 8, st) != hipSuccess) {
    cleanup();
    return fail(ctx, SHCI_ERR_HIP, "export memset");
  }
  k_export_count<<<div_up(n, 4), 256, 0, st>>>(n, P<int64_t>(ctx->row_ptr), P<int64_t>(ctx->row_end), P<uint32_t>(ctx->col),
                                                P<uint32_t>(ctx->perm), P<unsigned long long>(cnt));
  rc = excl_scan(ctx, P<unsigned long long>(cnt), P<unsigned long long>(orp), n + 1);
  unsigned long long total = 0;
  if (rc == SHCI_OK && (hipMemcpyAsync(&total, P<unsigned long long>(orp) + n, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
                        hipStreamSynchronize(st) != hipSuccess))
    rc = fail(ctx, SHCI_ERR_HIP, "export count");
  if (rc == SHCI_OK && total != m) rc = fail(ctx, SHCI_ERR_STATE, "stored pattern is not symmetric");
  if (rc == SHCI_OK) rc = ensure(ctx, oc, m * 8);
  if (rc == SHCI_OK) rc = ensure(ctx, ov, m * 8);
  if (rc == SHCI_OK) {
    k_export_fill<<<div_up(n, 4), 256, 0, st>>>(n, P<int64_t>(ctx->row_ptr), P<int64_t>(ctx->row_end), P<uint32_t>(ctx->col), P<double>(ctx->v64),
                                               P<double>(ctx->diag), P<uint32_t>(ctx->perm),
                                               P<unsigned long long>(orp), P<int64_t>(oc), P<double>(ov));
    if (hipMemcpyAsync(row_ptr, orp.p, (n + 1) * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        (cols && hipMemcpyAsync(cols, oc.p, m * 8, hipMemcpyDeviceToHost, st) != hipSuccess) ||
        (vals && hipMemcpyAsync(vals, ov.p, m * 8, hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = fail(ctx, SHCI_ERR_HIP, "export copy");
  }
  cleanup();
  return rc;
}

int shci_select(shci_ctx* ctx, const uint64_t* up, const uint64_t* dn, uint64_t n_dets, const double* coefs,
                uint32_t n_states, double eps_var, double* eps_tried_prev, uint64_t* n_new, uint32_t flags) {
  if (!ctx) return SHCI_ERR_ARG;
  if (!up || !dn || !coefs || !eps_tried_prev || !n_new || n_dets == 0 || n_states == 0)
    return fail(ctx, SHCI_ERR_ARG, "selection: empty or null arguments");
  if (!ctx->have_ints) return fail(ctx, SHCI_ERR_STATE, "selection: no integrals");
  if (n_dets >= 0xFFFFFFFFull) return fail(ctx, SHCI_ERR_UNSUPPORTED, "selection: n_dets >= 2^32 - 1");
  HIPC(hipSetDevice(ctx->device));
  *n_new = 0;
  ctx->sel_n = 0;
  DevBuf bufs[10];
  if ((flags & SHCI_TIME_SYM) && __builtin_popcountll(up[0]) != __builtin_popcountll(dn[0]))
    return fail(ctx, SHCI_ERR_ARG, "selection: time_sym needs n_up == n_dn");
  const int rc = select_sweep(ctx, up, dn, n_dets, coefs, n_states, eps_var, eps_tried_prev, n_new,
                              (flags & SHCI_TIME_SYM) != 0, bufs);
  for (DevBuf& b : bufs) release(ctx, b);
  return rc;
}

int shci_pt_dtm(shci_ctx* ctx, const uint64_t* up, const uint64_t* dn, uint64_t n_dets, const double* coefs,
                double eps_pt_dtm, double energy_var, uint32_t n_batches, double* energy_pt, uint64_t* n_pt_dets) {
  if (!ctx) return SHCI_ERR_ARG;
  if (!up || !dn || !coefs || !energy_pt || !n_pt_dets || n_dets == 0 || !(eps_pt_dtm > 0.0))
    return fail(ctx, SHCI_ERR_ARG, "PT: empty or null arguments, or eps_pt_dtm <= 0");
  if (!ctx->have_ints) return fail(ctx, SHCI_ERR_STATE, "PT: no integrals");
  if (n_dets >= 0xFFFFFFFFull) return fail(ctx, SHCI_ERR_UNSUPPORTED, "PT: n_dets >= 2^32 - 1");
  HIPC(hipSetDevice(ctx->device));
  *energy_pt = 0.0;
  *n_pt_dets = 0;
  DevBuf bufs[15];
  const int rc = pt_dtm_sweep(ctx, up, dn, n_dets, coefs, eps_pt_dtm, energy_var, n_batches, energy_pt, n_pt_dets,
                              bufs);
  for (DevBuf& b : bufs) release(ctx, b);
  return rc;
}

int shci_select_timings(shci_ctx* ctx, float* ms) {
  if (!ctx || !ms) return SHCI_ERR_ARG;
  for (int k = 0; k < 4; k++) ms[k] = ctx->sel_ms[k];
  return SHCI_OK;
}

int shci_select_result(shci_ctx* ctx, uint64_t* new_up, uint64_t* new_dn, uint64_t n) {
  if (!ctx) return SHCI_ERR_ARG;
  if (n != ctx->sel_n) return fail(ctx, SHCI_ERR_ARG, "selection result: n differs from the last sweep's count");
  if (n == 0) return SHCI_OK;
  if (!new_up || !new_dn) return fail(ctx, SHCI_ERR_ARG, "selection result: null output");
  HIPC(hipSetDevice(ctx->device));
  HIPC(hipMemcpy(new_up, ctx->sel_up.p, n * 8, hipMemcpyDeviceToHost));
  HIPC(hipMemcpy(new_dn, ctx->sel_dn.p, n * 8, hipMemcpyDeviceToHost));
  return SHCI_OK;
}

int shci_clear(shci_ctx* ctx) {
  if (!ctx) return SHCI_ERR_ARG;
  HIPC(hipSetDevice(ctx->device));
  HIPC(hipStreamSynchronize(ctx->stream));
  DevBuf* bufs[] = {&ctx->up_c, &ctx->dn_c, &ctx->up_s, &ctx->dn_s, &ctx->perm, &ctx->iperm, &ctx->up_gid,
                    &ctx->up_gstart, &ctx->dn_gid, &ctx->dn_gstart, &ctx->dnord_idx, &ctx->dnord_up,
                    &ctx->m1a_id, &ctx->m1b_id, &ctx->dm1_vals, &ctx->dm1_rstart, &ctx->dm1_seg, &ctx->diag,
                    &ctx->mA_id, &ctx->mB_id, &ctx->seg_desc, &ctx->seg_type, &ctx->seg_core, &ctx->seg_mem,
                    &ctx->sv32, &ctx->sv64, &ctx->dnpos, &ctx->xb, &ctx->yb, &ctx->tile_rb, &ctx->tile_re,
                    &ctx->tile_u0, &ctx->tile_nu, &ctx->tile_tab, &ctx->seg_loc, &ctx->tile_roff, &ctx->tile_rec,
                    &ctx->cnt, &ctx->row_ptr, &ctx->row_end, &ctx->row_end, &ctx->col, &ctx->v64,
                    &ctx->v32, &ctx->len, &ctx->tmp, &ctx->ka, &ctx->kb, &ctx->va, &ctx->vb, &ctx->xv,
                    &ctx->yv, &ctx->x32, &ctx->y32, &ctx->dav, &ctx->cap};
  for (DevBuf* b : bufs) release(ctx, *b);
  CHECK_RC(cache_trim(ctx));
  pool_release(ctx);
  ctx->cap_n = 0;  // the captured vectors belonged to the dropped det list
  ctx->n = 0;
  ctx->nnz = 0;
  ctx->n_seg = ctx->n_mem = ctx->n_slots = ctx->n_pairs = ctx->n_tiles = 0;
  for (auto& c : ctx->seg_cs) c = 0;
  for (auto& c : ctx->seg_maxm) c = 0;
  ctx->segl = false;
  ctx->built = false;
  ctx->ts = false;
  ctx->n_red = 0;
  ctx->ts_e.clear();
  ctx->ts_pair.clear();
  return SHCI_OK;
}

int shci_get_build_timings(shci_ctx* ctx, float* phases) {
  if (!ctx || !phases) return SHCI_ERR_ARG;
  std::memcpy(phases, ctx->phase_ms, sizeof(ctx->phase_ms));
  return SHCI_OK;
}

int shci_time_spmv(shci_ctx* ctx, uint32_t flags, int reps, float* ms_avg, double* bytes) {
  if (!ctx || reps <= 0) return SHCI_ERR_ARG;
  if (!ctx->built) return fail(ctx, SHCI_ERR_STATE, "no Hamiltonian built");
  HIPC(hipSetDevice(ctx->device));
  const uint64_t n = ctx->n;
  const bool fp32 = flags & SHCI_FP32;
  const bool f64v = fp32 && (flags & SHCI_F64_VECTORS);
  if (fp32 && !has_fp32(ctx)) return fail(ctx, SHCI_ERR_STATE, "FP32 store not built");
  if (!fp32 && !has_fp64(ctx)) return fail(ctx, SHCI_ERR_STATE, "FP64 store not built");
  CHECK_RC(ensure(ctx, ctx->xv, n * 8));
  CHECK_RC(ensure(ctx, ctx->yv, n * 8));
  const hipStream_t st = ctx->stream;
  // x = 1/sqrt(n) everywhere (values do not change the traffic).
  if (fp32 && !f64v) {
    std::vector<float> h(n, (float)(1.0 / std::sqrt((double)n)));
    HIPC(hipMemcpyAsync(ctx->xv.p, h.data(), n * 4, hipMemcpyHostToDevice, st));
    HIPC(hipStreamSynchronize(st));
  } else {
    std::vector<double> h(n, 1.0 / std::sqrt((double)n));
    HIPC(hipMemcpyAsync(ctx->xv.p, h.data(), n * 8, hipMemcpyHostToDevice, st));
    HIPC(hipStreamSynchronize(st));
  }
  auto once = [&]() -> int {
    if (f64v)
      return launch_spmv<float, double, double>(ctx, P<float>(ctx->v32), 0.0, P<double>(ctx->xv), P<double>(ctx->yv));
    if (fp32)
      return launch_spmv<float, float, float>(ctx, P<float>(ctx->v32), 0.0, reinterpret_cast<float*>(ctx->xv.p),
                                              reinterpret_cast<float*>(ctx->yv.p));
    return launch_spmv<double, double, double>(ctx, P<double>(ctx->v64), 0.0, P<double>(ctx->xv), P<double>(ctx->yv));
  };
  CHECK_RC(once());  // warm-up
  HIPC(hipEventRecord(ctx->ev[0], st));
  for (int r = 0; r < reps; r++) CHECK_RC(once());
  HIPC(hipEventRecord(ctx->ev[1], st));
  HIPC(hipEventSynchronize(ctx->ev[1]));
  float ms = 0.f;
  HIPC(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  if (ms_avg) *ms_avg = ms / reps;
  if (bytes) {
    const double F = (double)ctx->nnz;  // off-diagonal full-symmetric entries
    const double vs = fp32 ? 4.0 : 8.0;  // value bytes
    const double xs = fp32 && !f64v ? 4.0 : 8.0;  // vector element bytes
    if (ctx->segl) {
      // every value slot once, member ids, segment descriptors, FP64
      // diagonal, x once, y once (the per-member x gathers and y atomics
      // reuse x and y from the caches and are not counted).
      *bytes = (double)ctx->n_slots * vs + (double)ctx->n_mem * 4.0 + (double)ctx->n_seg * sizeof(SegDesc) +
               (double)n * 8.0 + 2.0 * (double)n * xs;
    } else {
      // values + columns, row start/end (int64 each), FP64 diagonal, x once, y once.
      *bytes = F * (vs + 4.0) + (double)n * 16.0 + (double)n * 8.0 + 2.0 * (double)n * xs;
    }
  }
  return SHCI_OK;
}

int shci_capture_ritz(shci_ctx* ctx, uint32_t max_vectors) {
  if (!ctx) return SHCI_ERR_ARG;
  HIPC(hipSetDevice(ctx->device));
  HIPC(hipStreamSynchronize(ctx->stream));
  ctx->cap_max = max_vectors;
  ctx->cap_n = 0;
  if (!max_vectors) release(ctx, ctx->cap);
  return SHCI_OK;
}

int shci_ritz_energies(shci_ctx* ctx, double* e, uint32_t cap, uint32_t* n_captured) {
  if (!ctx) return SHCI_ERR_ARG;
  if (!ctx->built) return fail(ctx, SHCI_ERR_STATE, "no Hamiltonian built");
  if (!has_fp64(ctx)) return fail(ctx, SHCI_ERR_STATE, "the FP64 store is needed for the FP64 energies");
  HIPC(hipSetDevice(ctx->device));
  const uint64_t n = ctx->n;
  const size_t nst = (n + 63) & ~(uint64_t)63;
  if (ctx->cap_n && ctx->cap.bytes < (size_t)ctx->cap_n * nst * 8)
    return fail(ctx, SHCI_ERR_STATE, "the captured vectors belong to a smaller determinant list");
  if (n_captured) *n_captured = ctx->cap_n;
  const uint32_t m = std::min(cap, ctx->cap_n);
  if (m == 0) return SHCI_OK;
  if (!e) return fail(ctx, SHCI_ERR_ARG, "no output array");
  CHECK_RC(ensure(ctx, ctx->yv, n * 8));
  for (uint32_t k = 0; k < m; k++) {
    const double* x = P<double>(ctx->cap) + (size_t)k * nst;
    CHECK_RC(launch_spmv<double, double, double>(ctx, P<double>(ctx->v64), 0.0, x, P<double>(ctx->yv)));
    const double* b1[1] = {P<double>(ctx->yv)};
    const double* b2[1] = {x};
    double num, den;
    CHECK_RC(dots<double>(ctx, x, b1, 1, &num));
    CHECK_RC(dots<double>(ctx, x, b2, 1, &den));
    e[k] = num / den;
  }
  return SHCI_OK;
}

int shci_time_spmm(shci_ctx* ctx, uint32_t k, int reps, float* ms_avg, double* bytes) {
  if (!ctx || reps <= 0 || k != 4) return ctx ? fail(ctx, SHCI_ERR_ARG, "shci_time_spmm: k must be 4") : SHCI_ERR_ARG;
  if (!ctx->built) return fail(ctx, SHCI_ERR_STATE, "no Hamiltonian built");
  if (!has_fp32(ctx)) return fail(ctx, SHCI_ERR_STATE, "FP32 store not built");
  HIPC(hipSetDevice(ctx->device));
  const uint64_t n = ctx->n;
  const hipStream_t st = ctx->stream;
  TmpBuf xi(ctx), yi(ctx);
  CHECK_RC(ensure(ctx, xi.b, (size_t)4 * n * 4));
  CHECK_RC(ensure(ctx, yi.b, (size_t)4 * n * 4));
  {
    std::vector<float> h((size_t)4 * n, (float)(1.0 / std::sqrt((double)n)));
    HIPC(hipMemcpyAsync(xi.b.p, h.data(), (size_t)4 * n * 4, hipMemcpyHostToDevice, st));
    HIPC(hipStreamSynchronize(st));
  }
  auto once = [&]() -> int {
    return ctx->segl ? launch_seg_spmv<4, float, float>(ctx, 0.0, xi.p<float>(), yi.p<float>())
                     : launch_spmm4(ctx, 0.0, xi.p<float>(), yi.p<float>());
  };
  CHECK_RC(once());  // warm-up
  HIPC(hipEventRecord(ctx->ev[0], st));
  for (int r = 0; r < reps; r++) CHECK_RC(once());
  HIPC(hipEventRecord(ctx->ev[1], st));
  HIPC(hipEventSynchronize(ctx->ev[1]));
  float ms = 0.f;
  HIPC(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  if (ms_avg) *ms_avg = ms / reps;
  if (bytes) {
    // values (and columns or member ids) once for all 4 vectors; 4 x and 4 y values per det
    const double vec = 2.0 * 4.0 * 4.0 * (double)n + 8.0 * (double)n;
    if (ctx->segl)
      *bytes = (double)ctx->n_slots * 4.0 + (double)ctx->n_mem * 4.0 + (double)ctx->n_seg * sizeof(SegDesc) + vec;
    else
      *bytes = (double)ctx->nnz * 8.0 + (double)n * 16.0 + vec;
  }
  return SHCI_OK;
}

int shci_layout_stats(shci_ctx* ctx, uint64_t* stats) {
  if (!ctx || !stats) return SHCI_ERR_ARG;
  stats[0] = ctx->n;
  stats[1] = ctx->nnz;
  stats[2] = ctx->nnz_alloc;
  stats[3] = ctx->segl ? 1 : 0;
  stats[4] = ctx->n_seg;
  stats[5] = ctx->n_mem;
  stats[6] = ctx->seg_cs[9];  // segments with more than 64 members
  stats[7] = ctx->segl ? ctx->n_slots : 0;
  stats[8] = ctx->n_dm1_runs;
  stats[9] = ctx->n_ugroups;
  stats[10] = ctx->n_dgroups;
  stats[11] = (uint64_t)ctx->lanes;
  stats[12] = ctx->segl ? ctx->n_tiles : 0;
  stats[13] = 0;  // distinct dets summed over the tiles (the tile kernel's gathers and atomics)
  if (ctx->segl && ctx->n_tiles) {
    CHECK_RC(ensure(ctx, ctx->scal, 64 * 8));
    unsigned long long* out = reinterpret_cast<unsigned long long*>(ctx->scal.p) + 4;
    HIPC(hipMemsetAsync(out, 0, 8, ctx->stream));
    k_sum_u32<<<std::min<uint64_t>(1024, div_up(ctx->n_tiles, 256)), 256, 0, ctx->stream>>>(
        P<uint32_t>(ctx->tile_nu), ctx->n_tiles, out, 0x7FFFFFFFu);
    unsigned long long v = 0;
    CHECK_RC(copy_back(ctx, &v, out));
    stats[13] = v;
  }
  return SHCI_OK;
}

int shci_segment_kept(shci_ctx* ctx, uint64_t* out) {
  if (!ctx || !out) return SHCI_ERR_ARG;
  if (!ctx->segl || !ctx->seg_desc.p) return fail(ctx, SHCI_ERR_STATE, "segment layout not built");
  for (int r = 0; r < kSegRanges; r++) out[r] = ctx->seg_kept[r];
  return SHCI_OK;
}

int shci_segment_ranges(shci_ctx* ctx, uint64_t* out) {
  if (!ctx || !out) return SHCI_ERR_ARG;
  if (!ctx->segl || !ctx->seg_desc.p) return fail(ctx, SHCI_ERR_STATE, "segment layout not built");
  // members and slots are laid out in segment order: range totals are
  // differences of the first descriptors of consecutive ranges
  uint64_t moff[kSegRanges + 1], voff[kSegRanges + 1];
  for (int r = 0; r <= kSegRanges; r++) {
    if (ctx->seg_cs[r] >= ctx->n_seg) {
      moff[r] = ctx->n_mem;
      voff[r] = ctx->n_slots;
      continue;
    }
    SegDesc D;
    HIPC(hipMemcpyAsync(&D, P<SegDesc>(ctx->seg_desc) + ctx->seg_cs[r], sizeof(D), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    moff[r] = D.moff;
    voff[r] = D.voff;
  }
  for (int r = 0; r < kSegRanges; r++) {
    out[3 * r] = ctx->seg_cs[r + 1] - ctx->seg_cs[r];
    out[3 * r + 1] = moff[r + 1] - moff[r];
    out[3 * r + 2] = voff[r + 1] - voff[r];
  }
  return SHCI_OK;
}

// ---- time-reversal symmetry: reduced basis <-> expanded list (host) ----
// P maps reduced det i to (e_d + e_rev) / sqrt2 (up != dn) or e_d; expand is
// x_E = P x, contract y = P^T y_E (base_system.h:108-122 unpacks coefficients
// the same way, c / sqrt2 on both).
constexpr double kSqrt1_2 = 0.7071067811865475;  // Util::SQRT2_INV (src/util.h:60), as written there

static void ts_expand(const shci_ctx* ctx, const double* x, double* xe) {
  for (uint64_t i = 0; i < ctx->n_red; i++) {
    const uint64_t e = ctx->ts_e[i];
    if (ctx->ts_pair[i]) {
      xe[e] = xe[e + 1] = x[i] * kSqrt1_2;
    } else {
      xe[e] = x[i];
    }
  }
}

static void ts_contract(const shci_ctx* ctx, const double* ye, double* y) {
  for (uint64_t i = 0; i < ctx->n_red; i++) {
    const uint64_t e = ctx->ts_e[i];
    y[i] = ctx->ts_pair[i] ? (ye[e] + ye[e + 1]) * kSqrt1_2 : ye[e];
  }
}

// The reduced diagonal of H_ts from the current integrals (k_ts_diag).
static int ts_diag_update(shci_ctx* ctx) {
  const uint64_t n = ctx->n_red;
  DevBuf de, dp;
  int rc = SHCI_OK;
  if ((rc = ensure(ctx, de, n * 8)) || (rc = ensure(ctx, dp, n)) || (rc = ensure(ctx, ctx->ts_diag, n * 8))) {
    release(ctx, de);
    release(ctx, dp);
    return rc;
  }
  if (hipMemcpy(de.p, ctx->ts_e.data(), n * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dp.p, ctx->ts_pair.data(), n, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(ctx, SHCI_ERR_HIP, "time_sym maps");
  if (rc == SHCI_OK) {
    k_ts_diag<<<div_up(n, 256), 256, 0, ctx->stream>>>(make_ints(ctx), P<uint64_t>(ctx->up_c), P<uint64_t>(ctx->dn_c),
                                                      P<uint64_t>(de), P<uint8_t>(dp), n, P<double>(ctx->ts_diag));
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = fail(ctx, SHCI_ERR_HIP, "time_sym diagonal");
  }
  release(ctx, de);
  release(ctx, dp);
  return rc;
}

static int build_ts(shci_ctx* ctx, const uint64_t* up, const uint64_t* dn, uint64_t n, uint64_t n_prev,
                    uint32_t flags) {
  if (flags & SHCI_DEVICE_PTRS) return fail(ctx, SHCI_ERR_UNSUPPORTED, "time_sym: host determinant arrays only");
  if (!up || !dn || n == 0) return fail(ctx, SHCI_ERR_ARG, "empty determinant list");
  if (__builtin_popcountll(up[0]) != __builtin_popcountll(dn[0]))
    return fail(ctx, SHCI_ERR_ARG, "time_sym needs n_up == n_dn");
  for (uint64_t i = 0; i < n; i++)
    if (up[i] > dn[i]) return fail(ctx, SHCI_ERR_ARG, "time_sym: every det needs up <= dn (the canonical form)");
  if (ctx->ts && ctx->built && n == n_prev && n == ctx->n_red) return SHCI_OK;  // hamiltonian.h:91
  if (!ctx->ts) n_prev = 0;
  std::vector<uint64_t> eu, ed, e(n);
  std::vector<uint8_t> pair(n);
  eu.reserve(2 * n);
  ed.reserve(2 * n);
  for (uint64_t i = 0; i < n; i++) {
    e[i] = eu.size();
    eu.push_back(up[i]);
    ed.push_back(dn[i]);
    pair[i] = up[i] != dn[i];
    if (pair[i]) {
      eu.push_back(dn[i]);
      ed.push_back(up[i]);
    }
  }
  const uint64_t nE = eu.size();
  const uint64_t nE_prev = n_prev == 0 ? 0 : (n_prev < n ? e[n_prev] : nE);
  ctx->ts = false;
  CHECK_RC(build_full(ctx, eu.data(), ed.data(), nE, nE_prev, flags));
  // internal E index -> internal index of its reverse
  std::vector<uint32_t> perm(nE), iperm(nE), partner(nE), rev(nE);
  HIPC(hipMemcpy(perm.data(), ctx->perm.p, nE * 4, hipMemcpyDeviceToHost));
  for (uint64_t k = 0; k < nE; k++) iperm[perm[k]] = (uint32_t)k;
  for (uint64_t i = 0; i < n; i++) {
    partner[e[i]] = (uint32_t)(pair[i] ? e[i] + 1 : e[i]);
    if (pair[i]) partner[e[i] + 1] = (uint32_t)e[i];
  }
  for (uint64_t k = 0; k < nE; k++) rev[k] = iperm[partner[perm[k]]];
  CHECK_RC(ensure(ctx, ctx->ts_rev, nE * 4));
  HIPC(hipMemcpy(ctx->ts_rev.p, rev.data(), nE * 4, hipMemcpyHostToDevice));
  ctx->n_red = n;
  ctx->ts_e = std::move(e);
  ctx->ts_pair = std::move(pair);
  CHECK_RC(ts_diag_update(ctx));
  ctx->ts = true;
  return SHCI_OK;
}

int shci_build_hamiltonian(shci_ctx* ctx, const uint64_t* up, const uint64_t* dn, uint64_t n_dets,
                           uint64_t n_dets_prev, uint32_t flags) {
  if (!ctx) return SHCI_ERR_ARG;
  if (flags & SHCI_TIME_SYM) return build_ts(ctx, up, dn, n_dets, n_dets_prev, flags & ~SHCI_TIME_SYM);
  if (ctx->ts) {  // a plain build after a symmetric one starts afresh
    ctx->ts = false;
    n_dets_prev = 0;
  }
  return build_full(ctx, up, dn, n_dets, n_dets_prev, flags);
}

int shci_spmv(shci_ctx* ctx, const double* x, double* y, uint32_t flags) {
  if (!ctx || !x || !y) return SHCI_ERR_ARG;
  if (!ctx->ts) return spmv_full(ctx, x, y, flags);
  if (flags & SHCI_DEVICE_PTRS) return fail(ctx, SHCI_ERR_UNSUPPORTED, "time_sym: host vectors only");
  std::vector<double> xe(ctx->n), ye(ctx->n);
  ts_expand(ctx, x, xe.data());
  CHECK_RC(spmv_full(ctx, xe.data(), ye.data(), flags));
  ts_contract(ctx, ye.data(), y);
  return SHCI_OK;
}

int shci_davidson(shci_ctx* ctx, uint32_t n_states, const double* init, double tol, uint32_t mode,
                  uint32_t max_calls, double* evals, double* evecs, int* converged, double* traj_e, double* traj_t,
                  int traj_cap, int* n_traj) {
  if (!ctx || !init || !evals) return ctx ? fail(ctx, SHCI_ERR_ARG, "bad davidson args") : SHCI_ERR_ARG;
  if (!ctx->ts || !ctx->built)
    return davidson_full(ctx, n_states, init, tol, mode, max_calls, evals, evecs, converged, traj_e, traj_t, traj_cap,
                         n_traj);
  const uint64_t n = ctx->n_red, nE = ctx->n;
  std::vector<double> ie((size_t)n_states * nE), ee(evecs ? (size_t)n_states * nE : 0);
  for (uint32_t s = 0; s < n_states; s++) ts_expand(ctx, init + (size_t)s * n, ie.data() + (size_t)s * nE);
  CHECK_RC(davidson_full(ctx, n_states, ie.data(), tol, mode, max_calls, evals, evecs ? ee.data() : nullptr, converged,
                         traj_e, traj_t, traj_cap, n_traj));
  if (evecs)
    for (uint32_t s = 0; s < n_states; s++) ts_contract(ctx, ee.data() + (size_t)s * nE, evecs + (size_t)s * n);
  return SHCI_OK;
}

int shci_rdm(shci_ctx* ctx, uint32_t n_states, const double* coefs, double* rdm1, double* rdm2) {
  if (!ctx || !coefs || n_states == 0) return SHCI_ERR_ARG;
  if (!ctx->ts || !ctx->built) return rdm_full(ctx, n_states, coefs, rdm1, rdm2);
  const uint64_t n = ctx->n_red, nE = ctx->n;
  std::vector<double> ce((size_t)n_states * nE);  // unpack_time_sym, base_system.h:108-122
  for (uint32_t s = 0; s < n_states; s++) ts_expand(ctx, coefs + (size_t)s * n, ce.data() + (size_t)s * nE);
  return rdm_full(ctx, n_states, ce.data(), rdm1, rdm2);
}

int shci_memory(shci_ctx* ctx, uint64_t* bytes) {
  if (!ctx || !bytes) return SHCI_ERR_ARG;
  *bytes = ctx->mem_in_use;
  return SHCI_OK;
}

}  // extern "C"
