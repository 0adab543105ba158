// batch.hip — batched GROUP BY: many independent queries in one persistent
// work-queue pipeline, replayed as a hipGraph, with an autotuner.
//
// No reference counterpart (the reference runs one execute_groupby per query,
// src/exec/groupby.rs:295).  Per query the semantics are those of
// bolt_groupby, whose device pieces (key encoding, accumulator words, exact
// 128-bit integer atomics, the global open-addressing table, finalisation)
// are shared through groupby_dev.h.
//
// A run is seven kernels and two small copies, captured once:
//   reset    zero the work counters, the sketches and the output validity;
//   sketch   persistent work queue over (query, row tile) items: a 2^10
//            register HyperLogLog per query;
//   plan     one workgroup: per query the group estimate, hash-table size
//            (power of two, >= 2.5x the estimate), LDS table size, arena
//            offset and extraction items (block-wide scans) — on the device,
//            so sizing needs no host round trip;
//   init     fills the arena's tables only when the plan differs from the
//            one the tables were left clean for;
//   aggregate persistent work queue over (query, row tile) items, largest
//            queries first: each item aggregates its tile in an LDS table;
//            rows whose key finds no LDS slot (the tail of a high-cardinality
//            query) update the query's HBM table directly; the LDS table is
//            then merged into the HBM table;
//   extract  persistent over (query, slot tile) items: occupied slots are
//            finalised into the query's output window and reset to empty, so
//            the next run on the same plan needs no table initialisation;
//   finish   records the plan the tables are now clean for;
//   then the control block and per-query plans are copied to pinned host
//   memory.
// Every query's hash table, specials and outputs live in one HBM arena whose
// per-query windows start on 64-slot boundaries.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/bolt_batch.h"
#include "groupby_dev.h"

namespace bolt {

constexpr int kBHllBits = 10;
constexpr int kBHllRegs = 1 << kBHllBits;
constexpr int kBExTile = 4096;  // slots per extraction item
constexpr int kBPad = 64;       // window padding (2 special slots + alignment)
constexpr int kBAux = 256;      // threads of the sketch / extract kernels

struct BQuery {
  GPlan P;
  KeySpec ks;
  int64_t n;
  int64_t item0;  // first (sketch / aggregate) work item
};

struct BPlan {
  uint64_t cap;   // hash slots, power of two
  uint64_t off;   // window start in the arena (slots)
  uint64_t ext0;  // first extraction item
  int32_t S;      // LDS slots
  int32_t pad;
  unsigned long long n_groups;
  uint64_t prev_cap, prev_off;  // the plan the tables are clean for
};

struct BCtl {
  unsigned long long next[3];  // work counters: sketch, aggregate, extract
  unsigned long long need;     // arena slots the plan needs
  unsigned long long ext_items;
  uint32_t flags;              // 1 arena too small, 2 tables need initialising
  uint32_t planned;            // the tables are clean for plan[].prev_*
  uint32_t err[4];             // [0] probe bound exceeded, [2] integer overflow
};

struct BOut {
  OutCol key[2];
  OutCol agg[kMaxAggs];
};

__device__ __forceinline__ void bhll_add(uint32_t* regs, uint64_t key) {
  uint64_t h = mix64(key ^ 0x9E3779B97F4A7C15ull);
  uint32_t idx = (uint32_t)(h >> (64 - kBHllBits));
  uint64_t rest = (h << kBHllBits) | (1ull << (kBHllBits - 1));
  atomicMax(&regs[idx], (uint32_t)__clzll(rest) + 1);
}

__device__ __forceinline__ int dtype_size_dev(int dt) {
  return dt == BOLT_I64 || dt == BOLT_F64 ? 8 : dt == BOLT_DEC128 ? 16 : dt == BOLT_BOOL ? 1 : 4;
}

// last query q with first[q] <= item (first[] ascending)
template <typename F>
__device__ __forceinline__ int find_owner(int nq, long long item, F first) {
  int lo = 0, hi = nq - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if ((long long)first(mid) <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void bg_reset_kernel(BCtl* __restrict__ ctl, uint32_t* __restrict__ hll, size_t hll_words,
                                uint32_t* __restrict__ vbits, size_t vwords) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (size_t i = i0; i < hll_words; i += stride) hll[i] = 0;
  for (size_t i = i0; i < vwords; i += stride) vbits[i] = 0;
  if (i0 == 0) {
    ctl->next[0] = ctl->next[1] = ctl->next[2] = 0;
    ctl->flags = 0;
    for (int k = 0; k < 4; k++) ctl->err[k] = 0;
  }
}

__global__ __launch_bounds__(kBAux) void bg_sketch_kernel(const BQuery* __restrict__ qs, int nq, long long items,
                                                          int tile_rows, BCtl* __restrict__ ctl,
                                                          uint32_t* __restrict__ hll) {
  __shared__ uint32_t regs[kBHllRegs];
  __shared__ long long sh_item;
  __shared__ int sh_q;
  for (;;) {
    if (threadIdx.x == 0) {
      long long it = (long long)atomicAdd(&ctl->next[0], 1ull);
      sh_item = it;
      sh_q = it < items ? find_owner(nq, it, [&](int q) { return qs[q].item0; }) : 0;
    }
    for (int i = threadIdx.x; i < kBHllRegs; i += kBAux) regs[i] = 0;
    __syncthreads();
    const long long it = sh_item;
    if (it >= items) break;
    const BQuery& Q = qs[sh_q];
    const int64_t r0 = (it - Q.item0) * (int64_t)tile_rows;
    const int64_t r1 = min(Q.n, r0 + (int64_t)tile_rows);
    for (int64_t row = r0 + threadIdx.x; row < r1; row += kBAux) {
      uint64_t key;
      if (row_key(Q.ks, row, &key) == 1) bhll_add(regs, key);
    }
    __syncthreads();
    uint32_t* g = hll + (size_t)sh_q * kBHllRegs;
    for (int i = threadIdx.x; i < kBHllRegs; i += kBAux)
      if (regs[i]) atomicMax(&g[i], regs[i]);
    __syncthreads();
  }
}

// One workgroup of 1024 threads: query q is planned by thread q % 1024.
__global__ __launch_bounds__(1024) void bg_plan_kernel(const BQuery* __restrict__ qs, int nq,
                                                       const uint32_t* __restrict__ hll, BPlan* __restrict__ plan,
                                                       BCtl* __restrict__ ctl, double mult, int s_max,
                                                       unsigned long long arena_slots) {
  __shared__ unsigned long long sreg[1024], sext[1024];
  __shared__ unsigned long long carry, ecarry;
  __shared__ uint32_t changed;
  const int t = threadIdx.x;
  if (t == 0) {
    carry = 0;
    ecarry = 0;
    changed = ctl->planned ? 0u : 1u;
  }
  __syncthreads();
  for (int base = 0; base < nq; base += 1024) {
    const int q = base + t;
    unsigned long long region = 0, ext = 0, cap = 0;
    int S = 0;
    if (q < nq) {
      const uint32_t* r = hll + (size_t)q * kBHllRegs;
      double sum = 0;
      int zeros = 0;
      for (int i = 0; i < kBHllRegs; i++) {
        uint32_t v = r[i];
        sum += ldexp(1.0, -(int)v);
        zeros += v == 0;
      }
      const double m = kBHllRegs;
      double e = (0.7213 / (1.0 + 1.079 / m)) * m * m / sum;
      if (e <= 2.5 * m && zeros > 0) e = m * log(m / zeros);
      double want = e * mult + 16.0;
      const double lim = 2.0 * (double)qs[q].n + 64.0;  // never more groups than rows
      if (want > lim) want = lim;
      cap = 64;
      while ((double)cap < want) cap <<= 1;
      S = (int)min(cap, (unsigned long long)s_max);
      region = cap + kBPad;
      ext = (region + kBExTile - 1) / kBExTile;
    }
    sreg[t] = region;
    sext[t] = ext;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
      unsigned long long a = t >= d ? sreg[t - d] : 0, b = t >= d ? sext[t - d] : 0;
      __syncthreads();
      sreg[t] += a;
      sext[t] += b;
      __syncthreads();
    }
    if (q < nq) {
      BPlan& p = plan[q];
      const unsigned long long off = carry + sreg[t] - region;
      if (p.prev_cap != cap || p.prev_off != off) changed = 1u;
      p.cap = cap;
      p.off = off;
      p.S = S;
      p.ext0 = ecarry + sext[t] - ext;
      p.n_groups = 0;
    }
    __syncthreads();
    if (t == 0) {
      carry += sreg[1023];
      ecarry += sext[1023];
    }
    __syncthreads();
  }
  if (t == 0) {
    ctl->need = carry;
    ctl->ext_items = ecarry;
    uint32_t fl = 0;
    if (carry > arena_slots) fl |= 1u;
    if (changed) fl |= 2u;
    ctl->flags = fl;
  }
}

__global__ void bg_init_kernel(const BCtl* __restrict__ ctl, GPlan P, unsigned long long* __restrict__ keys,
                               unsigned long long* __restrict__ words, unsigned long long arena_slots,
                               uint32_t* __restrict__ specials, int nq) {
  if ((ctl->flags & 3u) != 2u) return;
  const unsigned long long need = ctl->need;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < need; i += stride) {
    keys[i] = kEmpty;
    for (int w = 0; w < P.words; w++) words[(size_t)w * arena_slots + i] = word_identity(P, w);
    if (i < (size_t)nq * 2) specials[i] = 0;
  }
}

constexpr int kBRows = 4;  // rows per lane per iteration of the aggregation pass

__global__ __launch_bounds__(1024) void bg_agg_kernel(const BQuery* __restrict__ qs, int nq, long long items,
                                                      int tile_rows, const BPlan* __restrict__ plan,
                                                      BCtl* __restrict__ ctl, unsigned long long* __restrict__ keys,
                                                      unsigned long long* __restrict__ words,
                                                      unsigned long long arena_slots, uint32_t* __restrict__ specials) {
  extern __shared__ unsigned long long sm[];
  __shared__ long long sh_item;
  __shared__ int sh_q;
  if (ctl->flags & 1u) return;
  const int nt = blockDim.x;
  for (;;) {
    if (threadIdx.x == 0) {
      long long it = (long long)atomicAdd(&ctl->next[1], 1ull);
      sh_item = it;
      sh_q = it < items ? find_owner(nq, it, [&](int q) { return qs[q].item0; }) : 0;
    }
    __syncthreads();
    const long long it = sh_item;
    if (it >= items) break;
    const int q = __builtin_amdgcn_readfirstlane(sh_q);
    const BQuery& Q = qs[q];
    const GPlan& P = Q.P;
    const BPlan pl = plan[q];
    const int S = pl.S, SS = S + 2, W = P.words;
    unsigned long long* skeys = sm;
    unsigned long long* swords = sm + SS;
    uint32_t* sused = reinterpret_cast<uint32_t*>(swords + (size_t)W * SS);
    for (int i = threadIdx.x; i < SS; i += nt) skeys[i] = kEmpty;
    for (int w = 0; w < W; w++) {
      const unsigned long long id = word_identity(P, w);
      for (int i = threadIdx.x; i < SS; i += nt) swords[(size_t)w * SS + i] = id;
    }
    if (threadIdx.x < 3) sused[threadIdx.x] = 0;  // [0..1] specials used, [2] occupied slots
    __syncthreads();
    GTable t;
    t.keys = keys + pl.off;
    t.words = words + pl.off;
    t.used_special = specials + 2 * q;
    t.err = ctl->err;
    t.cap = pl.cap;
    const int64_t r0 = (it - Q.item0) * (int64_t)tile_rows;
    const int64_t r1 = min(Q.n, r0 + (int64_t)tile_rows);
    const uint32_t smask = (uint32_t)S - 1;
    const int npf = P.n < 4 ? P.n : 4;
    // kBRows rows per lane, nt apart: keys and the first aggregates' inputs
    // of all of them are loaded, then their home slots read, before the
    // first update, so a lane has several rows' HBM reads in flight
    for (int64_t base = r0 + threadIdx.x; base < r1; base += (int64_t)nt * kBRows) {
      uint64_t key[kBRows], pw0[kBRows], pw1[kBRows], pw2[kBRows], pw3[kBRows];
      int kr[kBRows];
#pragma unroll
      for (int j = 0; j < kBRows; j++) {
        const int64_t row = base + (int64_t)j * nt;
        key[j] = 0;
        kr[j] = row < r1 ? row_key(Q.ks, row, &key[j]) : 0;
        const int64_t rc = row < r1 ? row : r0;
        pw0[j] = npf > 0 ? pf_word(P.a[0], rc) : 0;
        pw1[j] = npf > 1 ? pf_word(P.a[1], rc) : 0;
        pw2[j] = npf > 2 ? pf_word(P.a[2], rc) : 0;
        pw3[j] = npf > 3 ? pf_word(P.a[3], rc) : 0;
      }
      uint32_t home[kBRows];
      unsigned long long cur[kBRows];
#pragma unroll
      for (int j = 0; j < kBRows; j++) {
        home[j] = (uint32_t)(mix64(key[j]) >> 32) & smask;
        cur[j] = skeys[home[j]];
      }
#pragma unroll
      for (int j = 0; j < kBRows; j++) {
        if (kr[j] == 0) continue;
        const int64_t row = base + (int64_t)j * nt;
        int slot = -1;
        if (kr[j] == 2) {
          slot = S;
          sused[0] = 1;
        } else if (key[j] == kEmpty) {
          slot = S + 1;
          sused[1] = 1;
        } else if (cur[j] == key[j]) {
          slot = (int)home[j];
        } else {
          // the shared probe (power-of-two S here): kLdsProbes slots while the
          // table has room, 2 once 15/16 of it is taken
          slot = gb_lds_probe(skeys, S, key[j], home[j], gb_lds_probes(S, &sused[2]), &sused[2]);
        }
        if (slot >= 0) {
          update_slot_pf(P, swords, SS, (uint32_t)slot, row, pw0[j], pw1[j], pw2[j], pw3[j]);
        } else {
          const int64_t g = gslot(t, 0, key[j]);
          if (g >= 0) update_slot_pf(P, t.words, arena_slots, (uint32_t)g, row, pw0[j], pw1[j], pw2[j], pw3[j]);
        }
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < SS; i += nt) {
      bool used;
      int tag = 0;
      uint64_t key = 0;
      if (i < S) {
        key = skeys[i];
        used = key != kEmpty;
      } else {
        used = sused[i - S] != 0;
        tag = i - S + 1;
      }
      if (!used) continue;
      const int64_t g = gslot(t, tag, key);
      if (g >= 0) merge_slot(P, t.words, arena_slots, (uint32_t)g, swords + i, SS);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBAux) void bg_extract_kernel(const BQuery* __restrict__ qs, int nq,
                                                           BPlan* __restrict__ plan, BCtl* __restrict__ ctl,
                                                           unsigned long long* __restrict__ keys,
                                                           unsigned long long* __restrict__ words,
                                                           unsigned long long arena_slots,
                                                           uint32_t* __restrict__ specials, BOut o) {
  __shared__ long long sh_item;
  __shared__ int sh_q;
  if (ctl->flags & 1u) return;
  const long long items = (long long)ctl->ext_items;
  for (;;) {
    if (threadIdx.x == 0) {
      long long it = (long long)atomicAdd(&ctl->next[2], 1ull);
      sh_item = it;
      sh_q = it < items ? find_owner(nq, it, [&](int q) { return plan[q].ext0; }) : 0;
    }
    __syncthreads();
    const long long it = sh_item;
    if (it >= items) break;
    const int q = sh_q;
    const BQuery& Q = qs[q];
    const GPlan& P = Q.P;
    const BPlan pl = plan[q];
    GTable t;
    t.keys = keys + pl.off;
    t.words = words + pl.off;
    t.used_special = specials + 2 * q;
    t.err = ctl->err;
    t.cap = pl.cap;
    const uint64_t lim = pl.cap + 2;
    const uint64_t s0 = (uint64_t)(it - (long long)pl.ext0) * kBExTile;
    for (int j = 0; j < kBExTile; j += kBAux) {
      const uint64_t s = s0 + j + threadIdx.x;
      bool occ = false;
      uint64_t packed = 0;
      if (s < pl.cap) {
        packed = t.keys[s];
        occ = packed != kEmpty;
      } else if (s < lim) {
        occ = t.used_special[s - pl.cap] != 0;
        packed = s == pl.cap ? 0x7FFFFFFFFFFFFFFFull : kEmpty;
      }
      const unsigned long long pos = wave_append(occ, &plan[q].n_groups);
      if (!occ) continue;
      const bool null_group = s == pl.cap;
      OutCol k0 = o.key[0];
      k0.data += pl.off * (uint64_t)dtype_size_dev(k0.dtype);
      if (k0.valid) k0.valid += pl.off / 32;
      if (Q.ks.n_keys == 1) {
        write_key(k0, pos, null_group ? 0 : packed);
        if (k0.valid && !null_group) set_valid(k0.valid, pos);
      } else {
        OutCol k1 = o.key[1];
        k1.data += pl.off * (uint64_t)dtype_size_dev(k1.dtype);
        write_key(k0, pos, packed >> 32);
        write_key(k1, pos, packed & 0xFFFFFFFFull);
      }
      for (int a = 0; a < P.n; a++) {
        OutCol c = o.agg[a];
        c.data += pl.off * (uint64_t)dtype_size_dev(c.dtype);
        if (c.valid) c.valid += pl.off / 32;
        finalize_to(t, P.a[a], c, arena_slots, s, pos);
      }
      // leave the slot empty for the next run on this plan
      if (s < pl.cap) t.keys[s] = kEmpty;
      else t.used_special[s - pl.cap] = 0;
      for (int w = 0; w < P.words; w++) t.words[(size_t)w * arena_slots + s] = word_identity(P, w);
    }
    __syncthreads();
  }
}

__global__ void bg_finish_kernel(BPlan* __restrict__ plan, int nq, BCtl* __restrict__ ctl) {
  if (ctl->flags & 1u) return;
  for (int q = threadIdx.x; q < nq; q += blockDim.x) {
    plan[q].prev_cap = plan[q].cap;
    plan[q].prev_off = plan[q].off;
  }
  if (threadIdx.x == 0) ctl->planned = 1u;
}

// ---- autotuner state ------------------------------------------------------------

std::mutex g_tune_mu;
std::map<std::string, bolt_gb_tune_t> g_tuned;
bool g_tune_loaded = false;

void tune_load_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char* path = getenv("BOLT_TUNE_FILE");
  if (!path) return;
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string key;
    bolt_gb_tune_t t{};
    if (ls >> key >> t.block >> t.tile_rows >> t.lds_slots >> t.grid) g_tuned[key] = t;
  }
}

void tune_save_locked() {
  const char* path = getenv("BOLT_TUNE_FILE");
  if (!path) return;
  std::string tmp = std::string(path) + ".tmp";
  {
    std::ofstream out(tmp);
    for (auto& kv : g_tuned)
      out << kv.first << ' ' << kv.second.block << ' ' << kv.second.tile_rows << ' ' << kv.second.lds_slots << ' '
          << kv.second.grid << '\n';
  }
  std::rename(tmp.c_str(), path);
}

int ilog2_bucket(double x) { return x < 1 ? 0 : (int)std::floor(std::log2(x)); }

}  // namespace bolt

using namespace bolt;

struct bolt_gb_batch {
  hipStream_t s = nullptr;
  hipStream_t own = nullptr;  // private stream when the caller passes the NULL stream (not capturable)
  int nq = 0, n_keys = 0, n_aggs = 0;
  std::vector<int> pos_of;  // caller query -> position in the (size-sorted) batch
  std::vector<int64_t> rows;
  long long items = 0;
  GPlan P0;
  int W = 0;
  int32_t out_dt[kMaxAggs];
  int32_t key_dt[2] = {0, 0};
  bool key_nullable = false;
  bool agg_nullable[kMaxAggs];
  std::string shape_key;
  bolt_gb_tune_t tune{};
  double mult = 2.5;
  int lds_limit = 0;
  int n_cus = 256;
  // device state
  DevBuf<BQuery> d_q;
  DevBuf<BPlan> d_plan;
  DevBuf<BCtl> d_ctl;
  DevBuf<uint32_t> d_hll, d_spec;
  DevBuf<unsigned long long> d_keys, d_words;
  std::vector<void*> d_out;  // n_keys + n_aggs output columns
  DevBuf<uint32_t> d_vbits;
  size_t vwords_per_col = 0;
  int n_valid_cols = 0;
  std::vector<int> valid_slot;  // output column -> index of its bitmap (-1 none)
  unsigned long long arena = 0;
  BCtl* h_ctl = nullptr;
  BPlan* h_plan = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool have_result = false;

  ~bolt_gb_batch() {
    drop_graph();
    free_arena();
    d_q.reset();
    d_plan.reset();
    d_ctl.reset();
    d_hll.reset();
    d_spec.reset();
    if (h_ctl) (void)hipHostFree(h_ctl);
    if (h_plan) (void)hipHostFree(h_plan);
    if (own) {
      (void)hipStreamSynchronize(own);
      cache_forget_stream(own);
      (void)hipStreamDestroy(own);
    }
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  void free_arena() {
    for (void* p : d_out)
      if (p) dev_free(p, s);
    d_out.clear();
    d_keys.reset();
    d_words.reset();
    d_vbits.reset();
    arena = 0;
  }
  int lds_bytes(int slots) const { return (slots + 2) * 8 * (1 + W) + 16; }
  int s_max() const {
    int S = tune.lds_slots;
    while (S > 64 && lds_bytes(S) > lds_limit) S >>= 1;
    return S;
  }
  unsigned agg_grid() const { return tune.grid > 0 ? (unsigned)tune.grid : (unsigned)(2 * n_cus); }

  int alloc_arena(unsigned long long slots) {
    free_arena();
    slots = (slots + 63) / 64 * 64;
    if (slots < 64) slots = 64;
    arena = slots;
    BOLT_TRY(d_keys.alloc(slots, s));
    BOLT_TRY(d_words.alloc(slots * (size_t)(W > 0 ? W : 1), s));
    d_out.assign(n_keys + n_aggs, nullptr);
    for (int c = 0; c < n_keys + n_aggs; c++) {
      int dt = c < n_keys ? key_dt[c] : out_dt[c - n_keys];
      BOLT_TRY(dev_alloc(&d_out[c], (size_t)slots * dtype_size(dt), s));
    }
    vwords_per_col = slots / 32;
    BOLT_TRY(d_vbits.alloc(vwords_per_col * (size_t)(n_valid_cols > 0 ? n_valid_cols : 1), s));
    // tables start out clean for no plan: the first run initialises them
    BCtl zero{};
    BOLT_HIP(hipMemcpyAsync(d_ctl.ptr, &zero, sizeof zero, hipMemcpyHostToDevice, s));
    return BOLT_OK;
  }

  BOut out_desc() const {
    BOut o;
    memset(&o, 0, sizeof o);
    for (int c = 0; c < n_keys + n_aggs; c++) {
      OutCol oc;
      oc.data = static_cast<uint8_t*>(d_out[c]);
      oc.valid = valid_slot[c] >= 0 ? d_vbits.ptr + (size_t)valid_slot[c] * vwords_per_col : nullptr;
      oc.dtype = c < n_keys ? key_dt[c] : out_dt[c - n_keys];
      oc.pad = 0;
      if (c < n_keys) o.key[c] = oc;
      else o.agg[c - n_keys] = oc;
    }
    return o;
  }

  // Enqueue one run on the stream (captured into the graph).
  int enqueue(bool with_copies) {
    const size_t hll_words = (size_t)nq * kBHllRegs;
    const size_t vwords = vwords_per_col * (size_t)n_valid_cols;
    bg_reset_kernel<<<grid_for((int64_t)std::max(hll_words, vwords), 256, 4, 1024), 256, 0, s>>>(
        d_ctl.ptr, d_hll.ptr, hll_words, d_vbits.ptr, vwords);
    bg_sketch_kernel<<<4 * n_cus, kBAux, 0, s>>>(d_q.ptr, nq, items, tune.tile_rows, d_ctl.ptr, d_hll.ptr);
    bg_plan_kernel<<<1, 1024, 0, s>>>(d_q.ptr, nq, d_hll.ptr, d_plan.ptr, d_ctl.ptr, mult, s_max(), arena);
    bg_init_kernel<<<4 * n_cus, 256, 0, s>>>(d_ctl.ptr, P0, d_keys.ptr, d_words.ptr, arena, d_spec.ptr, nq);
    const int S = s_max();
    bg_agg_kernel<<<agg_grid(), tune.block, lds_bytes(S), s>>>(d_q.ptr, nq, items, tune.tile_rows, d_plan.ptr,
                                                                 d_ctl.ptr, d_keys.ptr, d_words.ptr, arena,
                                                                 d_spec.ptr);
    bg_extract_kernel<<<4 * n_cus, kBAux, 0, s>>>(d_q.ptr, nq, d_plan.ptr, d_ctl.ptr, d_keys.ptr, d_words.ptr, arena,
                                                  d_spec.ptr, out_desc());
    bg_finish_kernel<<<1, 1024, 0, s>>>(d_plan.ptr, nq, d_ctl.ptr);
    BOLT_HIP(hipGetLastError());
    if (with_copies) {
      BOLT_HIP(hipMemcpyAsync(h_ctl, d_ctl.ptr, sizeof(BCtl), hipMemcpyDeviceToHost, s));
      BOLT_HIP(hipMemcpyAsync(h_plan, d_plan.ptr, sizeof(BPlan) * nq, hipMemcpyDeviceToHost, s));
    }
    return BOLT_OK;
  }

  int capture() {
    drop_graph();
    BOLT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(bg_agg_kernel),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, lds_limit));
    BOLT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int st = enqueue(true);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (st != BOLT_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (e != hipSuccess) return set_error(BOLT_ERR_DEVICE, "batch graph capture failed: %s", hipGetErrorString(e));
    graph = g;
    BOLT_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    return BOLT_OK;
  }

  // One run; re-plans (arena growth, larger tables) until the flags are clear.
  int run() {
    have_result = false;
    for (int attempt = 0; attempt < 6; attempt++) {
      if (!exec) BOLT_TRY(capture());
      BOLT_HIP(hipGraphLaunch(exec, s));
      BOLT_HIP(hipStreamSynchronize(s));
      if (h_ctl->flags & 1u) {  // the plan needs a larger arena
        BOLT_TRY(alloc_arena(h_ctl->need + h_ctl->need / 4));
        drop_graph();
        continue;
      }
      if (h_ctl->err[0]) {  // a table filled up: estimates were low
        mult *= 4;
        drop_graph();
        continue;
      }
      if (h_ctl->err[2] & 1u) return set_error(BOLT_ERR_TYPE, "SUM(integer) overflow: accumulator exceeds i64 range");
      if (h_ctl->err[2] & 2u)
        return set_error(BOLT_ERR_TYPE, "SUM(Decimal128) precision overflow: accumulator exceeds i128 range");
      have_result = true;
      return BOLT_OK;
    }
    return set_error(BOLT_ERR_MEMORY, "batched GROUP BY could not size its hash tables");
  }
};

static void default_tune(bolt_gb_tune_t* t) {
  t->block = 1024;
  t->tile_rows = 65536;
  t->lds_slots = 2048;
  t->grid = 0;
}

static bool tune_ok(const bolt_gb_tune_t& t) {
  auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
  return (t.block == 256 || t.block == 512 || t.block == 1024) && pow2(t.tile_rows) && t.tile_rows >= 4096 &&
         t.tile_rows <= (1 << 20) && pow2(t.lds_slots) && t.lds_slots >= 64 && t.lds_slots <= 16384 && t.grid >= 0;
}

extern "C" int bolt_gb_batch_create(bolt_stream_t stream, const bolt_column_t* keys, int32_t n_keys,
                                    const bolt_column_t* inputs, int32_t n_inputs, int32_t n_queries,
                                    const bolt_agg_t* aggs, int32_t n_aggs, const bolt_gb_tune_t* tune,
                                    bolt_gb_batch_t** out) {
  BOLT_CHECK_DEVICE();
  clear_error();
  if (!out) return set_error(BOLT_ERR_PLAN, "bolt_gb_batch_create: out is NULL");
  *out = nullptr;
  if (n_queries < 1) return set_error(BOLT_ERR_PLAN, "bolt_gb_batch_create: no queries");
  if (n_keys < 1 || n_keys > 2 || !keys) return set_error(BOLT_ERR_PLAN, "bolt_gb_batch_create: 1 or 2 key columns required");
  if (n_inputs < 0 || (n_inputs > 0 && !inputs)) return set_error(BOLT_ERR_PLAN, "bolt_gb_batch_create: bad inputs");
  auto b = new bolt_gb_batch();
  std::unique_ptr<bolt_gb_batch> guard(b);
  b->s = (hipStream_t)stream;
  if (!b->s) {
    BOLT_HIP(hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking));
    b->s = b->own;
  }
  b->nq = n_queries;
  b->n_keys = n_keys;
  b->n_aggs = n_aggs;
  int dev = 0;
  BOLT_HIP(hipGetDevice(&dev));
  BOLT_HIP(hipDeviceGetAttribute(&b->n_cus, hipDeviceAttributeMultiprocessorCount, dev));
  int lds = 0;
  BOLT_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  b->lds_limit = std::min(lds, 96 * 1024);
  // table size = estimate x mult (tests lower it to exercise the re-plan path)
  if (const char* m = getenv("BOLT_BATCH_EST_MULT")) {
    double v = atof(m);
    if (v > 0) b->mult = v;
  }

  // per-query plans and key specs, checked against query 0
  std::vector<BQuery> hq(n_queries);
  std::vector<int64_t> n_of(n_queries);
  for (int q = 0; q < n_queries; q++) {
    const bolt_column_t* kq = keys + (size_t)q * n_keys;
    const bolt_column_t* iq = inputs ? inputs + (size_t)q * n_inputs : nullptr;
    const int64_t n = kq[0].length;
    if (n > 0xFFFFFFFFll) return set_error(BOLT_ERR_PLAN, "query %d: %lld rows exceeds the u32 row-id limit", q, (long long)n);
    KeySpec& ks = hq[q].ks;
    memset(&ks, 0, sizeof ks);
    ks.n_keys = n_keys;
    for (int k = 0; k < n_keys; k++) {
      const int dt = kq[k].dtype;
      if (!(dt == BOLT_I32 || dt == BOLT_I64 || dt == BOLT_F32 || dt == BOLT_F64 || dt == BOLT_DATE32))
        return set_error(BOLT_ERR_TYPE, "GROUP BY key dtype %s not supported", dtype_name(dt));
      if (kq[k].length != n) return set_error(BOLT_ERR_PLAN, "query %d: key columns differ in length", q);
      if (n_keys == 2 && dtype_size(dt) != 4)
        return set_error(BOLT_ERR_UNSUPPORTED, "multi-column GROUP BY with > 64 bits of key width not yet supported");
      if (q == 0) b->key_dt[k] = dt;
      else if (dt != b->key_dt[k]) return set_error(BOLT_ERR_PLAN, "query %d: key dtype differs from query 0", q);
    }
    ks.d0 = kq[0].data;
    ks.v0 = kq[0].validity;
    ks.t0 = kq[0].dtype;
    if (n_keys == 2) {
      ks.d1 = kq[1].data;
      ks.v1 = kq[1].validity;
      ks.t1 = kq[1].dtype;
    }
    ks.null_group = (n_keys == 1 && kq[0].validity != nullptr) ? 1 : 0;
    if (q == 0) b->key_nullable = ks.null_group != 0;
    else if ((ks.null_group != 0) != b->key_nullable)
      return set_error(BOLT_ERR_PLAN, "query %d: key nullability differs from query 0", q);
    int32_t odt[kMaxAggs];
    BOLT_TRY(make_gplan(aggs, n_aggs, iq, n_inputs, n, &hq[q].P, odt));
    if (q == 0) {
      b->P0 = hq[0].P;
      b->W = hq[0].P.words;
      memcpy(b->out_dt, odt, sizeof odt);
    } else {
      for (int a = 0; a < n_aggs; a++)
        if (odt[a] != b->out_dt[a] || hq[q].P.a[a].kind != b->P0.a[a].kind || hq[q].P.a[a].nw != b->P0.a[a].nw)
          return set_error(BOLT_ERR_PLAN, "query %d: aggregate %d differs in dtype or nullability from query 0", q, a);
    }
    for (int a = 0; a < n_aggs; a++)
      if (hq[q].P.a[a].kind == K_VAR)
        return set_error(BOLT_ERR_UNSUPPORTED, "batched GROUP BY: VAR / STDDEV need the two-pass per-query path");
    hq[q].n = n;
    n_of[q] = n;
  }
  for (int a = 0; a < n_aggs; a++) {
    const GAgg& g = b->P0.a[a];
    b->agg_nullable[a] = (g.nw >= 0 && g.kind != K_COUNT);
  }
  // output validity bitmaps: key 0 when it has a NULL group, nullable aggregates
  b->valid_slot.assign(n_keys + n_aggs, -1);
  if (b->key_nullable) b->valid_slot[0] = b->n_valid_cols++;
  for (int a = 0; a < n_aggs; a++)
    if (b->agg_nullable[a]) b->valid_slot[n_keys + a] = b->n_valid_cols++;

  // parameters: caller's, else the tuned value for this shape class, else defaults
  int64_t total_rows = 0;
  for (int64_t n : n_of) total_rows += n;
  {
    std::ostringstream k;
    k << "gb:q" << ilog2_bucket(n_queries) << ":r" << ilog2_bucket((double)total_rows) << ":w" << b->W << ":k"
      << b->key_dt[0];
    b->shape_key = k.str();
  }
  default_tune(&b->tune);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_locked();
    auto it = g_tuned.find(b->shape_key);
    if (it != g_tuned.end()) b->tune = it->second;
  }
  if (tune) {
    if (tune->block) b->tune.block = tune->block;
    if (tune->tile_rows) b->tune.tile_rows = tune->tile_rows;
    if (tune->lds_slots) b->tune.lds_slots = tune->lds_slots;
    if (tune->grid) b->tune.grid = tune->grid;
  }
  if (!tune_ok(b->tune)) return set_error(BOLT_ERR_PLAN, "bolt_gb_batch_create: invalid tuning parameters");

  // largest queries first, so the long items start early
  std::vector<int> order(n_queries);
  for (int q = 0; q < n_queries; q++) order[q] = q;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n_of[x] > n_of[y]; });
  b->pos_of.assign(n_queries, 0);
  std::vector<BQuery> sq(n_queries);
  for (int p = 0; p < n_queries; p++) {
    sq[p] = hq[order[p]];
    b->pos_of[order[p]] = p;
  }
  BOLT_TRY(b->d_q.alloc(n_queries, b->s));
  BOLT_TRY(b->d_plan.alloc(n_queries, b->s));
  BOLT_TRY(b->d_ctl.alloc(1, b->s));
  BOLT_TRY(b->d_hll.alloc((size_t)n_queries * kBHllRegs, b->s));
  BOLT_TRY(b->d_spec.alloc((size_t)n_queries * 2, b->s));
  BOLT_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->h_ctl), sizeof(BCtl), hipHostMallocDefault));
  BOLT_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->h_plan), sizeof(BPlan) * n_queries, hipHostMallocDefault));
  b->rows = n_of;
  // upload the (sorted) query table with item offsets
  auto layout = [&](bolt_gb_batch* bb) -> int {
    long long it = 0;
    for (int p = 0; p < bb->nq; p++) {
      sq[p].item0 = it;
      it += (sq[p].n + bb->tune.tile_rows - 1) / bb->tune.tile_rows;
    }
    bb->items = it;
    BOLT_HIP(hipMemcpyAsync(bb->d_q.ptr, sq.data(), sizeof(BQuery) * bb->nq, hipMemcpyHostToDevice, bb->s));
    return BOLT_OK;
  };
  BOLT_TRY(layout(b));
  BPlan zp;
  memset(&zp, 0, sizeof zp);
  std::vector<BPlan> zplan(n_queries, zp);
  BOLT_HIP(hipMemcpyAsync(b->d_plan.ptr, zplan.data(), sizeof(BPlan) * n_queries, hipMemcpyHostToDevice, b->s));
  BCtl zc{};
  BOLT_HIP(hipMemcpyAsync(b->d_ctl.ptr, &zc, sizeof zc, hipMemcpyHostToDevice, b->s));
  // first sizing: one uncaptured run against an empty arena reports the need
  BOLT_TRY(b->alloc_arena(64));
  BOLT_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(bg_agg_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               b->lds_limit));
  BOLT_TRY(b->enqueue(true));
  BOLT_HIP(hipStreamSynchronize(b->s));
  BOLT_TRY(b->alloc_arena(b->h_ctl->need + b->h_ctl->need / 4));
  // keep the item layout with the batch (the autotuner re-lays it out)
  b->rows.assign(n_queries, 0);
  for (int p = 0; p < n_queries; p++) b->rows[p] = sq[p].n;
  guard.release();
  *out = b;
  return BOLT_OK;
}

// Re-lay the work items for the batch's current tile size (queries already
// uploaded in sorted order; only item0 changes).
static int relayout(bolt_gb_batch* b) {
  std::vector<BQuery> sq(b->nq);
  BOLT_HIP(hipMemcpyAsync(sq.data(), b->d_q.ptr, sizeof(BQuery) * b->nq, hipMemcpyDeviceToHost, b->s));
  BOLT_HIP(hipStreamSynchronize(b->s));
  long long it = 0;
  for (int p = 0; p < b->nq; p++) {
    sq[p].item0 = it;
    it += (sq[p].n + b->tune.tile_rows - 1) / b->tune.tile_rows;
  }
  b->items = it;
  BOLT_HIP(hipMemcpyAsync(b->d_q.ptr, sq.data(), sizeof(BQuery) * b->nq, hipMemcpyHostToDevice, b->s));
  BOLT_HIP(hipStreamSynchronize(b->s));
  return BOLT_OK;
}

extern "C" int bolt_gb_batch_run(bolt_gb_batch_t* b) {
  BOLT_CHECK_DEVICE();
  clear_error();
  if (!b) return set_error(BOLT_ERR_PLAN, "bolt_gb_batch_run: batch is NULL");
  return b->run();
}

extern "C" int bolt_gb_batch_num_groups(const bolt_gb_batch_t* b, int32_t q, int64_t* n_groups) {
  if (!b || q < 0 || q >= b->nq) return set_error(BOLT_ERR_PLAN, "bolt_gb_batch_num_groups: bad query index");
  if (!b->have_result) return set_error(BOLT_ERR_PLAN, "bolt_gb_batch_num_groups: the batch has no result");
  *n_groups = (int64_t)b->h_plan[b->pos_of[q]].n_groups;
  return BOLT_OK;
}

extern "C" int bolt_gb_batch_column(const bolt_gb_batch_t* b, int32_t q, int32_t i, bolt_column_t* col) {
  if (!b || q < 0 || q >= b->nq) return set_error(BOLT_ERR_PLAN, "bolt_gb_batch_column: bad query index");
  if (i < 0 || i >= b->n_keys + b->n_aggs) return set_error(BOLT_ERR_PLAN, "bolt_gb_batch_column: bad column index");
  if (!b->have_result) return set_error(BOLT_ERR_PLAN, "bolt_gb_batch_column: the batch has no result");
  const BPlan& p = b->h_plan[b->pos_of[q]];
  const int dt = i < b->n_keys ? b->key_dt[i] : b->out_dt[i - b->n_keys];
  col->data = static_cast<uint8_t*>(b->d_out[i]) + p.off * (uint64_t)dtype_size(dt);
  col->validity = b->valid_slot[i] >= 0
                      ? reinterpret_cast<const uint8_t*>(b->d_vbits.ptr + (size_t)b->valid_slot[i] * b->vwords_per_col) +
                            p.off / 8
                      : nullptr;
  col->length = (int64_t)p.n_groups;
  col->dtype = dt;
  col->scale = 0;
  return BOLT_OK;
}

extern "C" int bolt_gb_batch_get_tune(const bolt_gb_batch_t* b, bolt_gb_tune_t* tune, int64_t* arena_bytes) {
  if (!b) return set_error(BOLT_ERR_PLAN, "bolt_gb_batch_get_tune: batch is NULL");
  if (tune) {
    *tune = b->tune;
    tune->lds_slots = b->s_max();
    if (tune->grid == 0) tune->grid = (int32_t)b->agg_grid();
  }
  if (arena_bytes) {
    size_t per_slot = 8 * (1 + (size_t)b->W);
    for (int c = 0; c < b->n_keys + b->n_aggs; c++)
      per_slot += dtype_size(c < b->n_keys ? b->key_dt[c] : b->out_dt[c - b->n_keys]);
    *arena_bytes = (int64_t)(b->arena * per_slot + b->arena / 8 * b->n_valid_cols);
  }
  return BOLT_OK;
}

extern "C" void bolt_gb_batch_release(bolt_gb_batch_t* b) { delete b; }

extern "C" int bolt_gb_batch_autotune(bolt_gb_batch_t* b, int32_t reps, bolt_gb_tune_t* best, float* best_ms) {
  BOLT_CHECK_DEVICE();
  clear_error();
  if (!b) return set_error(BOLT_ERR_PLAN, "bolt_gb_batch_autotune: batch is NULL");
  if (reps < 1) reps = 1;
  const bolt_gb_tune_t start = b->tune;
  std::vector<bolt_gb_tune_t> cands;
  for (int block : {256, 512, 1024})
    for (int tile : {16384, 65536, 262144})
      for (int lds : {512, 1024, 2048, 4096}) {
        bolt_gb_tune_t t{block, tile, lds, 0};
        if (b->lds_bytes(lds) > b->lds_limit) continue;
        cands.push_back(t);
      }
  hipEvent_t e0, e1;
  BOLT_HIP(hipEventCreate(&e0));
  BOLT_HIP(hipEventCreate(&e1));
  float best_t = 1e30f;
  bolt_gb_tune_t win = start;
  int st = BOLT_OK;
  for (const auto& c : cands) {
    b->tune = c;
    b->drop_graph();
    if ((st = relayout(b)) != BOLT_OK) break;
    if ((st = b->run()) != BOLT_OK) break;  // warm-up (and capture)
    float total = 0;
    for (int r = 0; r < reps && st == BOLT_OK; r++) {
      (void)hipEventRecord(e0, b->s);
      st = b->run();
      (void)hipEventRecord(e1, b->s);
      (void)hipEventSynchronize(e1);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      total += ms;
    }
    if (st != BOLT_OK) break;
    const float mean = total / reps;
    if (mean < best_t) {
      best_t = mean;
      win = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->tune = win;
  b->drop_graph();
  int st2 = relayout(b);
  if (st != BOLT_OK) return st;
  BOLT_TRY(st2);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_locked();
    g_tuned[b->shape_key] = win;
    tune_save_locked();
  }
  if (best) *best = win;
  if (best_ms) *best_ms = best_t;
  return BOLT_OK;
}
