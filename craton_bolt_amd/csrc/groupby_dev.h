// groupby_dev.h — device-side pieces of hash GROUP BY shared by the
// per-query operator (groupby.hip) and the batched form (batch.hip): key
// encoding, the accumulator word layout, LDS / HBM slot updates with exact
// integer atomics, global open-addressing table and finalisation.
#pragma once

#include "ops_common.h"
// after ops_common.h: gb_lds.h builds on device_acc.h
#include "gb_lds.h"

namespace bolt {

constexpr int kGlobalProbes = 1 << 16;

enum Kind { K_ISUM = 1, K_DSUM, K_FSUM, K_AVG, K_COUNT, K_MINMAX, K_VAR };

struct GAgg {
  const void* data;
  const uint8_t* validity;
  int32_t op;
  int32_t dtype;
  int32_t kind;
  int32_t w;   // first word
  int32_t nw;  // count word (-1 = none)
  int32_t pad;
};
struct GPlan {
  GAgg a[kMaxAggs];
  int32_t n;
  int32_t words;
};

struct KeySpec {
  const void* d0;
  const uint8_t* v0;
  const void* d1;
  const uint8_t* v1;
  int32_t t0, t1;
  int32_t n_keys;
  int32_t null_group;
};

// ---- keys ---------------------------------------------------------------------

__device__ __forceinline__ uint64_t key_bits(const void* d, int dt, int64_t row) {
  switch (dt) {
    case BOLT_I32: case BOLT_DATE32: return (uint64_t)static_cast<const uint32_t*>(d)[row];
    case BOLT_I64: return static_cast<const uint64_t*>(d)[row];
    // a Decimal128 key whose values all fit i64 (checked before the operator
    // runs) packs as its low word, like an Int64 key
    case BOLT_DEC128: return static_cast<const uint64_t*>(d)[2 * row];
    case BOLT_F32: return canon_f32_bits(static_cast<const uint32_t*>(d)[row]);
    case BOLT_F64: return canon_f64_bits(static_cast<const uint64_t*>(d)[row]);
  }
  return 0;
}

// 0 = drop the row, 1 = regular key in *key, 2 = the NULL group
__device__ __forceinline__ int row_key(const KeySpec& ks, int64_t row, uint64_t* key) {
  bool v0 = bit_valid(ks.v0, row);
  if (ks.n_keys == 1) {
    if (!v0) return ks.null_group ? 2 : 0;
    *key = key_bits(ks.d0, ks.t0, row);
    return 1;
  }
  if (!v0 || !bit_valid(ks.v1, row)) return 0;
  *key = (key_bits(ks.d0, ks.t0, row) << 32) | key_bits(ks.d1, ks.t1, row);
  return 1;
}

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

// Fold row `row`'s inputs into slot `slot` of a word array with `stride`
// words per field (LDS or global: the same atomics, address space inferred).
__device__ __forceinline__ void update_slot(const GPlan& P, unsigned long long* words, size_t stride, uint32_t slot,
                                            int64_t row) {
  for (int a = 0; a < P.n; a++) {
    const GAgg& g = P.a[a];
    unsigned long long* w = words + (size_t)g.w * stride + slot;
    if (g.op == BOLT_AGG_COUNT_STAR) { atomicAdd(w, 1ull); continue; }
    if (!bit_valid(g.validity, row)) continue;
    if (g.kind == K_COUNT) { atomicAdd(w, 1ull); continue; }
    if (g.nw >= 0) atomicAdd(words + (size_t)g.nw * stride + slot, 1ull);
    switch (g.kind) {
      case K_ISUM: {
        int64_t v = g.dtype == BOLT_I64 ? static_cast<const int64_t*>(g.data)[row]
                                        : (int64_t) static_cast<const int32_t*>(g.data)[row];
        atomic_add128(w, w + stride, (uint64_t)v, (uint64_t)(v >> 63));
        break;
      }
      case K_DSUM: {
        const uint64_t* p = static_cast<const uint64_t*>(g.data) + 2 * row;
        uint64_t lo = p[0], hi = p[1];
        atomic_add192(w, stride, lo, hi, (uint64_t)((int64_t)hi >> 63));
        break;
      }
      case K_FSUM:
      case K_AVG:
      case K_VAR: {
        double v = g.dtype == BOLT_F64 ? static_cast<const double*>(g.data)[row]
                 : g.dtype == BOLT_F32 ? (double)static_cast<const float*>(g.data)[row]
                 : g.dtype == BOLT_I64 ? (double)static_cast<const int64_t*>(g.data)[row]
                                       : (double)static_cast<const int32_t*>(g.data)[row];
        unsafeAtomicAdd(reinterpret_cast<double*>(w), v);
        break;
      }
      case K_MINMAX: {
        uint64_t k;
        switch (g.dtype) {
          case BOLT_F64: k = f64_order_key(static_cast<const double*>(g.data)[row]); break;
          case BOLT_F32: k = f64_order_key((double)static_cast<const float*>(g.data)[row]); break;
          case BOLT_I64: k = (uint64_t) static_cast<const int64_t*>(g.data)[row] ^ 0x8000000000000000ull; break;
          default: k = (uint64_t)(int64_t) static_cast<const int32_t*>(g.data)[row] ^ 0x8000000000000000ull; break;
        }
        if (g.op == BOLT_AGG_MIN) atomicMin(w, (unsigned long long)k);
        else atomicMax(w, (unsigned long long)k);
        break;
      }
    }
  }
}

// The 4- or 8-byte input word of aggregate `g` at `row`, for loading ahead of
// the slot lookup (0 for COUNT(*), Bool and Decimal128, which update_slot_pf
// reads at update time).
__device__ __forceinline__ uint64_t pf_word(const GAgg& g, int64_t row) {
  if (g.op == BOLT_AGG_COUNT_STAR || g.data == nullptr) return 0;
  switch (g.dtype) {
    case BOLT_I64: case BOLT_F64: return static_cast<const uint64_t*>(g.data)[row];
    case BOLT_I32: case BOLT_F32: case BOLT_DATE32: return static_cast<const uint32_t*>(g.data)[row];
  }
  return 0;
}

// update_slot with the input words of the first four aggregates already
// loaded (pf_word); the same atomics, so results are those of update_slot.
__device__ __forceinline__ void update_slot_pf(const GPlan& P, unsigned long long* words, size_t stride,
                                               uint32_t slot, int64_t row, uint64_t w0, uint64_t w1, uint64_t w2,
                                               uint64_t w3) {
  for (int a = 0; a < P.n; a++) {
    const GAgg& g = P.a[a];
    if (a >= 4 || g.dtype == BOLT_DEC128 || g.dtype == BOLT_BOOL) {
      GPlan one;
      one.n = 1;
      one.words = P.words;
      one.a[0] = g;
      update_slot(one, words, stride, slot, row);
      continue;
    }
    const uint64_t rv = a == 0 ? w0 : a == 1 ? w1 : a == 2 ? w2 : w3;
    unsigned long long* w = words + (size_t)g.w * stride + slot;
    if (g.op == BOLT_AGG_COUNT_STAR) { atomicAdd(w, 1ull); continue; }
    if (!bit_valid(g.validity, row)) continue;
    if (g.kind == K_COUNT) { atomicAdd(w, 1ull); continue; }
    if (g.nw >= 0) atomicAdd(words + (size_t)g.nw * stride + slot, 1ull);
    const int64_t iv = g.dtype == BOLT_I64 ? (int64_t)rv : (int64_t)(int32_t)(uint32_t)rv;
    const double fv = g.dtype == BOLT_F64 ? as_f64(rv)
                    : g.dtype == BOLT_F32 ? (double)__uint_as_float((uint32_t)rv) : (double)iv;
    switch (g.kind) {
      case K_ISUM: atomic_add128(w, w + stride, (uint64_t)iv, (uint64_t)(iv >> 63)); break;
      case K_FSUM: case K_AVG: case K_VAR: unsafeAtomicAdd(reinterpret_cast<double*>(w), fv); break;
      case K_MINMAX: {
        const uint64_t k = (g.dtype == BOLT_F64 || g.dtype == BOLT_F32) ? f64_order_key(fv)
                                                                          : (uint64_t)iv ^ 0x8000000000000000ull;
        if (g.op == BOLT_AGG_MIN) atomicMin(w, (unsigned long long)k);
        else atomicMax(w, (unsigned long long)k);
        break;
      }
    }
  }
}

// Merge the words of a partial record (one value per field, `pstride` apart)
// into a slot of the global table.
__device__ __forceinline__ void merge_slot(const GPlan& P, unsigned long long* words, size_t stride, uint32_t slot,
                                           const unsigned long long* pw, size_t pstride) {
  for (int a = 0; a < P.n; a++) {
    const GAgg& g = P.a[a];
    unsigned long long* w = words + (size_t)g.w * stride + slot;
    const unsigned long long* q = pw + (size_t)g.w * pstride;
    if (g.nw >= 0 && g.nw != g.w) {
      unsigned long long c = pw[(size_t)g.nw * pstride];
      if (c) atomicAdd(words + (size_t)g.nw * stride + slot, c);
    }
    switch (g.kind) {
      case K_COUNT: if (q[0]) atomicAdd(w, q[0]); break;
      case K_ISUM: atomic_add128(w, w + stride, q[0], q[pstride]); break;
      case K_DSUM: atomic_add192(w, stride, q[0], q[pstride], q[2 * pstride]); break;
      case K_FSUM: case K_AVG: case K_VAR: unsafeAtomicAdd(reinterpret_cast<double*>(w), as_f64(q[0])); break;
      case K_MINMAX:
        if (g.op == BOLT_AGG_MIN) atomicMin(w, q[0]);
        else atomicMax(w, q[0]);
        break;
    }
  }
}

// merge_slot for a plan of at most two words, whose words v0, v1 are in
// registers (word k of the record is v0 for k = 0, else v1).
__device__ __forceinline__ void merge_slot2(const GPlan& P, unsigned long long* words, size_t stride, uint32_t slot,
                                            unsigned long long v0, unsigned long long v1) {
  for (int a = 0; a < P.n; a++) {
    const GAgg& g = P.a[a];
    unsigned long long* w = words + (size_t)g.w * stride + slot;
    const unsigned long long q0 = g.w == 0 ? v0 : v1;
    if (g.nw >= 0 && g.nw != g.w) {
      const unsigned long long c = g.nw == 0 ? v0 : v1;
      if (c) atomicAdd(words + (size_t)g.nw * stride + slot, c);
    }
    switch (g.kind) {
      case K_COUNT: if (q0) atomicAdd(w, q0); break;
      case K_ISUM: atomic_add128(w, w + stride, q0, v1); break;  // two words: the sum is words 0 and 1
      case K_FSUM: case K_AVG: case K_VAR: unsafeAtomicAdd(reinterpret_cast<double*>(w), as_f64(q0)); break;
      case K_MINMAX:
        if (g.op == BOLT_AGG_MIN) atomicMin(w, q0);
        else atomicMax(w, q0);
        break;
    }
  }
}

__device__ __forceinline__ unsigned long long word_identity(const GPlan& P, int w) {
  for (int a = 0; a < P.n; a++)
    if (P.a[a].kind == K_MINMAX && P.a[a].w == w) return P.a[a].op == BOLT_AGG_MIN ? ~0ull : 0ull;
  return 0ull;
}

// Global table insert: returns the slot of `key` (claiming one if absent), or
// 0xFFFFFFFF when the probe bound is exceeded (the table is too small).
__device__ __forceinline__ uint32_t global_insert(unsigned long long* keys, uint64_t cap_mask, uint64_t key) {
  uint64_t s = mix64(key) & cap_mask;
  for (uint64_t p = 0; p < (uint64_t)kGlobalProbes && p <= cap_mask; p++) {
    unsigned long long cur = keys[s];
    if (cur == key) return (uint32_t)s;
    if (cur == kEmpty) {
      unsigned long long prev = atomicCAS(&keys[s], kEmpty, (unsigned long long)key);
      if (prev == kEmpty || prev == key) return (uint32_t)s;
    }
    s = (s + 1) & cap_mask;
  }
  return 0xFFFFFFFFu;
}

__device__ __forceinline__ uint32_t global_find(const unsigned long long* keys, uint64_t cap_mask, uint64_t key) {
  uint64_t s = mix64(key) & cap_mask;
  for (uint64_t p = 0; p < (uint64_t)kGlobalProbes && p <= cap_mask; p++) {
    unsigned long long cur = keys[s];
    if (cur == key) return (uint32_t)s;
    if (cur == kEmpty) return 0xFFFFFFFFu;
    s = (s + 1) & cap_mask;
  }
  return 0xFFFFFFFFu;
}

struct GTable {
  unsigned long long* keys;   // [cap]
  unsigned long long* words;  // [words][cap + 2]
  uint32_t* used_special;     // [2]
  uint32_t* err;              // [0] probe overflow, [1] lookback, [2] integer overflow
  uint64_t cap;               // power of two
};

__device__ __forceinline__ int64_t gslot(const GTable& t, int kind_tag, uint64_t key) {
  if (kind_tag == 1) { t.used_special[0] = 1; return (int64_t)t.cap; }
  if (kind_tag == 2 || key == kEmpty) { t.used_special[1] = 1; return (int64_t)t.cap + 1; }
  uint32_t s = global_insert(t.keys, t.cap - 1, key);
  if (s == 0xFFFFFFFFu) { atomicOr(&t.err[0], 1u); return -1; }
  return s;
}

struct OutCol {
  uint8_t* data;
  uint32_t* valid;  // zeroed bitmap words or nullptr
  int32_t dtype;
  int32_t pad;
};
struct OutSpec {
  OutCol key[2];
  OutCol agg[kMaxAggs];
  unsigned long long* packed;  // packed key per group (for ordering), i64 order; NULL group = i64::MAX
  uint32_t* slot_of;           // slot per output row
};

__device__ __forceinline__ void set_valid(uint32_t* bm, uint64_t pos) { atomicOr(&bm[pos >> 5], 1u << (pos & 31)); }

__device__ __forceinline__ void write_key(const OutCol& c, uint64_t pos, uint64_t bits) {
  if (c.dtype == BOLT_DEC128) {  // the low word, sign-extended
    reinterpret_cast<uint64_t*>(c.data)[2 * pos] = bits;
    reinterpret_cast<uint64_t*>(c.data)[2 * pos + 1] = (uint64_t)((int64_t)bits >> 63);
  } else if (c.dtype == BOLT_I64 || c.dtype == BOLT_F64) {
    reinterpret_cast<uint64_t*>(c.data)[pos] = bits;
  } else {
    reinterpret_cast<uint32_t*>(c.data)[pos] = (uint32_t)bits;
  }
}

__device__ __forceinline__ void finalize_to(const GTable& t, const GAgg& g, const OutCol& c, size_t total, uint64_t s,
                                            uint64_t pos) {
  auto W = [&](int w) { return t.words[(size_t)w * total + s]; };
  uint64_t n = g.nw >= 0 ? W(g.nw) : 1;
  bool valid = true;
  switch (g.kind) {
    case K_COUNT: reinterpret_cast<int64_t*>(c.data)[pos] = (int64_t)W(g.w); break;
    case K_ISUM: {
      uint64_t lo = W(g.w);
      int64_t hi = (int64_t)W(g.w + 1);
      if (hi != ((int64_t)lo >> 63)) atomicOr(&t.err[2], 1u);
      reinterpret_cast<int64_t*>(c.data)[pos] = (int64_t)lo;
      valid = n > 0;
      break;
    }
    case K_DSUM: {
      uint64_t lo = W(g.w), mid = W(g.w + 1);
      int64_t top = (int64_t)W(g.w + 2);
      if (top != ((int64_t)mid >> 63)) atomicOr(&t.err[2], 2u);
      reinterpret_cast<uint64_t*>(c.data)[2 * pos] = lo;
      reinterpret_cast<uint64_t*>(c.data)[2 * pos + 1] = mid;
      valid = n > 0;
      break;
    }
    case K_FSUM: {
      double v = as_f64(W(g.w));
      if (g.dtype == BOLT_F32) reinterpret_cast<float*>(c.data)[pos] = (float)v;
      else reinterpret_cast<double*>(c.data)[pos] = v;
      valid = n > 0;
      break;
    }
    case K_AVG: {
      double v = n ? as_f64(W(g.w)) / (double)n : 0.0;
      reinterpret_cast<double*>(c.data)[pos] = v;
      valid = n > 0;
      break;
    }
    case K_MINMAX: {
      uint64_t k = W(g.w);
      if (g.dtype == BOLT_F64) reinterpret_cast<double*>(c.data)[pos] = f64_from_order_key(k);
      else if (g.dtype == BOLT_F32) reinterpret_cast<float*>(c.data)[pos] = (float)f64_from_order_key(k);
      else if (g.dtype == BOLT_I64) reinterpret_cast<int64_t*>(c.data)[pos] = (int64_t)(k ^ 0x8000000000000000ull);
      else reinterpret_cast<int32_t*>(c.data)[pos] = (int32_t)(int64_t)(k ^ 0x8000000000000000ull);
      valid = n > 0;
      break;
    }
    case K_VAR: {
      bool samp = g.op == BOLT_AGG_VAR_SAMP || g.op == BOLT_AGG_STDDEV_SAMP;
      double m2 = as_f64(W(g.w + 1));
      valid = samp ? n > 1 : n > 0;
      double var = valid ? m2 / (double)(samp ? n - 1 : n) : 0.0;
      if (g.op == BOLT_AGG_STDDEV_POP || g.op == BOLT_AGG_STDDEV_SAMP) var = sqrt(var);
      reinterpret_cast<double*>(c.data)[pos] = var;
      break;
    }
  }
  if (c.valid && valid) set_valid(c.valid, pos);
}

// ---- host helpers (groupby.hip) ----
int make_gplan(const bolt_agg_t* aggs, int32_t n_aggs, const bolt_column_t* inputs, int32_t n_inputs, int64_t n,
               GPlan* P, int32_t* out_dtype);
double hll_estimate(const uint32_t* regs);

// The partitioned HBM tier's state (gb_part.hip): the items' partition
// ids and per-workgroup counts, the partition starts ([np + 1]) and the
// record buffer the scatter fills (records of partition p at
// [start[p], start[p + 1])).
struct PartSrc {
  const unsigned long long* pkeys;
  const uint8_t* ptag;
  const unsigned long long* pwords;
  int64_t pcap;
  int64_t n_parts;
  const uint32_t* rows;  // spilled row ids, in the LDS tier's spill segments (see per / spre)
  int64_t n_items;
  // spilled keys (PartOut::spill_key), or null: then they are read by row id
  const unsigned long long* keys;
  // rows and keys sit in the LDS tier's spill segments: spilled row j is at
  // b * per + (j - spre[b]) for the segment b holding it (spre: exclusive
  // prefix of the segments' counts, nseg + 1 entries)
  int64_t per;
  const uint64_t* spre;
  int nseg;
};
struct GbPartitions {
  int np = 0, pbits = 0, nb = 0, stride = 0;
  int64_t chunk = 0, recs_cap = 0;
  PartSrc src{};
  KeySpec ks{};
  DevBuf<uint64_t> spre;  // prefix of the spill segments' counts
  DevBuf<uint16_t> pid;
  DevBuf<uint32_t> hist;
  DevBuf<unsigned long long> cursor;
  DevBuf<uint64_t> start;
  DevBuf<unsigned long long> rec;     // final layout: partition p at [start[p], start[p + 1])
  DevBuf<unsigned long long> coarse;  // first staged pass: buckets of the top 6 bits
  DevBuf<unsigned long long> ccursor; // [64] scatter cursors of the first pass
};
int gb_part_build(hipStream_t s, const KeySpec& ks, const GPlan& P, const unsigned long long* pkeys,
                  const uint8_t* ptag, const unsigned long long* pwords, int64_t pcap, int64_t n_parts,
                  const uint32_t* spill, const unsigned long long* spill_key, int64_t spill_per,
                  const uint32_t* spill_seg, int nseg, int64_t n_spill, double est, int S, int bits_delta,
                  GbPartitions* out);
int gb_part_bits(double est, int S, int bits_delta);
int gb_part_reduce(hipStream_t s, const GTable& t, const GPlan& P, GbPartitions& parts, int S);
uint64_t next_pow2(uint64_t x);

}  // namespace bolt
