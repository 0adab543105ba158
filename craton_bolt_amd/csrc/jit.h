// jit.h — per-query fused row kernels generated as HIP source and compiled
// with hipRTC (see jit.hip).
#pragma once

#include <functional>
#include <string>
#include <vector>

#include "ops_common.h"

namespace bolt {

// The shape a generated kernel is specialised for (everything that changes
// the emitted source; pointers and literals are arguments).
struct JitSpec {
  struct Col {
    int dtype;
    bool has_valid;
  };
  struct Term {
    int col;
    int op;
  };
  struct Agg {
    int op;
    int src;  // 0 column, 1 join carry, 2 COUNT(*)
    int col;
    int dtype;
  };
  bool probe = false;
  bool bloom = false;  // probe through the build side's Bloom filter first
  int unroll = 4;  // rows per lane per block iteration (1..8)
  int probe_slots = 1;  // slots of each probe sequence read before the first compare (1..8)
  int key_dtype = 0;
  bool key_valid = false;
  int carry_dtype = 0;
  std::vector<Col> cols;
  std::vector<Term> terms;
  std::vector<Agg> aggs;
};

// The shape of a generated GROUP BY LDS-tier kernel: key encoding, filter,
// and each aggregate's accumulator words (the layout of the precompiled
// kernel's plan, so both produce the same partial records).
struct GbJitSpec {
  struct Agg {
    int op;
    int kind;   // accumulator kind of the plan (integer / decimal / float sum, count, min-max, ...)
    int dtype;
    int col;    // input column (-1 for COUNT(*))
    int w, nw;  // first word, count word (-1 = none)
  };
  int unroll = 8;  // rows per lane per block iteration (1..8)
  int n_keys = 1;
  int t0 = 0, t1 = 0;
  bool kv0 = false, kv1 = false;
  int words = 0;
  std::vector<unsigned long long> ident;  // identity of each word
  std::vector<JitSpec::Col> cols;
  std::vector<JitSpec::Term> terms;
  std::vector<Agg> aggs;
};

// The shape of a generated partitioned-join scatter: the probe key, the
// filter, and which columns are carried into the partitioned copies.
struct JpJitSpec {
  int items = 8;  // rows per lane per tile (tile = 1024 * items rows)
  int key_dtype = 0;
  bool key_valid = false;
  std::vector<JitSpec::Col> cols;
  std::vector<JitSpec::Term> terms;
  std::vector<int> pay;  // column index carried into partitioned copy j
};

bool jit_enabled();
// Device helpers every generated kernel starts with (device_acc.h, gb_lds.h).
const char* jit_prelude_text();
std::string jit_row_agg_source(const JitSpec& sp);
// Compact description of everything in `sp` that shapes jit_row_agg_source's
// text, for jit_get_keyed.
std::string jit_row_agg_key(const JitSpec& sp);
int jit_get(const std::string& src, hipFunction_t* fn);
int jit_spec_from(const DevFilter& f, const AggList& aggs, bool probe, int key_dtype, bool key_valid, int carry_dtype,
                  JitSpec* sp, JitLits* lits, JitCols* cols);
std::string jit_gb_lds_source(const GbJitSpec& sp);
// Partitioned join: the filtered scatter of the probe side into per-partition
// runs, and the region-ordered probe + aggregate over those runs.
std::string jit_jp_scatter_source(const JpJitSpec& sp);
std::string jp_scatter_jit_key(const JpJitSpec& sp);
std::string jit_jp_probe_source(const JitSpec& sp, int runs);
std::string gb_jit_key(const GbJitSpec& sp);
int jit_get_keyed(const std::string& key, const std::function<std::string()>& gen, hipFunction_t* fn);
int jit_row_agg_col(const void* d, const uint8_t* v, int dt, std::vector<JitSpec::Col>* cols, JitCols* jc);
int jit_launch_row_agg(hipFunction_t fn, unsigned grid, hipStream_t s, const uint4* tbl, uint64_t mask,
                       const uint32_t* flags, int64_t n, const void* key, const uint8_t* kvalid, const JitLits& lits,
                       const JitCols& cols, Acc* parts, const uint4* bloom, uint64_t bloom_mask);

}  // namespace bolt
