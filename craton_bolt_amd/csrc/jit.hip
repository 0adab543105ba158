// jit.hip — per-query fused kernels emitted as HIP source and compiled with
// hipRTC for gfx950.
//
// Replaces the reference's PTX emitters for fused predicates / projections
// feeding aggregates (src/jit/ptx_gen.rs, src/jit/scan_kernel.rs:265,
// src/jit/agg_kernels.rs:242) and its KernelSpec-keyed module cache
// (src/exec/module_cache.rs:291, src/jit/jit_compiler.rs:553 — an in-process
// cache keyed by the emitted text, plus the opt-in disk cache of
// src/jit/disk_cache.rs, here BOLT_JIT_CACHE_DIR holding gfx950 code objects).
//
// A generated kernel evaluates a conjunction of predicate terms and folds any
// number of aggregates over the rows that pass, optionally after an INNER
// hash-join probe (the fused filter + join + aggregate form).  Types,
// operators and the column set are baked into the source, so every value
// lives in registers and the inner loop has no dtype or operator dispatch;
// predicate literals are a by-value kernel argument, so queries differing only
// in their constants share one compiled kernel.  The accumulator code is the
// same text the precompiled kernels use (device_acc.h), so results are the
// same whichever path runs.  If hipRTC fails, callers fall back to the
// precompiled generic kernels; BOLT_JIT=0 forces that fallback.
#include <dirent.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>

#include <cstdlib>
#include <fstream>
#include <functional>
#include <mutex>
#include <sstream>
#include <unordered_map>
#include <vector>

#include "groupby_dev.h"
#include "jit.h"
#include "jit_prelude.inc"

namespace bolt {

bool jit_enabled() {
  const char* e = getenv("BOLT_JIT");
  return !(e && e[0] == '0');
}

namespace {

struct JitEntry {
  hipModule_t mod;
  hipFunction_t fn;
};
std::mutex g_mu;
std::unordered_map<std::string, JitEntry> g_cache;

const char* ctype(int dt) {
  switch (dt) {
    case BOLT_I32: case BOLT_DATE32: return "int";
    case BOLT_I64: return "long long";
    case BOLT_F32: return "float";
    case BOLT_F64: return "double";
    case BOLT_BOOL: return "unsigned char";
  }
  return "unsigned long long";
}

const char* cmp_op(int op) {
  switch (op) {
    case BOLT_CMP_LT: return "<";
    case BOLT_CMP_LE: return "<=";
    case BOLT_CMP_GT: return ">";
    case BOLT_CMP_GE: return ">=";
    case BOLT_CMP_EQ: return "==";
    case BOLT_CMP_NE: return "!=";
  }
  return "==";
}

// Predicate term t of row r (row index variable `i`, loaded value x<col>_<r>).
std::string term_expr(const std::vector<JitSpec::Col>& cols, const JitSpec::Term& tm, int t, int r,
                      const std::string& i) {
  const int c = tm.col;
  const bool hv = cols[c].has_valid;
  const int dt = cols[c].dtype;
  const std::string x = "x" + std::to_string(c) + "_" + std::to_string(r);
  const std::string valid = hv ? "vbit(v" + std::to_string(c) + ", " + i + ")" : "true";
  if (tm.op == BOLT_CMP_IS_NULL) return "!" + valid;
  if (tm.op == BOLT_CMP_IS_NOT_NULL) return valid;
  std::ostringstream e;
  if (dt == BOLT_F32 || dt == BOLT_F64)
    e << "((double)" << x << " " << cmp_op(tm.op) << " lits.f[" << t << "])";
  else if (dt == BOLT_DEC128)
    e << "(cmp128((long long)" << x << "_hi, " << x << "_lo, lits.hi[" << t << "], (unsigned long long)lits.i[" << t
      << "]) " << cmp_op(tm.op) << " 0)";
  else
    e << "((long long)" << x << " " << cmp_op(tm.op) << " lits.i[" << t << "])";
  return hv ? "(" + valid + " && " + e.str() + ")" : e.str();
}

// Pointer declarations c<k> (and v<k> for nullable columns) of the column set.
void emit_col_ptrs(std::ostringstream& o, const std::vector<JitSpec::Col>& cols) {
  for (int c = 0; c < (int)cols.size(); c++) {
    if (cols[c].dtype == BOLT_DEC128)
      o << "  const unsigned long long* __restrict__ c" << c << " = (const unsigned long long*)cols.d[" << c << "];\n";
    else
      o << "  const " << ctype(cols[c].dtype) << "* __restrict__ c" << c << " = (const " << ctype(cols[c].dtype)
        << "*)cols.d[" << c << "];\n";
    if (cols[c].has_valid) o << "  const unsigned char* __restrict__ v" << c << " = cols.v[" << c << "];\n";
  }
}

}  // namespace

const char* jit_prelude_text() { return kJitPrelude; }

// ---- persistent code-object cache (BOLT_JIT_CACHE_DIR) ----------------------
//
// The role of src/jit/disk_cache.rs for this build: generated kernels survive
// the process as gfx950 code objects, one file per kernel.
//   * Opt-in: BOLT_JIT_CACHE_DIR names the directory; unset or empty = off.
//   * Trust (disk_cache.rs:70-90): the directory must be owned by the
//     effective uid and not be group- or world-writable, else the disk layer
//     is off (a code object read back is launched).
//   * Key: FNV-1a 64 of a salt (library version, target, hipRTC version, the
//     compile options) followed by the source.  The file starts with a header
//     holding a magic, that key, an independent digest and the length of the
//     source, and the code length; a file whose header does not match the
//     kernel being looked up, or whose body is short, is ignored and
//     recompiled (no collision or stale file is ever loaded).
//   * Atomicity: written to a per-process, per-thread temporary name in the
//     same directory, then renamed into place.
//   * Bounds: at most BOLT_JIT_CACHE_MAX_ENTRIES files (default 4096) and
//     BOLT_JIT_CACHE_MAX_BYTES bytes (default 256 MiB); the oldest files by
//     modification time go first.
namespace {

const char* const kJitOpts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                                "-packed-fp32-ops"};
constexpr int kJitNOpts = (int)(sizeof kJitOpts / sizeof kJitOpts[0]);
constexpr char kCoMagic[8] = {'B', 'O', 'L', 'T', 'C', 'O', '1', '\0'};

struct CoHeader {
  char magic[8];
  uint64_t key;
  uint64_t src_digest;
  uint64_t src_len;
  uint64_t code_len;
};

std::atomic<int64_t> g_compiled{0}, g_disk_hits{0}, g_disk_stores{0};

uint64_t fnv1a_from(uint64_t h, const std::string& s) {
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ull;
  }
  return h;
}

const std::string& cache_salt() {
  static const std::string salt = [] {
    int maj = 0, min = 0;
    hiprtcVersion(&maj, &min);
    std::ostringstream o;
    o << "craton_bolt_amd 0.1|gfx950|hiprtc " << maj << '.' << min << '|';
    for (int i = 0; i < kJitNOpts; i++) o << kJitOpts[i] << ' ';
    o << '\n';
    return o.str();
  }();
  return salt;
}

// The cache directory, or "" when the disk layer is off.
std::string cache_dir() {
  const char* dir = getenv("BOLT_JIT_CACHE_DIR");
  if (!dir || !*dir) return "";
  struct stat st;
  if (stat(dir, &st) != 0 || !S_ISDIR(st.st_mode) || st.st_uid != geteuid() || (st.st_mode & 022) != 0) return "";
  return dir;
}

int64_t env_i64(const char* name, int64_t dflt) {
  const char* e = getenv(name);
  return e && *e ? strtoll(e, nullptr, 10) : dflt;
}

// Removes the oldest cache files until both bounds hold.
void cache_evict(const std::string& dir) {
  const int64_t max_n = env_i64("BOLT_JIT_CACHE_MAX_ENTRIES", 4096);
  const int64_t max_b = env_i64("BOLT_JIT_CACHE_MAX_BYTES", (int64_t)256 << 20);
  DIR* d = opendir(dir.c_str());
  if (!d) return;
  struct Ent {
    std::string path;
    int64_t mtime, size;
  };
  std::vector<Ent> ents;
  int64_t total = 0;
  while (struct dirent* e = readdir(d)) {
    const std::string name = e->d_name;
    if (name.size() < 9 || name.compare(0, 5, "bolt_") != 0 || name.compare(name.size() - 3, 3, ".co") != 0) continue;
    struct stat st;
    const std::string path = dir + "/" + name;
    if (stat(path.c_str(), &st) != 0) continue;
    ents.push_back({path, (int64_t)st.st_mtime, (int64_t)st.st_size});
    total += st.st_size;
  }
  closedir(d);
  if ((max_n <= 0 || (int64_t)ents.size() <= max_n) && (max_b <= 0 || total <= max_b)) return;
  std::sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) { return a.mtime < b.mtime; });
  int64_t n = (int64_t)ents.size();
  for (const Ent& e : ents) {
    if ((max_n <= 0 || n <= max_n) && (max_b <= 0 || total <= max_b)) break;
    if (unlink(e.path.c_str()) == 0) {
      n--;
      total -= e.size;
    }
  }
}

bool cache_read(const std::string& path, uint64_t key, uint64_t digest, uint64_t src_len, std::vector<char>* code) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return false;
  CoHeader h;
  if (!in.read(reinterpret_cast<char*>(&h), sizeof h)) return false;
  if (memcmp(h.magic, kCoMagic, 8) != 0 || h.key != key || h.src_digest != digest || h.src_len != src_len ||
      h.code_len == 0 || h.code_len > ((uint64_t)1 << 30))
    return false;
  code->resize(h.code_len);
  if (!in.read(code->data(), (std::streamsize)h.code_len)) return false;
  return in.peek() == std::char_traits<char>::eof();
}

void cache_write(const std::string& dir, const std::string& path, const CoHeader& h, const std::vector<char>& code) {
  std::ostringstream tmp;
  tmp << path << ".tmp." << getpid() << '.' << std::hash<std::thread::id>{}(std::this_thread::get_id());
  {
    std::ofstream out(tmp.str(), std::ios::binary);
    out.write(reinterpret_cast<const char*>(&h), sizeof h);
    out.write(code.data(), (std::streamsize)code.size());
    if (!out) {
      out.close();
      unlink(tmp.str().c_str());
      return;
    }
  }
  if (std::rename(tmp.str().c_str(), path.c_str()) != 0) {
    unlink(tmp.str().c_str());
    return;
  }
  g_disk_stores++;
  cache_evict(dir);
}

}  // namespace

int jit_get(const std::string& src, hipFunction_t* fn) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cache.find(src);
  if (it != g_cache.end()) {
    *fn = it->second.fn;
    return BOLT_OK;
  }
  std::vector<char> code;
  std::string path;
  const std::string dir = cache_dir();
  CoHeader hdr;
  memset(&hdr, 0, sizeof hdr);
  if (!dir.empty()) {
    memcpy(hdr.magic, kCoMagic, 8);
    hdr.key = fnv1a_from(fnv1a_from(1469598103934665603ull, cache_salt()), src);
    hdr.src_digest = fnv1a_from(0x6a09e667f3bcc908ull, src);  // a second basis: an independent digest
    hdr.src_len = src.size();
    char name[64];
    snprintf(name, sizeof name, "/bolt_%016llx.co", (unsigned long long)hdr.key);
    path = dir + name;
    if (cache_read(path, hdr.key, hdr.src_digest, hdr.src_len, &code)) g_disk_hits++;
    else code.clear();
  }
  if (code.empty()) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "bolt_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
      return set_error(BOLT_ERR_OTHER, "hipRTC: program creation failed");
    hiprtcResult r = hiprtcCompileProgram(prog, kJitNOpts, kJitOpts);
    if (r != HIPRTC_SUCCESS) {
      size_t ls = 0;
      hiprtcGetProgramLogSize(prog, &ls);
      std::string log(ls + 1, '\0');
      hiprtcGetProgramLog(prog, &log[0]);
      hiprtcDestroyProgram(&prog);
      return set_error(BOLT_ERR_OTHER, "hipRTC compile failed: %s", log.substr(0, 800).c_str());
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    g_compiled++;
    if (!path.empty()) {
      hdr.code_len = code.size();
      cache_write(dir, path, hdr, code);
    }
  }
  JitEntry e;
  hipError_t he = hipModuleLoadData(&e.mod, code.data());
  if (he == hipSuccess) {
    he = hipModuleGetFunction(&e.fn, e.mod, "bolt_jit_kernel");
    if (he != hipSuccess) (void)hipModuleUnload(e.mod);
  }
  if (he != hipSuccess) {
    // callers fall back to precompiled kernels: leave no HIP error pending for
    // their launch checks to report
    (void)hipGetLastError();
    return set_error(BOLT_ERR_DEVICE, "hipRTC module load: %s (%d)", hipGetErrorString(he), (int)he);
  }
  g_cache.emplace(src, e);
  *fn = e.fn;
  return BOLT_OK;
}

// Kernels looked up by a compact description of their spec, so that a query
// repeated with the same shape neither regenerates nor hashes the source.
int jit_get_keyed(const std::string& key, const std::function<std::string()>& gen, hipFunction_t* fn) {
  static std::mutex mu;
  static std::unordered_map<std::string, hipFunction_t> by_key;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = by_key.find(key);
    if (it != by_key.end()) {
      *fn = it->second;
      return BOLT_OK;
    }
  }
  BOLT_TRY(jit_get(gen(), fn));
  std::lock_guard<std::mutex> lk(mu);
  by_key.emplace(key, *fn);
  return BOLT_OK;
}

std::string jit_row_agg_key(const JitSpec& sp) {
  std::ostringstream k;
  k << "row:" << sp.probe << sp.bloom << ':' << sp.unroll << ':' << sp.probe_slots << ':' << sp.key_dtype << ':'
    << sp.key_valid << ':' << sp.carry_dtype << '|';
  for (const JitSpec::Col& c : sp.cols) k << c.dtype << (c.has_valid ? 'v' : 'n') << ',';
  k << '|';
  for (const JitSpec::Term& t : sp.terms) k << t.col << '.' << t.op << ',';
  k << '|';
  for (const JitSpec::Agg& g : sp.aggs) k << g.op << '.' << g.src << '.' << g.col << '.' << g.dtype << ',';
  return k.str();
}

std::string gb_jit_key(const GbJitSpec& sp) {
  std::ostringstream k;
  k << "gb:" << sp.unroll << ':' << sp.n_keys << ':' << sp.t0 << ':' << sp.t1 << ':' << sp.kv0 << sp.kv1 << ':'
    << sp.words << '|';
  for (unsigned long long v : sp.ident) k << v << ',';
  k << '|';
  for (const JitSpec::Col& c : sp.cols) k << c.dtype << (c.has_valid ? 'v' : 'n') << ',';
  k << '|';
  for (const JitSpec::Term& t : sp.terms) k << t.col << '.' << t.op << ',';
  k << '|';
  for (const GbJitSpec::Agg& g : sp.aggs)
    k << g.op << '.' << g.kind << '.' << g.dtype << '.' << g.col << '.' << g.w << '.' << g.nw << ',';
  return k.str();
}

// ---- source generation ------------------------------------------------------------

namespace {

// Accumulator updates of row r (values x<col>_<r>, row index i<r>, join carry w),
// one statement per aggregate, guarded by the input's validity bit.
std::string agg_updates(const JitSpec& sp, int r, const char* ind) {
  const int na = (int)sp.aggs.size();
  std::ostringstream up;
  const std::string i = "i" + std::to_string(r);
  for (int a = 0; a < na; a++) {
    const JitSpec::Agg& g = sp.aggs[a];
    const std::string A = "acc" + std::to_string(a);
    if (g.op == BOLT_AGG_COUNT_STAR) {
      up << ind << A << ".n++;\n";
      continue;
    }
    std::string V;
    int dt = g.dtype;
    std::string guard;
    if (g.src == 1) {
      V = dt == BOLT_F32 ? "(double)__uint_as_float(w)" : "(long long)(int)w";
    } else {
      V = "x" + std::to_string(g.col) + "_" + std::to_string(r);
      if (sp.cols[g.col].has_valid) guard = "vbit(v" + std::to_string(g.col) + ", " + i + ")";
    }
    std::ostringstream s;
    switch (g.op) {
      case BOLT_AGG_COUNT: s << A << ".n++;"; break;
      case BOLT_AGG_SUM:
        if (dt == BOLT_DEC128) s << "add128(" << A << ", " << V << "_lo, (long long)" << V << "_hi); " << A << ".n++;";
        else if (dt == BOLT_F32 || dt == BOLT_F64) s << "neum_add(" << A << ", (double)" << V << "); " << A << ".n++;";
        else s << "{ const long long v_ = (long long)" << V << "; add128(" << A << ", (unsigned long long)v_, v_ >> 63); "
               << A << ".n++; }";
        break;
      case BOLT_AGG_AVG: s << "neum_add(" << A << ", (double)" << V << "); " << A << ".n++;"; break;
      case BOLT_AGG_MIN:
      case BOLT_AGG_MAX: {
        const char* is_min = g.op == BOLT_AGG_MIN ? "true" : "false";
        if (dt == BOLT_F32 || dt == BOLT_F64)
          s << "minmax_push(" << A << ", " << is_min << ", f64_order_key((double)" << V << "));";
        else
          s << "minmax_push(" << A << ", " << is_min << ", (unsigned long long)(long long)" << V
            << " ^ 0x8000000000000000ull);";
        break;
      }
      default: s << "welford_push(" << A << ", (double)" << V << ");"; break;
    }
    if (!guard.empty()) up << ind << "if (" << guard << ") { " << s.str() << " }\n";
    else up << ind << s.str() << "\n";
  }
  return up.str();
}

// Probe sequences of R rows whose keys are k<r> and home slots s<r>; row r
// probes when `act(r)` holds.  The first K slots of every row's sequence are
// read together (8 slots share a 128-byte line, so mostly one line), rows
// that do not probe reading slots from 0, which stay cached; each row then
// resolves its chain from registers, and only a chain longer than K slots
// goes on with dependent reads.  The walk is that of probe_chain (join_dev.h):
// stop at the first empty slot, or after the first match when first_only.
std::string probe_chains(const JitSpec& sp, int R, const std::function<std::string(int)>& act, const char* ind) {
  const int K = sp.probe_slots;
  std::ostringstream o;
  for (int r = 0; r < R; r++) {
    const std::string rs = std::to_string(r);
    o << ind << "const unsigned long long b" << rs << " = " << act(r) << " ? s" << rs << " : 0ull;\n";
    for (int j = 0; j < K; j++)
      o << ind << "const uint4 e" << rs << "_" << j << " = tbl[(b" << rs << " + " << j << "ull) & mask];\n";
  }
  auto match = [&](const std::string& e, int r) {
    return "(long long)(((unsigned long long)" + e + ".y << 32) | " + e + ".x) == k" + std::to_string(r);
  };
  const std::string in2 = std::string(ind) + "  ", in4 = std::string(ind) + "    ", in6 = std::string(ind) + "      ";
  for (int r = 0; r < R; r++) {
    const std::string rs = std::to_string(r);
    o << ind << "if (" << act(r) << ") {\n" << in2 << "bool live = true;\n";
    for (int j = 0; j < K; j++) {
      const std::string e = "e" + rs + "_" + std::to_string(j);
      o << in2 << "if (live && " << j << "ull <= mask) {\n"
        << in4 << "if (" << e << ".w == 0u) live = false;\n"
        << in4 << "else if (" << match(e, r) << ") {\n"
        << in6 << "const unsigned int w = " << e << ".z;\n"
        << agg_updates(sp, r, in6.c_str())
        << in6 << "if (first_only) live = false;\n"
        << in4 << "}\n"
        << in2 << "}\n";
    }
    o << in2 << "if (live && " << K << "ull <= mask) {\n"
      << in4 << "unsigned long long sv = (s" << rs << " + " << K << "ull) & mask;\n"
      << in4 << "uint4 e = tbl[sv];\n"
      << in4 << "for (unsigned long long step = " << K << "; e.w != 0u && step <= mask; step++) {\n"
      << in6 << "if (" << match("e", r) << ") {\n"
      << in6 << "  const unsigned int w = e.z;\n"
      << agg_updates(sp, r, (in6 + "  ").c_str())
      << in6 << "  if (first_only) break;\n"
      << in6 << "}\n"
      << in6 << "sv = (sv + 1) & mask;\n"
      << in6 << "e = tbl[sv];\n"
      << in4 << "}\n"
      << in2 << "}\n"
      << ind << "}\n";
  }
  return o.str();
}

}  // namespace

std::string jit_row_agg_source(const JitSpec& sp) {
  std::ostringstream o;
  o << kJitPrelude << "\n";
  o << "extern \"C\" __global__ void __launch_bounds__(256) bolt_jit_kernel(\n"
       "    const uint4* __restrict__ tbl, unsigned long long mask, const unsigned int* __restrict__ flags,\n"
       "    long long n, const void* __restrict__ keyp, const unsigned char* __restrict__ kvalid,\n"
       "    bolt::JitLits lits, bolt::JitCols cols, bolt::Acc* __restrict__ parts,\n"
       "    const uint4* __restrict__ bloom, unsigned long long bmask) {\n"
       "  using namespace bolt;\n";
  const int nc = (int)sp.cols.size();
  emit_col_ptrs(o, sp.cols);
  const int na = (int)sp.aggs.size();
  const int R = sp.unroll;
  for (int a = 0; a < na; a++) o << "  Acc acc" << a << "; acc_init(acc" << a << ", " << sp.aggs[a].op << ");\n";
  if (sp.probe) o << "  const bool first_only = flags[0] == 0u;\n";
  // Each block iteration covers R * 256 rows; lane t owns rows base + r * 256
  // (r < R), so every load is a coalesced 256-row stripe and the R rows' first
  // table reads are independent and in flight together.
  o << "  const long long tile = 256LL * " << R << ";\n"
       "  for (long long base = (long long)blockIdx.x * tile + threadIdx.x; base < n; base += (long long)gridDim.x * tile) {\n";
  for (int r = 0; r < R; r++) {
    const std::string i = "i" + std::to_string(r);
    const std::string in = "in" + std::to_string(r);
    o << "    const long long " << i << " = base + " << 256 * r << ";\n"
      << "    const bool " << in << " = " << i << " < n;\n";
    if (sp.probe)
      o << "    const " << ctype(sp.key_dtype) << " kraw" << r << " = " << in << " ? __builtin_nontemporal_load((const "
        << ctype(sp.key_dtype) << "*)keyp + " << i << ") : 0;\n";
    for (int c = 0; c < nc; c++) {
      const std::string x = "x" + std::to_string(c) + "_" + std::to_string(r);
      if (sp.cols[c].dtype == BOLT_DEC128)
        o << "    const unsigned long long " << x << "_lo = " << in << " ? c" << c << "[2 * " << i << "] : 0, " << x
          << "_hi = " << in << " ? c" << c << "[2 * " << i << " + 1] : 0;\n";
      else if (sp.cols[c].dtype == BOLT_BOOL)
        o << "    const unsigned char " << x << " = " << in << " ? c" << c << "[" << i << "] : 0;\n";
      else
        o << "    const " << ctype(sp.cols[c].dtype) << " " << x << " = " << in << " ? __builtin_nontemporal_load(c" << c
          << " + " << i << ") : 0;\n";
    }
  }
  for (int r = 0; r < R; r++) {
    const std::string i = "i" + std::to_string(r);
    const std::string keep = "keep" + std::to_string(r);
    o << "    bool " << keep << " = in" << r << ";\n";
    if (sp.probe && sp.key_valid) o << "    " << keep << " = " << keep << " && vbit(kvalid, " << i << ");\n";
    for (int t = 0; t < (int)sp.terms.size(); t++)
      o << "    " << keep << " = " << keep << " && " << term_expr(sp.cols, sp.terms[t], t, r, i) << ";\n";
  }
  auto updates = [&](int r, const char* ind) { return agg_updates(sp, r, ind); };
  if (sp.probe) {
    // first reads of all R rows, then each row's chain
    for (int r = 0; r < R; r++)
      o << "    const long long k" << r << " = (long long)kraw" << r << ";\n"
        << "    const unsigned long long h" << r << " = fmix64((unsigned long long)k" << r << ");\n"
        << "    unsigned long long s" << r << " = h" << r << " & mask;\n";
    if (sp.bloom) {
      // Bloom blocks of all R rows first (cache resident), then the table
      for (int r = 0; r < R; r++)
        o << "    const uint4 b" << r << " = bloom[keep" << r << " ? ((h" << r << " >> 40) & bmask) : 0];\n";
      for (int r = 0; r < R; r++)
        o << "    keep" << r << " = keep" << r << " && bloom_hit(b" << r << ", h" << r << ");\n";
    }
    o << probe_chains(sp, R, [](int r) { return "keep" + std::to_string(r); }, "    ");
  } else {
    for (int r = 0; r < R; r++) o << "    if (keep" << r << ") {\n" << updates(r, "      ") << "    }\n";
  }
  o << "  }\n  __shared__ Acc sh[4];\n";
  for (int a = 0; a < na; a++)
    o << "  block_reduce_acc(acc" << a << ", " << sp.aggs[a].op << ", " << sp.aggs[a].dtype << ", sh, 4, parts + (long long)blockIdx.x * "
      << na << " + " << a << ");\n";
  o << "}\n";
  return o.str();
}

// Partitioned join, scatter (the generated form of jp_scatter_kernel in
// join_part.hip): workgroup g owns chunk g of the probe side and walks it in
// tiles of 1024 * R rows.  A lane's R rows' key and column loads are issued
// together, and the next tile's are issued as soon as the current tile sits in
// LDS, so they are in flight while the tile is written out.  Filter terms and
// column types are baked in; ranking, staging and the write-out are the
// shared jp_tile.h steps.
std::string jit_jp_scatter_source(const JpJitSpec& sp) {
  std::ostringstream o;
  const int R = sp.items;
  const int np = (int)sp.pay.size();
  const int nc = (int)sp.cols.size();
  o << kJitPrelude << "\n";
  o << "extern \"C\" __global__ void __launch_bounds__(1024) bolt_jit_kernel(\n"
       "    const void* __restrict__ keyp, const unsigned char* __restrict__ kvalid, long long n, long long per,\n"
       "    bolt::JitLits lits, bolt::JitCols cols, unsigned long long mask, int shift, int P,\n"
       "    const unsigned int* __restrict__ counts, const unsigned int* __restrict__ base,\n"
       "    long long* __restrict__ okeys, bolt::JpDst dst, unsigned int* __restrict__ actual) {\n"
       "  using namespace bolt;\n"
       "  constexpr int T = kJpBlock * " << R << ";\n"
       "  __shared__ JpTileLds<T, " << np << "> L;\n";
  emit_col_ptrs(o, sp.cols);
  o << "  const int g = blockIdx.x;\n"
       "  jp_tile_begin(L, P, g, counts, base);\n"
       "  const long long lo = (long long)g * per;\n"
       "  const long long hi = lo + per < n ? lo + per : n;\n";
  const char* kt = ctype(sp.key_dtype);
  for (int r = 0; r < R; r++) {
    o << "  " << kt << " kraw" << r << " = 0;\n";
    for (int c = 0; c < nc; c++) {
      const std::string x = "x" + std::to_string(c) + "_" + std::to_string(r);
      if (sp.cols[c].dtype == BOLT_DEC128) o << "  unsigned long long " << x << "_lo = 0, " << x << "_hi = 0;\n";
      else o << "  " << ctype(sp.cols[c].dtype) << " " << x << " = 0;\n";
    }
  }
  // loads of the tile starting at row `t` (every index clamped into the chunk)
  auto loads = [&](const std::string& t, const char* ind) {
    std::ostringstream l;
    for (int r = 0; r < R; r++) {
      const std::string ic = "ic" + std::to_string(r);
      l << ind << "{\n"
        << ind << "  const long long " << ic << "_ = " << t << " + " << 1024 * r << " + (long long)threadIdx.x;\n"
        << ind << "  const long long " << ic << " = " << ic << "_ < hi ? " << ic << "_ : hi - 1;\n"
        << ind << "  kraw" << r << " = __builtin_nontemporal_load((const " << kt << "*)keyp + " << ic << ");\n";
      for (int c = 0; c < nc; c++) {
        const std::string x = "x" + std::to_string(c) + "_" + std::to_string(r);
        if (sp.cols[c].dtype == BOLT_DEC128)
          l << ind << "  " << x << "_lo = c" << c << "[2 * " << ic << "]; " << x << "_hi = c" << c << "[2 * " << ic
            << " + 1];\n";
        else if (sp.cols[c].dtype == BOLT_BOOL)
          l << ind << "  " << x << " = c" << c << "[" << ic << "];\n";
        else
          l << ind << "  " << x << " = __builtin_nontemporal_load(c" << c << " + " << ic << ");\n";
      }
      l << ind << "}\n";
    }
    return l.str();
  };
  o << "  if (lo < hi) {\n" << loads("lo", "    ") << "  }\n";
  o << "  for (long long t0 = lo; t0 < hi; t0 += T) {\n"
       "    long long k[" << R << "];\n"
       "    int part[" << R << "];\n"
       "    unsigned long long pv[" << R << "][" << (np > 0 ? np : 1) << "];\n";
  for (int r = 0; r < R; r++) {
    const std::string i = "i" + std::to_string(r);
    const std::string keep = "keep" + std::to_string(r);
    o << "    {\n"
      << "      const long long " << i << " = t0 + " << 1024 * r << " + (long long)threadIdx.x;\n"
      << "      bool " << keep << " = " << i << " < hi;\n";
    if (sp.key_valid) o << "      " << keep << " = " << keep << " && vbit(kvalid, " << i << ");\n";
    for (int t = 0; t < (int)sp.terms.size(); t++)
      o << "      " << keep << " = " << keep << " && " << term_expr(sp.cols, sp.terms[t], t, r, i) << ";\n";
    o << "      k[" << r << "] = (long long)kraw" << r << ";\n"
      << "      part[" << r << "] = " << keep << " ? (int)((fmix64((unsigned long long)k[" << r
      << "]) & mask) >> shift) : -1;\n";
    for (int j = 0; j < np; j++) {
      const int c = sp.pay[j];
      const std::string x = "x" + std::to_string(c) + "_" + std::to_string(r);
      std::string bits;
      switch (sp.cols[c].dtype) {
        case BOLT_F64: bits = "(unsigned long long)__double_as_longlong(" + x + ")"; break;
        case BOLT_F32: bits = "(unsigned long long)__float_as_uint(" + x + ")"; break;
        case BOLT_I64: bits = "(unsigned long long)" + x; break;
        default: bits = "(unsigned long long)(unsigned int)" + x; break;
      }
      o << "      pv[" << r << "][" << j << "] = " << bits << ";\n";
    }
    o << "    }\n";
  }
  o << "    jp_tile_stage<T, " << np << ", " << R << ">(L, P, k, part, pv);\n"
       "    if (t0 + T < hi) {\n" << loads("(t0 + T)", "      ") << "    }\n"
       "    jp_tile_flush<T, " << np << ">(L, P, okeys, dst);\n"
       "  }\n"
       "  jp_tile_end<T, " << np << ">(L, P, g, counts, base, actual);\n"
       "}\n";
  return o.str();
}

std::string jp_scatter_jit_key(const JpJitSpec& sp) {
  std::ostringstream k;
  k << "jps:" << sp.items << ':' << sp.key_dtype << ':' << sp.key_valid << '|';
  for (const JitSpec::Col& c : sp.cols) k << c.dtype << (c.has_valid ? 'v' : 'n') << ',';
  k << '|';
  for (const JitSpec::Term& t : sp.terms) k << t.col << '.' << t.op << ',';
  k << '|';
  for (int c : sp.pay) k << c << ',';
  return k.str();
}

// Partitioned join, probe (the generated form of jp_probe_kernel in
// join_part.hip).  Workgroup b serves XCD b % 8 (workgroups are dispatched to
// the XCDs round-robin) and takes work items from that XCD's queue: an item
// is `runs` consecutive chunk runs of one partition, the XCD's partitions in
// order, so its workgroups probe one table region at a time, which stays in
// its L2.  A lane's R rows' keys and carried columns are loaded together
// (about half the rows match, so nearly every line of the runs is read
// anyway), then their first table slots, then each row's chain.  The
// columns of `sp` are the partitioned copies (no validity: rows with a NULL
// input never take this form).
std::string jit_jp_probe_source(const JitSpec& sp, int runs) {
  std::ostringstream o;
  const int R = sp.unroll;
  const int nc = (int)sp.cols.size();
  const int na = (int)sp.aggs.size();
  o << kJitPrelude << "\n";
  o << "extern \"C\" __global__ void __launch_bounds__(256) bolt_jit_kernel(\n"
       "    const uint4* __restrict__ tbl, unsigned long long mask, const unsigned int* __restrict__ flags,\n"
       "    const long long* __restrict__ pkeys, int P, int G, const unsigned int* __restrict__ counts,\n"
       "    const unsigned int* __restrict__ base, const unsigned int* __restrict__ actual,\n"
       "    unsigned int* __restrict__ queue, bolt::JitCols cols, bolt::Acc* __restrict__ parts) {\n"
       "  using namespace bolt;\n";
  emit_col_ptrs(o, sp.cols);
  for (int a = 0; a < na; a++) o << "  Acc acc" << a << "; acc_init(acc" << a << ", " << sp.aggs[a].op << ");\n";
  o << "  const bool first_only = flags[0] == 0u;\n"
       "  constexpr int CG = " << runs << ";\n"
       "  __shared__ unsigned int sh_item;\n"
       "  const int x = blockIdx.x & 7;\n"
       "  const unsigned int per_part = (unsigned int)(G / CG);\n"
       "  const unsigned int items = (unsigned int)(P / 8) * per_part;\n"
       "  for (;;) {\n"
       "    if (threadIdx.x == 0) sh_item = atomicAdd(&queue[x], 1u);\n"
       "    __syncthreads();\n"
       "    const unsigned int item = sh_item;\n"
       "    __syncthreads();\n"
       "    if (item >= items) break;\n"
       "    const int p = x + 8 * (int)(item / per_part);\n"
       "    const int g0 = (int)(item % per_part) * CG;\n"
       "    long long st[CG];\n"
       "    unsigned int pre[CG + 1];\n"
       "    pre[0] = 0u;\n"
       "#pragma unroll\n"
       "    for (int c = 0; c < CG; c++) {\n"
       "      const long long cell = (long long)(g0 + c) * P + p;\n"
       "      st[c] = (long long)base[p] + counts[cell];\n"
       "      pre[c + 1] = pre[c] + actual[cell];\n"
       "    }\n"
       "    const unsigned int len = pre[CG];\n"
       "    for (unsigned int r0 = 0; r0 < len; r0 += 256u * " << R << ") {\n";
  for (int r = 0; r < R; r++) {
    const std::string rs = std::to_string(r);
    o << "      const unsigned int v" << rs << " = r0 + " << 256 * r << "u + threadIdx.x;\n"
      << "      const bool in" << rs << " = v" << rs << " < len;\n"
      << "      const unsigned int u" << rs << " = in" << rs << " ? v" << rs << " : len - 1u;\n"
      << "      long long i" << rs << " = st[0] + u" << rs << ";\n"
      << "#pragma unroll\n"
      << "      for (int c = 1; c < CG; c++)\n"
      << "        if (u" << rs << " >= pre[c]) i" << rs << " = st[c] + (u" << rs << " - pre[c]);\n"
      << "      const long long k" << rs << " = __builtin_nontemporal_load(pkeys + i" << rs << ");\n";
    for (int c = 0; c < nc; c++)
      o << "      const " << ctype(sp.cols[c].dtype) << " x" << c << "_" << rs << " = __builtin_nontemporal_load(c" << c
        << " + i" << rs << ");\n";
    o << "      unsigned long long s" << rs << " = fmix64((unsigned long long)k" << rs << ") & mask;\n";
  }
  o << probe_chains(sp, R, [](int r) { return "in" + std::to_string(r); }, "      ");
  o << "    }\n  }\n  __shared__ Acc sh[4];\n";
  for (int a = 0; a < na; a++)
    o << "  block_reduce_acc(acc" << a << ", " << sp.aggs[a].op << ", " << sp.aggs[a].dtype
      << ", sh, 4, parts + (long long)blockIdx.x * " << na << " + " << a << ");\n";
  o << "}\n";
  return o.str();
}

// GROUP BY LDS tier (the generated form of gb_lds_kernel in groupby.hip):
// each lane owns R rows per block iteration, 1024 apart, whose key, filter
// and input loads are all issued before the first row touches the LDS table,
// so R rows' HBM reads are in flight per lane instead of one.  Slot lookup,
// spilling and the flush are the shared gb_lds.h helpers; the accumulator
// updates are the same LDS atomics as update_slot, specialised to the plan.
std::string jit_gb_lds_source(const GbJitSpec& sp) {
  std::ostringstream o;
  const int R = sp.unroll;
  const int W = sp.words;
  o << kJitPrelude << "\n";
  o << "extern \"C\" __global__ void __launch_bounds__(1024) bolt_jit_kernel(\n"
       "    long long n, int S, const void* __restrict__ k0p, const unsigned char* __restrict__ k0v,\n"
       "    const void* __restrict__ k1p, const unsigned char* __restrict__ k1v, bolt::JitLits lits,\n"
       "    bolt::JitCols cols, bolt::PartOut out) {\n"
       "  using namespace bolt;\n"
       "  extern __shared__ u64_t sm[];\n"
       "  const int SS = S + 2;\n"
       "  u64_t* skeys = sm;\n"
       "  u64_t* swords = sm + SS;\n"
       "  u32_t* shll = (u32_t*)(swords + (u64_t)" << W << " * SS);\n"
       "  u32_t* sused = shll + kHllRegs;\n"
       "  for (int i = threadIdx.x; i < SS; i += 1024) {\n"
       "    skeys[i] = kEmpty;\n";
  for (int w = 0; w < W; w++) o << "    swords[" << w << " * SS + i] = " << sp.ident[w] << "ull;\n";
  o << "  }\n"
       "  for (int i = threadIdx.x; i < kHllRegs; i += 1024) shll[i] = 0;\n"
       "  if (threadIdx.x < 4) sused[threadIdx.x] = 0;\n"
       "  __syncthreads();\n";
  auto kty = [](int dt) { return dtype_size(dt) == 4 ? "u32_t" : "u64_t"; };
  auto kcanon = [](int dt, const std::string& v) -> std::string {
    if (dt == BOLT_F32) return "(u64_t)canon_f32_bits(" + v + ")";
    if (dt == BOLT_F64) return "canon_f64_bits(" + v + ")";
    return "(u64_t)" + v;
  };
  o << "  const " << kty(sp.t0) << "* __restrict__ k0 = (const " << kty(sp.t0) << "*)k0p;\n";
  if (sp.n_keys == 2) o << "  const " << kty(sp.t1) << "* __restrict__ k1 = (const " << kty(sp.t1) << "*)k1p;\n";
  emit_col_ptrs(o, sp.cols);
  o << "  const long long tile = 1024LL * " << R << ";\n"
       "  const long long end = ((n + tile - 1) / tile) * tile;  // wave-uniform trip count\n"
       "  for (long long base = (long long)blockIdx.x * tile + threadIdx.x; base < end; base += (long long)gridDim.x * tile) {\n";
  for (int r = 0; r < R; r++) {
    const std::string rs = std::to_string(r), i = "i" + rs, j = "j" + rs;
    o << "    const long long " << i << " = base + " << 1024 * r << ";\n"
      << "    const bool in" << rs << " = " << i << " < n;\n"
      << "    const long long " << j << " = in" << rs << " ? " << i << " : n - 1;  // clamped: loads stay branch-free\n"
      << "    const " << kty(sp.t0) << " ka" << rs << " = __builtin_nontemporal_load(k0 + "
      << (sp.t0 == BOLT_DEC128 ? "2 * " + j : j) << ");  // a Decimal128 key: its low word\n";
    if (sp.n_keys == 2) o << "    const " << kty(sp.t1) << " kb" << rs << " = __builtin_nontemporal_load(k1 + " << j << ");\n";
    for (int c = 0; c < (int)sp.cols.size(); c++) {
      const std::string x = "x" + std::to_string(c) + "_" + rs;
      if (sp.cols[c].dtype == BOLT_DEC128)
        o << "    const unsigned long long " << x << "_lo = c" << c << "[2 * " << j << "], " << x << "_hi = c" << c
          << "[2 * " << j << " + 1];\n";
      else if (sp.cols[c].dtype == BOLT_BOOL)
        o << "    const unsigned char " << x << " = c" << c << "[" << j << "];\n";
      else
        o << "    const " << ctype(sp.cols[c].dtype) << " " << x << " = __builtin_nontemporal_load(c" << c << " + " << j
          << ");\n";
    }
  }
  for (int r = 0; r < R; r++) {
    const std::string rs = std::to_string(r), i = "i" + rs;
    o << "    bool keep" << rs << " = in" << rs << ";\n";
    for (int t = 0; t < (int)sp.terms.size(); t++)
      o << "    keep" << rs << " = keep" << rs << " && " << term_expr(sp.cols, sp.terms[t], t, r, i) << ";\n";
    // 0 drop, 1 regular key, 2 the NULL group (single nullable key)
    if (sp.n_keys == 1) {
      o << "    const u64_t key" << rs << " = " << kcanon(sp.t0, "ka" + rs) << ";\n";
      if (sp.kv0) o << "    const int kind" << rs << " = keep" << rs << " ? (vbit(k0v, " << i << ") ? 1 : 2) : 0;\n";
      else o << "    const int kind" << rs << " = keep" << rs << " ? 1 : 0;\n";
    } else {
      o << "    const u64_t key" << rs << " = (" << kcanon(sp.t0, "ka" + rs) << " << 32) | " << kcanon(sp.t1, "kb" + rs)
        << ";\n";
      std::string kv = "keep" + rs;
      if (sp.kv0) kv += " && vbit(k0v, " + i + ")";
      if (sp.kv1) kv += " && vbit(k1v, " + i + ")";
      o << "    const int kind" << rs << " = (" << kv << ") ? 1 : 0;\n";
    }
  }
  // home slots of all R rows read together (one LDS round trip); the rows
  // found at home, the common case once the table is warm, skip the probe loop
  for (int r = 0; r < R; r++) {
    const std::string rs = std::to_string(r);
    o << "    const u32_t home" << rs << " = gb_lds_home(key" << rs << ", S);\n";
  }
  for (int r = 0; r < R; r++) {
    const std::string rs = std::to_string(r);
    o << "    const u64_t cur" << rs << " = skeys[home" << rs << "];\n";
  }
  for (int r = 0; r < R; r++) {
    const std::string rs = std::to_string(r), i = "i" + rs;
    o << "    {\n"
         "      int slot = -1;\n"
         "      bool spill = false;\n"
         "      if (kind" << rs << " == 2) { slot = S; sused[0] = 1; }\n"
         "      else if (kind" << rs << " == 1) {\n"
         "        if (key" << rs << " == kEmpty) { slot = S + 1; sused[1] = 1; }\n"
         "        else if (cur" << rs << " == key" << rs << ") slot = (int)home" << rs << ";\n"
         "        else {\n"
         "          const bool taken = cur" << rs << " != kEmpty;\n"
         "          const u32_t from = taken ? (home" << rs << " + 1 == (u32_t)S ? 0u : home" << rs << " + 1) : home" << rs << ";\n"
         "          const int lim = gb_lds_probes(S, &sused[3]);\n"
         "          slot = gb_lds_probe(skeys, S, key" << rs << ", from, taken ? lim - 1 : lim, &sused[3]);\n"
         "          spill = slot < 0;\n"
         "        }\n"
         "      }\n"
         "      if (slot >= 0) {\n";
    for (const GbJitSpec::Agg& g : sp.aggs) {
      auto word = [&](int w) { return "swords[" + std::to_string(w) + " * SS + slot]"; };
      if (g.op == BOLT_AGG_COUNT_STAR) {
        o << "        atomicAdd(&" << word(g.w) << ", 1ull);\n";
        continue;
      }
      const std::string x = "x" + std::to_string(g.col) + "_" + rs;
      std::ostringstream u;
      if (g.kind == K_COUNT) {
        u << "atomicAdd(&" << word(g.w) << ", 1ull);";
      } else {
        if (g.nw >= 0) u << "atomicAdd(&" << word(g.nw) << ", 1ull); ";
        switch (g.kind) {
          case K_ISUM:
            // exact without returning atomics: word w sums the low 32 bits
            // (unsigned), word w + 1 the high 32 bits (signed); fewer than
            // 2^32 rows cannot overflow either.  Recombined before the flush.
            u << "{ const long long v_ = (long long)" << x << "; atomicAdd(&" << word(g.w)
              << ", (u64_t)(u32_t)v_); if (v_ >> 32) atomicAdd(&" << word(g.w + 1) << ", (u64_t)(v_ >> 32)); }";
            break;
          case K_DSUM:
            u << "atomic_add192(&" << word(g.w) << ", (u64_t)SS, " << x << "_lo, " << x << "_hi, (u64_t)((long long)"
              << x << "_hi >> 63));";
            break;
          case K_MINMAX: {
            std::string k = (g.dtype == BOLT_F32 || g.dtype == BOLT_F64)
                                ? "f64_order_key((double)" + x + ")"
                                : "((u64_t)(long long)" + x + " ^ 0x8000000000000000ull)";
            u << (g.op == BOLT_AGG_MIN ? "atomicMin(&" : "atomicMax(&") << word(g.w) << ", " << k << ");";
            break;
          }
          default:  // float SUM, AVG, VAR / STDDEV (sum and count here, M2 in a second pass)
            u << "unsafeAtomicAdd((double*)&" << word(g.w) << ", (double)" << x << ");";
            break;
        }
      }
      if (sp.cols[g.col].has_valid) o << "        if (vbit(v" << g.col << ", " << i << ")) { " << u.str() << " }\n";
      else o << "        " << u.str() << "\n";
    }
    o << "      }\n"
         "      gb_spill(spill, " << i << ", key" << rs << ", sused, shll, out);\n"
         "    }\n";
  }
  o << "  }\n";
  // integer sums back to the (lo, hi) 128-bit pair the partial records carry
  bool any_isum = false;
  for (const GbJitSpec::Agg& g : sp.aggs) any_isum |= g.kind == K_ISUM;
  if (any_isum) {
    o << "  __syncthreads();\n"
         "  for (int i = threadIdx.x; i < SS; i += 1024) {\n";
    for (const GbJitSpec::Agg& g : sp.aggs) {
      if (g.kind != K_ISUM) continue;
      const std::string lo = "swords[" + std::to_string(g.w) + " * SS + i]";
      const std::string hi = "swords[" + std::to_string(g.w + 1) + " * SS + i]";
      o << "    { const u64_t u_ = " << lo << "; const long long t_ = (long long)" << hi << ";\n"
        << "      const u64_t lo0 = (u64_t)t_ << 32, lo1 = lo0 + u_;\n"
        << "      " << lo << " = lo1; " << hi << " = (u64_t)(t_ >> 32) + (lo1 < lo0 ? 1ull : 0ull); }\n";
    }
    o << "  }\n";
  }
  o << "  gb_lds_flush(skeys, swords, S, " << W << ", shll, sused, out);\n"
       "}\n";
  return o.str();
}

int jit_row_agg_col(const void* d, const uint8_t* v, int dt, std::vector<JitSpec::Col>* cols, JitCols* jc) {
  for (int c = 0; c < (int)cols->size(); c++)
    if (jc->d[c] == d && jc->v[c] == v && (*cols)[c].dtype == dt) return c;
  if ((int)cols->size() >= kJitMaxCols) return -1;
  const int c = (int)cols->size();
  cols->push_back({dt, v != nullptr});
  jc->d[c] = d;
  jc->v[c] = v;
  return c;
}

int jit_spec_from(const DevFilter& f, const AggList& aggs, bool probe, int key_dtype, bool key_valid, int carry_dtype,
                  JitSpec* sp, JitLits* lits, JitCols* cols) {
  *sp = JitSpec();
  memset(lits, 0, sizeof *lits);
  memset(cols, 0, sizeof *cols);
  sp->probe = probe;
  sp->key_dtype = key_dtype;
  sp->key_valid = key_valid;
  sp->carry_dtype = carry_dtype;
  if (const char* u = getenv("BOLT_JIT_UNROLL")) {
    int v = atoi(u);
    if (v >= 1 && v <= 8) sp->unroll = v;
  }
  if (const char* u = getenv("BOLT_JIT_PROBE_SLOTS")) {
    int v = atoi(u);
    if (v >= 1 && v <= 8) sp->probe_slots = v;
  }
  auto col_index = [&](const void* d, const uint8_t* v, int dt) -> int {
    return jit_row_agg_col(d, v, dt, &sp->cols, cols);
  };
  for (int t = 0; t < f.n; t++) {
    int c = col_index(f.t[t].data, f.t[t].validity, f.t[t].dtype);
    if (c < 0) return set_error(BOLT_ERR_UNSUPPORTED, "jit: too many columns");
    sp->terms.push_back({c, f.t[t].op});
    lits->i[t] = f.t[t].lit_i64;
    lits->f[t] = f.t[t].lit_f64;
    lits->hi[t] = f.t[t].lit_hi;
  }
  for (int a = 0; a < aggs.n; a++) {
    const AggDesc& d = aggs.a[a];
    JitSpec::Agg g{d.op, d.src, -1, d.dtype};
    if (d.op != BOLT_AGG_COUNT_STAR && d.src == 0) {
      g.col = col_index(d.data, d.validity, d.dtype);
      if (g.col < 0) return set_error(BOLT_ERR_UNSUPPORTED, "jit: too many columns");
    }
    if (d.op == BOLT_AGG_COUNT_STAR) g.src = 2;
    sp->aggs.push_back(g);
  }
  return BOLT_OK;
}

int jit_launch_row_agg(hipFunction_t fn, unsigned grid, hipStream_t s, const uint4* tbl, uint64_t mask,
                       const uint32_t* flags, int64_t n, const void* key, const uint8_t* kvalid, const JitLits& lits,
                       const JitCols& cols, Acc* parts, const uint4* bloom, uint64_t bloom_mask) {
  unsigned long long m = mask, bm = bloom_mask;
  long long nn = n;
  JitLits l = lits;
  JitCols c = cols;
  void* args[] = {(void*)&tbl, &m, (void*)&flags, &nn, (void*)&key, (void*)&kvalid, &l, &c, &parts, (void*)&bloom, &bm};
  BOLT_HIP(hipModuleLaunchKernel(fn, grid, 1, 1, 256, 1, 1, 0, s, args, nullptr));
  return BOLT_OK;
}

}  // namespace bolt

extern "C" int bolt_jit_cache_stats(int64_t* compiled, int64_t* disk_hits, int64_t* disk_stores) {
  if (compiled) *compiled = bolt::g_compiled.load();
  if (disk_hits) *disk_hits = bolt::g_disk_hits.load();
  if (disk_stores) *disk_stores = bolt::g_disk_stores.load();
  return BOLT_OK;
}
