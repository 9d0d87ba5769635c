// Graph ingestion (host C++20, std::thread): MatrixMarket coordinate files and
// the "NW" binary edge list / adjacency formats (include/nwgr_io.h).
//
// MatrixMarket: the reference reads the entry section line by line through a
// stringstream (read_mm / mm_fill, io/mmio.hpp:95-217).  Here the file is read
// whole, the entry section is cut into one byte range per thread at line
// starts, every thread parses its lines with a hand-written integer scanner
// into its own arrays, and the per-thread arrays are concatenated in file
// order -- the same edges, in the same order, as the reference produces.

#include <algorithm>
#include <cctype>
#include <climits>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <string>
#include <string_view>
#include <thread>
#include <vector>

#include "host_internal.hpp"
#include "nwgr_io.h"

struct nwgr_edge_list {
  uint32_t n = 0;
  int directedness = NWGR_DIRECTED;
  std::vector<uint32_t> src, dst;
  std::vector<double> values;
  bool has_values = false;
  std::vector<int32_t> ivalues;  // integer attribute (read_mm<..., int>, or the third array of an NW file)
  bool has_ivalues = false;
};

namespace {

using nwgr_host::fail;
using nwgr_host::pick_threads;

// The magic strings of the reference's binary formats, with their NUL
// (edge_list.hpp:190: char[27]; containers/compressed.hpp:62: char[34]).
constexpr char kEdgeListMagic[27] = "NW Graph index_edge_list";
constexpr char kAdjacencyMagic[34] = "NW GRAPH indexed_struct_of_arrays";

bool read_file(const char* path, std::string& buf) {
  std::FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (len < 0) {
    std::fclose(f);
    return false;
  }
  buf.resize(static_cast<size_t>(len));
  const size_t got = len ? std::fread(buf.data(), 1, buf.size(), f) : 0;
  std::fclose(f);
  return got == buf.size();
}

// Line [pos, end-of-line) and the position after its newline.
std::string_view next_line(const std::string& s, size_t& pos) {
  const size_t b = pos;
  size_t e = s.find('\n', b);
  if (e == std::string::npos) e = s.size();
  pos = e < s.size() ? e + 1 : e;
  size_t t = e;
  if (t > b && s[t - 1] == '\r') --t;
  return std::string_view(s.data() + b, t - b);
}

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// Unsigned decimal at p (after blanks); false if none.
inline bool scan_u64(const char*& p, const char* end, uint64_t& v) {
  while (p < end && is_space(*p)) ++p;
  if (p >= end || *p < '0' || *p > '9') return false;
  uint64_t x = 0;
  while (p < end && *p >= '0' && *p <= '9') {
    x = x * 10 + static_cast<uint64_t>(*p - '0');
    ++p;
  }
  v = x;
  return true;
}

struct Part {
  std::vector<uint32_t> src, dst;
  std::vector<double> val;
  std::vector<int32_t> ival;
  uint64_t max_id = 0;
  uint64_t lines = 0;
  std::string err;
};

// The integer attribute `stringstream >> int` reads from a value field: an
// optional sign and the digits after it (0 when there are none, clamped to
// the int range), 1 for a pattern file (mm_fill's `T v(1.0)`, io/mmio.hpp:109).
inline int32_t int_prefix(const char* q) {
  char* stop = nullptr;
  errno = 0;
  const long long v = std::strtoll(q, &stop, 10);
  if (stop == q) return 0;
  return static_cast<int32_t>(std::clamp<long long>(v, INT32_MIN, INT32_MAX));
}

// Parses the entry lines of s[b, e) (whole lines).  `mirror`: a symmetric
// file's off-diagonal entry also gives the reversed edge (directed lists);
// `ints`: also keep the integer attribute of every entry.
void parse_range(const std::string& s, size_t b, size_t e, bool mirror, bool pattern, bool ints, Part& out) {
  const char* p = s.data() + b;
  const char* end = s.data() + e;
  std::string tmp;
  while (p < end) {
    const char* eol = static_cast<const char*>(std::memchr(p, '\n', static_cast<size_t>(end - p)));
    if (!eol) eol = end;
    const char* q = p;
    while (q < eol && is_space(*q)) ++q;
    if (q == eol || *q == '%') {  // blank or comment line
      p = eol + 1;
      continue;
    }
    uint64_t i = 0, j = 0;
    if (!scan_u64(q, eol, i) || !scan_u64(q, eol, j) || i == 0 || j == 0) {
      out.err = "malformed MatrixMarket entry line: " + std::string(p, static_cast<size_t>(eol - p));
      return;
    }
    if (i > 0xffffffffULL || j > 0xffffffffULL) {
      out.err = "vertex id exceeds the uint32 range";
      return;
    }
    double v = 1.0;
    int32_t iv = 1;
    if (!pattern) {
      while (q < eol && is_space(*q)) ++q;
      tmp.assign(q, static_cast<size_t>(eol - q));
      char* stop = nullptr;
      errno = 0;
      v = std::strtod(tmp.c_str(), &stop);
      if (stop == tmp.c_str()) {
        out.err = "MatrixMarket entry without a value: " + std::string(p, static_cast<size_t>(eol - p));
        return;
      }
      if (ints) iv = int_prefix(tmp.c_str());
    }
    out.src.push_back(static_cast<uint32_t>(i - 1));
    out.dst.push_back(static_cast<uint32_t>(j - 1));
    if (!pattern) out.val.push_back(v);
    if (ints) out.ival.push_back(iv);
    if (mirror && i != j) {
      out.src.push_back(static_cast<uint32_t>(j - 1));
      out.dst.push_back(static_cast<uint32_t>(i - 1));
      if (!pattern) out.val.push_back(v);
      if (ints) out.ival.push_back(iv);
    }
    out.max_id = std::max({out.max_id, i, j});
    ++out.lines;
    p = eol + 1;
  }
}

}  // namespace

extern "C" {

int nwgr_read_mm(const char* path, int nthreads, nwgr_edge_list** out) {
  return nwgr_read_mm_as(path, nthreads, NWGR_DIRECTED, 0, out);
}

int nwgr_read_mm_as(const char* path, int nthreads, int directedness, int flags, nwgr_edge_list** out) {
  if (!path || !out) return fail("null argument");
  if (directedness != NWGR_DIRECTED && directedness != NWGR_UNDIRECTED) return fail("bad directedness");
  if (flags & ~NWGR_MM_INT_ATTRIBUTE) return fail("unknown read_mm flags");
  const bool ints = (flags & NWGR_MM_INT_ATTRIBUTE) != 0;
  *out = nullptr;
  try {
    std::string s;
    if (!read_file(path, s)) return fail(std::string("cannot read ") + path);
    size_t pos = 0;
    // header: %%MatrixMarket matrix coordinate <field> <symmetry>
    std::string head(next_line(s, pos));
    char w[5][64] = {};
    if (std::sscanf(head.c_str(), "%63s %63s %63s %63s %63s", w[0], w[1], w[2], w[3], w[4]) != 5 ||
        std::strcmp(w[0], "%%MatrixMarket") != 0)
      return fail("unsupported format (no %%MatrixMarket header)");
    for (auto& word : w)
      for (char* c = word; *c; ++c) *c = static_cast<char>(std::tolower(static_cast<unsigned char>(*c)));
    if (std::strcmp(w[1], "matrix") != 0 || std::strcmp(w[2], "coordinate") != 0)
      return fail("only 'matrix coordinate' MatrixMarket files are supported");
    const bool pattern = std::strcmp(w[3], "pattern") == 0;
    if (!pattern && std::strcmp(w[3], "real") != 0 && std::strcmp(w[3], "integer") != 0 &&
        std::strcmp(w[3], "double") != 0)
      return fail(std::string("unsupported MatrixMarket field: ") + w[3]);
    bool symmetric;
    if (std::strcmp(w[4], "symmetric") == 0)
      symmetric = true;
    else if (std::strcmp(w[4], "general") == 0)
      symmetric = false;
    else
      return fail(std::string("bad MatrixMarket symmetry: ") + w[4]);
    // comments, then the size line
    std::string_view line;
    do {
      if (pos >= s.size()) return fail("MatrixMarket file ends before its size line");
      line = next_line(s, pos);
    } while (!line.empty() && line[0] == '%');
    uint64_t n0 = 0, n1 = 0, nnz = 0;
    {
      const char* q = line.data();
      const char* e = q + line.size();
      if (!scan_u64(q, e, n0) || !scan_u64(q, e, n1) || !scan_u64(q, e, nnz))
        return fail("malformed MatrixMarket size line");
    }
    if (n0 > 0xffffffffULL || n1 > 0xffffffffULL) return fail("vertex count exceeds the uint32 range");
    // entry section [pos, size) cut at line starts
    const int T = std::max(1, std::min<int>(pick_threads(nthreads), static_cast<int>((s.size() - pos) / 65536 + 1)));
    std::vector<size_t> cut(static_cast<size_t>(T) + 1, s.size());
    cut[0] = pos;
    for (int t = 1; t < T; ++t) {
      size_t c = pos + (s.size() - pos) * static_cast<size_t>(t) / static_cast<size_t>(T);
      c = std::max(c, cut[t - 1]);
      while (c < s.size() && c > pos && s[c - 1] != '\n') ++c;
      cut[t] = c;
    }
    std::vector<Part> parts(static_cast<size_t>(T));
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
      th.emplace_back([&, t] {
        parse_range(s, cut[t], cut[t + 1], symmetric && directedness == NWGR_DIRECTED, pattern, ints, parts[t]);
      });
    for (auto& x : th) x.join();
    uint64_t lines = 0, m = 0, max_id = 0;
    for (auto& pt : parts) {
      if (!pt.err.empty()) return fail(pt.err);
      lines += pt.lines;
      m += pt.src.size();
      max_id = std::max(max_id, pt.max_id);
    }
    if (lines != nnz)
      return fail("MatrixMarket file has " + std::to_string(lines) + " entries, its size line says " +
                  std::to_string(nnz));
    if (m > 0xffffffffULL) return fail("edge count exceeds the uint32 offset range");
    auto el = std::make_unique<nwgr_edge_list>();
    el->n = static_cast<uint32_t>(std::max<uint64_t>(n0, max_id));  // ids are 1-based: max_id = largest id + 1
    el->directedness = directedness;
    el->has_values = !pattern;
    el->has_ivalues = ints;
    el->src.reserve(m);
    el->dst.reserve(m);
    if (!pattern) el->values.reserve(m);
    if (ints) el->ivalues.reserve(m);
    for (auto& pt : parts) {
      el->src.insert(el->src.end(), pt.src.begin(), pt.src.end());
      el->dst.insert(el->dst.end(), pt.dst.begin(), pt.dst.end());
      if (!pattern) el->values.insert(el->values.end(), pt.val.begin(), pt.val.end());
      if (ints) el->ivalues.insert(el->ivalues.end(), pt.ival.begin(), pt.ival.end());
    }
    (void)n1;
    *out = el.release();
    return 0;
  } catch (const std::bad_alloc&) {
    return fail("host allocation failed");
  }
}

int nwgr_write_nw_edge_list(const char* path, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                            int directedness) {
  return nwgr_write_nw_edge_list_attr(path, n, m, src, dst, nullptr, directedness);
}

int nwgr_write_nw_edge_list_attr(const char* path, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                                 const int32_t* attr, int directedness) {
  if (!path || (m && (!src || !dst))) return fail("null argument");
  if (directedness != NWGR_DIRECTED && directedness != NWGR_UNDIRECTED) return fail("bad directedness");
  std::ofstream f(path, std::ios::binary);
  if (!f) return fail(std::string("cannot write ") + path);
  const int32_t d = directedness;
  const uint64_t N = n, st = m, arrays = attr ? 3 : 2, el = sizeof(uint32_t);
  const bool is_open = false;
  f.write(kEdgeListMagic, sizeof(kEdgeListMagic));
  f.write(reinterpret_cast<const char*>(&d), sizeof(d));
  f.write(reinterpret_cast<const char*>(&N), sizeof(N));
  f.write(reinterpret_cast<const char*>(&is_open), sizeof(is_open));
  f.write(reinterpret_cast<const char*>(&st), sizeof(st));
  f.write(reinterpret_cast<const char*>(&arrays), sizeof(arrays));
  for (const void* a : {static_cast<const void*>(src), static_cast<const void*>(dst), static_cast<const void*>(attr)}) {
    if (!a) continue;
    f.write(reinterpret_cast<const char*>(&st), sizeof(st));
    f.write(reinterpret_cast<const char*>(&el), sizeof(el));
    if (m) f.write(static_cast<const char*>(a), static_cast<std::streamsize>(m * sizeof(uint32_t)));
  }
  return f ? 0 : fail(std::string("write failed: ") + path);
}

int nwgr_read_nw_edge_list(const char* path, nwgr_edge_list** out) {
  if (!path || !out) return fail("null argument");
  *out = nullptr;
  try {
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(std::string("cannot read ") + path);
    char magic[sizeof(kEdgeListMagic)];
    int32_t d = 0;
    uint64_t N = 0, st = 0, arrays = 0;
    bool is_open = false;
    f.read(magic, sizeof(magic));
    if (!f || std::memcmp(magic, kEdgeListMagic, sizeof(magic)) != 0) return fail("not an NW index_edge_list file");
    f.read(reinterpret_cast<char*>(&d), sizeof(d));
    f.read(reinterpret_cast<char*>(&N), sizeof(N));
    f.read(reinterpret_cast<char*>(&is_open), sizeof(is_open));
    f.read(reinterpret_cast<char*>(&st), sizeof(st));
    f.read(reinterpret_cast<char*>(&arrays), sizeof(arrays));
    // the reference reads as many arrays as its own edge list type has and
    // ignores the count (soa.hpp:302-309): ids first, then attributes
    if (!f || arrays < 2) return fail("NW edge list: fewer than two arrays");
    if (N > 0xffffffffULL || st > 0xffffffffULL) return fail("NW edge list exceeds the uint32 id/offset range");
    auto el = std::make_unique<nwgr_edge_list>();
    el->n = static_cast<uint32_t>(N);
    el->directedness = d;
    for (std::vector<uint32_t>* a : {&el->src, &el->dst}) {
      uint64_t len = 0, esz = 0;
      f.read(reinterpret_cast<char*>(&len), sizeof(len));
      f.read(reinterpret_cast<char*>(&esz), sizeof(esz));
      if (!f || len != st || esz != sizeof(uint32_t)) return fail("NW edge list: unexpected array header");
      a->resize(len);
      if (len) f.read(reinterpret_cast<char*>(a->data()), static_cast<std::streamsize>(len * esz));
      if (!f) return fail("NW edge list: file truncated");
    }
    if (arrays >= 3) {
      // a 4-byte attribute (process_edge_list writes int weights) is kept;
      // other attribute arrays are skipped
      uint64_t len = 0, esz = 0;
      f.read(reinterpret_cast<char*>(&len), sizeof(len));
      f.read(reinterpret_cast<char*>(&esz), sizeof(esz));
      if (!f || len != st) return fail("NW edge list: unexpected attribute array header");
      if (esz == sizeof(int32_t)) {
        el->ivalues.resize(len);
        if (len) f.read(reinterpret_cast<char*>(el->ivalues.data()), static_cast<std::streamsize>(len * esz));
        el->has_ivalues = true;
      } else {
        f.seekg(static_cast<std::streamoff>(len * esz), std::ios::cur);
      }
      if (!f) return fail("NW edge list: file truncated");
    }
    for (uint64_t k = 0; k < st; ++k)
      if (el->src[k] >= el->n || el->dst[k] >= el->n) return fail("NW edge list: vertex id >= vertex count");
    *out = el.release();
    return 0;
  } catch (const std::bad_alloc&) {
    return fail("host allocation failed");
  }
}

int nwgr_write_nw_adjacency(const char* path, uint32_t n, uint64_t m, const uint32_t* offsets,
                            const uint32_t* indices) {
  if (!path || !offsets || (m && !indices)) return fail("null argument");
  if (offsets[n] != m) return fail("offsets[n] != m");
  std::ofstream f(path, std::ios::binary);
  if (!f) return fail(std::string("cannot write ") + path);
  const uint64_t N = n, n1 = static_cast<uint64_t>(n) + 1, st = m, arrays = 1, el = sizeof(uint32_t);
  f.write(kAdjacencyMagic, sizeof(kAdjacencyMagic));
  f.write(reinterpret_cast<const char*>(&N), sizeof(N));
  f.write(reinterpret_cast<const char*>(&n1), sizeof(n1));
  f.write(reinterpret_cast<const char*>(&el), sizeof(el));
  f.write(reinterpret_cast<const char*>(offsets), static_cast<std::streamsize>(n1 * sizeof(uint32_t)));
  // to_be_indexed_: a struct of one array (soa.hpp:289-295)
  f.write(reinterpret_cast<const char*>(&st), sizeof(st));
  f.write(reinterpret_cast<const char*>(&arrays), sizeof(arrays));
  f.write(reinterpret_cast<const char*>(&st), sizeof(st));
  f.write(reinterpret_cast<const char*>(&el), sizeof(el));
  if (m) f.write(reinterpret_cast<const char*>(indices), static_cast<std::streamsize>(m * sizeof(uint32_t)));
  return f ? 0 : fail(std::string("write failed: ") + path);
}

int nwgr_read_nw_adjacency(const char* path, nwgr_host_csr** out) {
  if (!path || !out) return fail("null argument");
  *out = nullptr;
  try {
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(std::string("cannot read ") + path);
    char magic[sizeof(kAdjacencyMagic)];
    uint64_t N = 0, n1 = 0, el = 0, st = 0, arrays = 0, len = 0, esz = 0;
    f.read(magic, sizeof(magic));
    if (!f || std::memcmp(magic, kAdjacencyMagic, sizeof(magic)) != 0)
      return fail("not an NW indexed_struct_of_arrays file");
    f.read(reinterpret_cast<char*>(&N), sizeof(N));
    f.read(reinterpret_cast<char*>(&n1), sizeof(n1));
    f.read(reinterpret_cast<char*>(&el), sizeof(el));
    if (!f || el != sizeof(uint32_t) || n1 == 0 || n1 - 1 > 0xffffffffULL)
      return fail("NW adjacency: only uint32 offsets are supported");
    auto g = std::make_unique<nwgr_host_csr>();
    g->n = static_cast<uint32_t>(n1 - 1);
    g->offsets.resize(n1);
    f.read(reinterpret_cast<char*>(g->offsets.data()), static_cast<std::streamsize>(n1 * el));
    f.read(reinterpret_cast<char*>(&st), sizeof(st));
    f.read(reinterpret_cast<char*>(&arrays), sizeof(arrays));
    f.read(reinterpret_cast<char*>(&len), sizeof(len));
    f.read(reinterpret_cast<char*>(&esz), sizeof(esz));
    if (!f || arrays != 1 || len != st || esz != sizeof(uint32_t) || g->offsets[g->n] != st)
      return fail("NW adjacency: only one uint32 index array without attributes is supported");
    g->m = st;
    g->indices.resize(st);
    if (st) f.read(reinterpret_cast<char*>(g->indices.data()), static_cast<std::streamsize>(st * esz));
    if (!f) return fail("NW adjacency: file truncated");
    if (g->offsets[0] != 0) return fail("NW adjacency: offsets[0] != 0");
    for (uint32_t v = 0; v < g->n; ++v)
      if (g->offsets[v + 1] < g->offsets[v]) return fail("NW adjacency: offsets decrease");
    g->outdeg.assign(g->n, 0);
    for (uint64_t k = 0; k < st; ++k) {
      if (g->indices[k] >= g->n) return fail("NW adjacency: index >= vertex count");
      ++g->outdeg[g->indices[k]];
    }
    (void)N;
    *out = g.release();
    return 0;
  } catch (const std::bad_alloc&) {
    return fail("host allocation failed");
  }
}

uint32_t nwgr_edge_list_n(const nwgr_edge_list* e) { return e ? e->n : 0; }
uint64_t nwgr_edge_list_m(const nwgr_edge_list* e) { return e ? e->src.size() : 0; }
int nwgr_edge_list_directedness(const nwgr_edge_list* e) { return e ? e->directedness : -1; }
const uint32_t* nwgr_edge_list_src(const nwgr_edge_list* e) { return e ? e->src.data() : nullptr; }
const uint32_t* nwgr_edge_list_dst(const nwgr_edge_list* e) { return e ? e->dst.data() : nullptr; }
const double* nwgr_edge_list_values(const nwgr_edge_list* e) {
  return e && e->has_values ? e->values.data() : nullptr;
}
const int32_t* nwgr_edge_list_int_values(const nwgr_edge_list* e) {
  return e && e->has_ivalues ? e->ivalues.data() : nullptr;
}
void nwgr_edge_list_free(nwgr_edge_list* e) { delete e; }

}  // extern "C"
