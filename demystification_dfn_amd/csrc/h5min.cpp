// Minimal HDF5 reader for the reference's input files, without libhdf5.
//
// The reference reads its configuration with the HDF5 Fortran library
// (src/configuration.f90:70-379): scalar and 1-D/2-D FP64 and int64 datasets and
// variable-length strings, in groups "hyperparameter", "material_loading" and
// "faultN".  The files its Python generator writes use the original layout of
// the format: superblock version 0, groups as symbol tables (version-1 B-trees
// of symbol-table nodes, names in local heaps), version-1 object headers,
// contiguous or compact storage, and variable-length strings in global heaps.
// This reader implements exactly that subset, from the published HDF5 file-format
// specification; anything else (chunked or filtered storage, version-2 object
// headers, other superblock versions) is reported as unsupported.
//
// The writer produces the reference's output files (src/solver.f90:227-317):
// FP64 datasets of rank 1 or 2 in the root group, in the same subset (superblock
// 0, one symbol-table node under a version-1 B-tree, a local heap of names,
// version-1 object headers with dataspace, IEEE little-endian FP64 datatype,
// fill-value and contiguous layout messages), encoded as libhdf5 encodes them.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/dfn.h"

namespace {

constexpr uint64_t kUndef = ~0ull;

struct Dataset {
  int cls = -1;        // 0 integer, 1 floating point, 2 string
  int size = 0;        // bytes per element (16 for variable-length references)
  bool vlen = false;
  bool is_signed = true;
  bool big_endian = false;
  std::vector<uint64_t> dims;
  int layout = -1;     // 0 compact, 1 contiguous
  uint64_t addr = kUndef, nbytes = 0;
  std::vector<uint8_t> compact;
  bool filtered = false;
};

// Thrown by the reader's byte accessors for a read outside the file; the entry
// points turn it into DFN_ERR_UNSUPPORTED (a truncated or corrupt file).
struct H5Malformed {};

struct H5 {
  std::vector<uint8_t> d;
  uint64_t root_oh = 0;
  std::string err;

  bool in(uint64_t off, uint64_t n) const { return off <= d.size() && n <= d.size() - off; }
  uint64_t u(uint64_t off, int n) const {  // little-endian unsigned of n bytes
    if (!in(off, (uint64_t)n)) throw H5Malformed{};
    uint64_t v = 0;
    for (int i = n - 1; i >= 0; i--) v = (v << 8) | d[off + i];
    return v;
  }
  uint8_t b(uint64_t off) const {  // one byte
    if (off >= d.size()) throw H5Malformed{};
    return d[off];
  }

  int open(const char *path) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return DFN_ERR_ARG;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 96) { std::fclose(f); return DFN_ERR_UNSUPPORTED; }
    d.resize((size_t)n);
    const size_t got = std::fread(d.data(), 1, d.size(), f);
    std::fclose(f);
    if (got != d.size()) return DFN_ERR_ARG;
    static const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    if (std::memcmp(d.data(), sig, 8) != 0) return DFN_ERR_UNSUPPORTED;
    // superblock version 0: sizes of offsets and lengths 8, root symbol-table entry at 56
    if (d[8] != 0 || d[13] != 8 || d[14] != 8) return DFN_ERR_UNSUPPORTED;
    root_oh = u(56 + 8, 8);
    return in(root_oh, 16) ? DFN_OK : DFN_ERR_UNSUPPORTED;
  }

  // messages of a version-1 object header (following continuation blocks)
  bool messages(uint64_t oh, std::vector<std::pair<int, uint64_t>> &out, std::vector<uint16_t> &sizes) const {
    if (!in(oh, 16) || b(oh) != 1) return false;
    const int nmsg = (int)u(oh + 2, 2);
    std::vector<std::pair<uint64_t, uint64_t>> blocks{{oh + 16, u(oh + 8, 4)}};
    int seen = 0;
    for (size_t bi = 0; bi < blocks.size() && seen < nmsg; bi++) {
      uint64_t p = blocks[bi].first;
      const uint64_t end = blocks[bi].first + blocks[bi].second;
      while (p + 8 <= end && seen < nmsg) {
        if (!in(p, 8)) return false;
        const int type = (int)u(p, 2);
        const uint16_t sz = (uint16_t)u(p + 2, 2);
        if (!in(p + 8, sz)) return false;
        if (type == 0x10) blocks.push_back({u(p + 8, 8), u(p + 16, 8)});  // continuation
        out.push_back({type, p + 8});
        sizes.push_back(sz);
        seen++;
        p += 8 + sz;
      }
    }
    return true;
  }

  // object header address of `name` in the group whose symbol table is (btree, heap)
  uint64_t group_find(uint64_t btree, uint64_t heap, const std::string &name, int depth = 0) const {
    if (depth > 32 || !in(btree, 24) || std::memcmp(&d[btree], "TREE", 4) != 0 || b(btree + 4) != 0) return kUndef;
    if (!in(heap, 32) || std::memcmp(&d[heap], "HEAP", 4) != 0) return kUndef;
    const uint64_t hdata = u(heap + 24, 8), hsize = u(heap + 8, 8);
    const int level = b(btree + 5), used = (int)u(btree + 6, 2);
    for (int e = 0; e < used; e++) {
      const uint64_t child = u(btree + 24 + 8 + (uint64_t)e * 16, 8);
      if (level > 0) {
        const uint64_t r = group_find(child, heap, name, depth + 1);
        if (r != kUndef) return r;
        continue;
      }
      if (!in(child, 8) || std::memcmp(&d[child], "SNOD", 4) != 0) return kUndef;
      const int nsym = (int)u(child + 6, 2);
      for (int s = 0; s < nsym; s++) {
        const uint64_t ent = child + 8 + (uint64_t)s * 40;
        if (!in(ent, 40)) return kUndef;
        const uint64_t noff = u(ent, 8);
        if (noff >= hsize || !in(hdata + noff, 1)) continue;
        const char *nm = reinterpret_cast<const char *>(&d[hdata + noff]);
        const size_t maxlen = (size_t)(hsize - noff);
        if (strnlen(nm, maxlen) == name.size() && std::memcmp(nm, name.data(), name.size()) == 0) return u(ent + 8, 8);
      }
    }
    return kUndef;
  }

  uint64_t resolve(const std::string &path) const {
    uint64_t oh = root_oh;
    size_t pos = 0;
    while (pos < path.size()) {
      size_t nxt = path.find('/', pos);
      if (nxt == std::string::npos) nxt = path.size();
      const std::string comp = path.substr(pos, nxt - pos);
      pos = nxt + 1;
      if (comp.empty()) continue;
      std::vector<std::pair<int, uint64_t>> ms;
      std::vector<uint16_t> sz;
      if (!messages(oh, ms, sz)) return kUndef;
      uint64_t bt = kUndef, hp = kUndef;
      for (auto &m : ms)
        if (m.first == 0x11) { bt = u(m.second, 8); hp = u(m.second + 8, 8); }
      if (bt == kUndef) return kUndef;  // not a group
      oh = group_find(bt, hp, comp);
      if (oh == kUndef) return kUndef;
    }
    return oh;
  }

  int dataset(const std::string &path, Dataset &ds) const {
    const uint64_t oh = resolve(path);
    if (oh == kUndef) return DFN_ERR_ARG;
    std::vector<std::pair<int, uint64_t>> ms;
    std::vector<uint16_t> sz;
    if (!messages(oh, ms, sz)) return DFN_ERR_UNSUPPORTED;
    bool have_space = false, have_type = false;
    for (size_t i = 0; i < ms.size(); i++) {
      const uint64_t p = ms[i].second, end = p + sz[i];  // every field read lies in [p, end)
      const auto fits = [&](uint64_t q, uint64_t n) { return q >= p && q <= end && n <= end - q; };
      switch (ms[i].first) {
        case 0x01: {  // dataspace
          if (!fits(p, 4)) return DFN_ERR_UNSUPPORTED;
          const int ver = b(p), rank = b(p + 1);
          const uint64_t q = ver == 1 ? p + 8 : p + 4;
          if (rank > 32 || !fits(q, 8ull * rank)) return DFN_ERR_UNSUPPORTED;
          ds.dims.clear();
          for (int r = 0; r < rank; r++) ds.dims.push_back(u(q + 8ull * r, 8));
          have_space = true;
          break;
        }
        case 0x03: {  // datatype
          if (!fits(p, 8)) return DFN_ERR_UNSUPPORTED;
          const int cls = b(p) & 0x0f;
          const uint32_t bits = (uint32_t)u(p + 1, 3);
          const uint64_t size = u(p + 4, 4);
          if (cls == 0) { ds.cls = 0; ds.big_endian = bits & 1; ds.is_signed = (bits >> 3) & 1; }
          else if (cls == 1) { ds.cls = 1; ds.big_endian = bits & 1; }
          else if (cls == 3) { ds.cls = 2; }
          else if (cls == 9 && (bits & 0x0f) == 1) { ds.cls = 2; ds.vlen = true; }
          else return DFN_ERR_UNSUPPORTED;
          // numbers of 1 to 8 bytes (floats of 4 or 8), fixed strings of 1 byte to 1 MiB,
          // variable-length strings as 16-byte heap references
          if (ds.cls == 0 && (size < 1 || size > 8)) return DFN_ERR_UNSUPPORTED;
          if (ds.cls == 1 && size != 4 && size != 8) return DFN_ERR_UNSUPPORTED;
          if (ds.cls == 2 && !ds.vlen && (size < 1 || size > (1u << 20))) return DFN_ERR_UNSUPPORTED;
          if (ds.vlen && size != 16) return DFN_ERR_UNSUPPORTED;
          ds.size = (int)size;
          have_type = true;
          break;
        }
        case 0x08: {  // data layout
          if (!fits(p, 4)) return DFN_ERR_UNSUPPORTED;
          const int ver = b(p);
          if (ver == 3) {
            ds.layout = b(p + 1);
            if (ds.layout == 0) {
              const uint64_t n = u(p + 2, 2);
              if (!fits(p + 4, n)) return DFN_ERR_UNSUPPORTED;
              ds.compact.assign(d.begin() + (long)(p + 4), d.begin() + (long)(p + 4 + n));
            } else if (ds.layout == 1) {
              if (!fits(p + 2, 16)) return DFN_ERR_UNSUPPORTED;
              ds.addr = u(p + 2, 8);
              ds.nbytes = u(p + 10, 8);
            } else {
              return DFN_ERR_UNSUPPORTED;  // chunked
            }
          } else if (ver == 1 || ver == 2) {
            const int rank = b(p + 1);
            ds.layout = b(p + 2);
            uint64_t q = p + 8;
            if (ds.layout == 1) {
              if (!fits(q, 8)) return DFN_ERR_UNSUPPORTED;
              ds.addr = u(q, 8);
              q += 8;
            } else if (ds.layout != 0) {
              return DFN_ERR_UNSUPPORTED;
            }
            q += 4ull * rank;
            if (ds.layout == 0) {
              if (!fits(q, 4)) return DFN_ERR_UNSUPPORTED;
              const uint64_t n = u(q, 4);
              if (!fits(q + 4, n)) return DFN_ERR_UNSUPPORTED;
              ds.compact.assign(d.begin() + (long)(q + 4), d.begin() + (long)(q + 4 + n));
            }
          } else {
            return DFN_ERR_UNSUPPORTED;
          }
          break;
        }
        case 0x0B:  // filter pipeline
          ds.filtered = true;
          break;
        default:
          break;
      }
    }
    if (!have_space || !have_type || ds.layout < 0) return DFN_ERR_UNSUPPORTED;
    if (ds.filtered) return DFN_ERR_UNSUPPORTED;
    uint64_t n = 0;
    if (!count(ds, n)) return DFN_ERR_UNSUPPORTED;  // overflowing extents
    uint64_t bytes = 0;
    if (__builtin_mul_overflow(n, (uint64_t)ds.size, &bytes)) return DFN_ERR_UNSUPPORTED;
    return DFN_OK;
  }

  // element count (false if the product of the extents overflows)
  static bool count(const Dataset &ds, uint64_t &n) {
    n = 1;
    for (uint64_t x : ds.dims)
      if (__builtin_mul_overflow(n, x, &n)) return false;
    return true;
  }
  static uint64_t count(const Dataset &ds) {
    uint64_t n = 0;
    return count(ds, n) ? n : 0;
  }

  // raw bytes of the dataset (count * size); unwritten contiguous data reads as zeros.
  // The bytes are checked to lie in the file before anything is allocated.
  bool raw(const Dataset &ds, std::vector<uint8_t> &out) const {
    uint64_t c = 0, n = 0;
    if (!count(ds, c) || __builtin_mul_overflow(c, (uint64_t)ds.size, &n)) return false;
    out.clear();
    if (n == 0) return true;  // an empty dataset
    if (ds.layout == 0) {
      if (ds.compact.size() < n) return false;
      out.assign(ds.compact.begin(), ds.compact.begin() + (long)n);
      return true;
    }
    if (ds.addr == kUndef) {
      if (n > d.size() + (1ull << 20)) return false;  // unwritten data larger than any real input
      out.assign(n, 0);
      return true;
    }
    if (!in(ds.addr, n)) return false;
    out.assign(d.begin() + (long)ds.addr, d.begin() + (long)(ds.addr + n));
    return true;
  }

  double number(const Dataset &ds, const uint8_t *p) const {
    if (ds.big_endian) return NAN;
    if (ds.cls == 1) {
      if (ds.size == 8) { double v; std::memcpy(&v, p, 8); return v; }
      if (ds.size == 4) { float v; std::memcpy(&v, p, 4); return v; }
      return NAN;
    }
    uint64_t v = 0;
    for (int i = ds.size - 1; i >= 0; i--) v = (v << 8) | p[i];
    if (ds.is_signed && ds.size < 8 && (v >> (8 * ds.size - 1)) & 1) v |= ~0ull << (8 * ds.size);
    return ds.is_signed ? (double)(int64_t)v : (double)v;
  }

  int64_t integer(const Dataset &ds, const uint8_t *p) const {
    if (ds.cls == 1) return (int64_t)number(ds, p);
    uint64_t v = 0;
    for (int i = ds.size - 1; i >= 0; i--) v = (v << 8) | p[i];
    if (ds.is_signed && ds.size < 8 && (v >> (8 * ds.size - 1)) & 1) v |= ~0ull << (8 * ds.size);
    return (int64_t)v;
  }

  // the first string of a string dataset
  bool string(const Dataset &ds, std::string &s) const {
    std::vector<uint8_t> r;
    if (!raw(ds, r) || r.empty() || (ds.vlen && r.size() < 16)) return false;
    if (!ds.vlen) {
      s.assign(reinterpret_cast<const char *>(r.data()), strnlen(reinterpret_cast<const char *>(r.data()), ds.size));
      return true;
    }
    // variable length: length (4), global heap collection address (8), object index (4)
    const uint64_t len = (uint64_t)r[0] | ((uint64_t)r[1] << 8) | ((uint64_t)r[2] << 16) | ((uint64_t)r[3] << 24);
    uint64_t col = 0;
    for (int i = 7; i >= 0; i--) col = (col << 8) | r[4 + i];
    const uint32_t idx = (uint32_t)r[12] | ((uint32_t)r[13] << 8) | ((uint32_t)r[14] << 16) | ((uint32_t)r[15] << 24);
    if (!in(col, 16) || std::memcmp(&d[col], "GCOL", 4) != 0) return false;
    const uint64_t csize = u(col + 8, 8);
    uint64_t p = col + 16;
    while (p + 16 <= col + csize && in(p, 16)) {
      const uint32_t oi = (uint32_t)u(p, 2);
      const uint64_t osz = u(p + 8, 8);
      if (oi == 0) break;
      if (oi == idx) {
        if (!in(p + 16, len)) return false;
        s.assign(reinterpret_cast<const char *>(&d[p + 16]), (size_t)std::min(len, osz));
        return true;
      }
      p += 16 + ((osz + 7) / 8) * 8;
    }
    return false;
  }
};

// ---------------------------------------------------------------------------
// Writer
struct Writer {
  std::vector<uint8_t> b;
  static constexpr uint64_t kLeafK = 4, kInternalK = 16;  // symbol-table node and group B-tree ranks

  uint64_t alloc(uint64_t n) {  // 8-byte aligned block of n zero bytes
    const uint64_t off = (b.size() + 7) & ~7ull;
    b.resize(off + n, 0);
    return off;
  }
  void put(uint64_t off, uint64_t v, int n) {
    for (int i = 0; i < n; i++) b[off + i] = (uint8_t)(v >> (8 * i));
  }
  // one object-header message: type, size (a multiple of 8), flags
  static void message(std::vector<uint8_t> &m, int type, const std::vector<uint8_t> &body, int flags) {
    const size_t sz = (body.size() + 7) & ~size_t(7);
    const size_t at = m.size();
    m.resize(at + 8 + sz, 0);
    m[at] = (uint8_t)type;
    m[at + 1] = (uint8_t)(type >> 8);
    m[at + 2] = (uint8_t)sz;
    m[at + 3] = (uint8_t)(sz >> 8);
    m[at + 4] = (uint8_t)flags;
    if (!body.empty()) std::copy(body.begin(), body.end(), m.begin() + (long)(at + 8));
  }
  static void le(std::vector<uint8_t> &v, uint64_t x, int n) {
    for (int i = 0; i < n; i++) v.push_back((uint8_t)(x >> (8 * i)));
  }
  // version-1 object header holding the given messages
  uint64_t object_header(const std::vector<uint8_t> &msgs, int nmsg) {
    const uint64_t oh = alloc(16 + msgs.size());
    b[oh] = 1;                      // version
    put(oh + 2, (uint64_t)nmsg, 2);
    put(oh + 4, 1, 4);              // reference count
    put(oh + 8, msgs.size(), 4);    // bytes of message data
    std::memcpy(&b[oh + 16], msgs.data(), msgs.size());
    return oh;
  }
};

}  // namespace

// FP64 datasets (rank 1 or 2, C order) as the root group's members; names must be
// distinct.  Returns DFN_ERR_ARG for bad arguments, DFN_ERR_UNSUPPORTED for more
// datasets than one symbol-table node holds (8) or a name over 255 bytes.
static int h5_write_f64(const char *path, int n, const char *const *names, const int32_t *ranks,
                        const int64_t *dims, const double *const *data) {
  if (!path || n < 1 || !names || !ranks || !dims || !data) return DFN_ERR_ARG;
  if (n > (int)(2 * Writer::kLeafK)) return DFN_ERR_UNSUPPORTED;
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) {
    if (!names[i] || !data[i] || ranks[i] < 1 || ranks[i] > 2) return DFN_ERR_ARG;
    const size_t len = strnlen(names[i], 256);
    if (len == 0 || len > 255) return DFN_ERR_UNSUPPORTED;
    for (int r = 0; r < ranks[i]; r++)
      if (dims[2 * i + r] < 0) return DFN_ERR_ARG;
    for (int j = 0; j < i; j++)
      if (std::strcmp(names[i], names[j]) == 0) return DFN_ERR_ARG;
    order[i] = i;
  }
  // symbol-table entries are kept in name order
  std::sort(order.begin(), order.end(), [&](int x, int y) { return std::strcmp(names[x], names[y]) < 0; });
  Writer w;
  w.alloc(96);  // superblock and root symbol-table entry
  // root group: symbol-table message -> B-tree + local heap
  std::vector<uint8_t> rm, body;
  Writer::le(body, 0, 8);  // B-tree address, patched below
  Writer::le(body, 0, 8);  // heap address, patched below
  Writer::message(rm, 0x11, body, 0);
  const uint64_t root = w.object_header(rm, 1);
  const uint64_t tree = w.alloc(24 + (2 * Writer::kInternalK + 1) * 8 + 2 * Writer::kInternalK * 8);
  // local heap: "" at offset 0, then the names (each padded to 8 bytes), then one free block
  std::vector<uint64_t> name_off(n);
  uint64_t hsize = 8;
  for (int i : order) {
    name_off[i] = hsize;
    hsize += (std::strlen(names[i]) + 1 + 7) & ~size_t(7);
  }
  const uint64_t free_off = hsize;
  hsize += 16;
  const uint64_t heap = w.alloc(32);
  const uint64_t hdata = w.alloc(hsize);
  std::memcpy(&w.b[heap], "HEAP", 4);
  w.put(heap + 8, hsize, 8);
  w.put(heap + 16, free_off, 8);
  w.put(heap + 24, hdata, 8);
  for (int i : order) std::memcpy(&w.b[hdata + name_off[i]], names[i], std::strlen(names[i]));
  w.put(hdata + free_off, 1, 8);    // end of the free list
  w.put(hdata + free_off + 8, 16, 8);
  // symbol-table node
  const uint64_t snod = w.alloc(8 + 2 * Writer::kLeafK * 40);
  std::memcpy(&w.b[snod], "SNOD", 4);
  w.b[snod + 4] = 1;
  w.put(snod + 6, (uint64_t)n, 2);
  // B-tree: one leaf node, keys = heap offsets ("" and the last name), child = the node
  std::memcpy(&w.b[tree], "TREE", 4);
  w.put(tree + 6, 1, 2);
  w.put(tree + 8, kUndef, 8);
  w.put(tree + 16, kUndef, 8);
  w.put(tree + 24, 0, 8);
  w.put(tree + 32, snod, 8);
  w.put(tree + 40, name_off[order[n - 1]], 8);
  w.put(root + 16 + 8, tree, 8);
  w.put(root + 16 + 16, heap, 8);
  // datasets
  std::vector<uint64_t> oh(n), raw(n), nbytes(n);
  for (int i = 0; i < n; i++) {
    const int rank = ranks[i];
    uint64_t cnt = 1;
    for (int r = 0; r < rank; r++) cnt *= (uint64_t)dims[2 * i + r];
    nbytes[i] = cnt * 8;
    std::vector<uint8_t> m, d;
    // dataspace, version 1, with maximum dimensions equal to the dimensions
    d = {1, (uint8_t)rank, 1, 0, 0, 0, 0, 0};
    for (int r = 0; r < rank; r++) Writer::le(d, (uint64_t)dims[2 * i + r], 8);
    for (int r = 0; r < rank; r++) Writer::le(d, (uint64_t)dims[2 * i + r], 8);
    Writer::message(m, 0x01, d, 0);
    // datatype: floating point, version 1, little endian, implied mantissa MSB,
    // sign bit 63, 8 bytes, precision 64, exponent at 52 (11 bits), mantissa at 0
    // (52 bits), bias 1023
    d = {0x11, 0x20, 0x3f, 0x00, 8, 0, 0, 0, 0, 0, 64, 0, 52, 11, 0, 52, 0xff, 0x03, 0, 0};
    Writer::message(m, 0x03, d, 1);
    // fill value, version 2: late allocation, written if set, defined with no value
    d = {2, 2, 2, 1, 0, 0, 0, 0};
    Writer::message(m, 0x05, d, 1);
    // contiguous layout, version 3 (address and size patched below)
    d = {3, 1};
    Writer::le(d, 0, 8);
    Writer::le(d, nbytes[i], 8);
    Writer::message(m, 0x08, d, 0);
    oh[i] = w.object_header(m, 4);
  }
  for (int i = 0; i < n; i++) {
    raw[i] = w.alloc(nbytes[i]);
    if (nbytes[i]) std::memcpy(&w.b[raw[i]], data[i], nbytes[i]);
    // layout message is the fourth: 16 (header) + 3 earlier messages
    const int rank = ranks[i];
    const uint64_t lay = oh[i] + 16 + (8 + 8 + 16ull * rank) + (8 + 24) + (8 + 8);
    w.put(lay + 8 + 2, nbytes[i] ? raw[i] : kUndef, 8);
  }
  for (int s = 0; s < n; s++) {
    const int i = order[s];
    const uint64_t ent = snod + 8 + 40ull * s;
    w.put(ent, name_off[i], 8);
    w.put(ent + 8, oh[i], 8);
  }
  // superblock version 0
  static const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
  std::memcpy(&w.b[0], sig, 8);
  w.b[13] = 8;  // size of offsets
  w.b[14] = 8;  // size of lengths
  w.put(16, Writer::kLeafK, 2);
  w.put(18, Writer::kInternalK, 2);
  w.put(24, 0, 8);            // base address
  w.put(32, kUndef, 8);       // free-space info
  w.put(40, w.b.size(), 8);   // end of file
  w.put(48, kUndef, 8);       // driver information
  w.put(56, 0, 8);            // root entry: name offset
  w.put(64, root, 8);         //   object header
  w.put(72, 1, 4);            //   cache type 1: B-tree and heap addresses follow
  w.put(80, tree, 8);
  w.put(88, heap, 8);
  FILE *f = std::fopen(path, "wb");
  if (!f) return DFN_ERR_ARG;
  const size_t put = std::fwrite(w.b.data(), 1, w.b.size(), f);
  const int closed = std::fclose(f);
  return (put == w.b.size() && closed == 0) ? DFN_OK : DFN_ERR_ARG;
}

struct dfn_h5_s {
  H5 f;
};

#define API extern "C" __attribute__((visibility("default")))
#include "api_guard.hpp"

// reads outside a truncated or corrupt file end the call with DFN_ERR_UNSUPPORTED
#define H5_READ_BEGIN try {
#define H5_READ_END                       \
  }                                       \
  catch (const H5Malformed &) {           \
    return DFN_ERR_UNSUPPORTED;           \
  }

API int dfn_h5_open(const char *path, dfn_h5 *out) {
  DFN_API_BEGIN
  H5_READ_BEGIN
  if (!path || !out) return DFN_ERR_ARG;
  *out = nullptr;
  dfn_h5 h = new (std::nothrow) dfn_h5_s;
  if (!h) return DFN_ERR_NOMEM;
  const int st = h->f.open(path);
  if (st != DFN_OK) { delete h; return st; }
  *out = h;
  return DFN_OK;
  H5_READ_END
  DFN_API_END
}

API int dfn_h5_close(dfn_h5 h) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  delete h;
  return DFN_OK;
  DFN_API_END
}

API int dfn_h5_info(dfn_h5 h, const char *path, int32_t *type_class, int32_t *rank, int64_t *dims) {
  DFN_API_BEGIN
  H5_READ_BEGIN
  if (!h || !path) return DFN_ERR_ARG;
  Dataset ds;
  const int st = h->f.dataset(path, ds);
  if (st != DFN_OK) return st;
  if (type_class) *type_class = ds.cls;
  if (rank) *rank = (int32_t)ds.dims.size();
  if (dims)
    for (size_t i = 0; i < ds.dims.size() && i < 8; i++) dims[i] = (int64_t)ds.dims[i];
  return DFN_OK;
  H5_READ_END
  DFN_API_END
}

API int dfn_h5_read_f64(dfn_h5 h, const char *path, double *buf, int64_t n) {
  DFN_API_BEGIN
  H5_READ_BEGIN
  if (!h || !path || !buf || n < 0) return DFN_ERR_ARG;
  Dataset ds;
  const int st = h->f.dataset(path, ds);
  if (st != DFN_OK) return st;
  if (ds.cls != 0 && ds.cls != 1) return DFN_ERR_UNSUPPORTED;
  if ((int64_t)H5::count(ds) != n) return DFN_ERR_ARG;
  std::vector<uint8_t> r;
  if (!h->f.raw(ds, r)) return DFN_ERR_UNSUPPORTED;
  for (int64_t i = 0; i < n; i++) buf[i] = h->f.number(ds, &r[(size_t)i * ds.size]);
  return DFN_OK;
  H5_READ_END
  DFN_API_END
}

API int dfn_h5_read_i64(dfn_h5 h, const char *path, int64_t *buf, int64_t n) {
  DFN_API_BEGIN
  H5_READ_BEGIN
  if (!h || !path || !buf || n < 0) return DFN_ERR_ARG;
  Dataset ds;
  const int st = h->f.dataset(path, ds);
  if (st != DFN_OK) return st;
  if (ds.cls != 0 && ds.cls != 1) return DFN_ERR_UNSUPPORTED;
  if ((int64_t)H5::count(ds) != n) return DFN_ERR_ARG;
  std::vector<uint8_t> r;
  if (!h->f.raw(ds, r)) return DFN_ERR_UNSUPPORTED;
  for (int64_t i = 0; i < n; i++) buf[i] = h->f.integer(ds, &r[(size_t)i * ds.size]);
  return DFN_OK;
  H5_READ_END
  DFN_API_END
}

API int dfn_h5_read_string(dfn_h5 h, const char *path, char *buf, int32_t cap) {
  DFN_API_BEGIN
  H5_READ_BEGIN
  if (!h || !path || !buf || cap < 1) return DFN_ERR_ARG;
  Dataset ds;
  const int st = h->f.dataset(path, ds);
  if (st != DFN_OK) return st;
  if (ds.cls != 2) return DFN_ERR_UNSUPPORTED;
  std::string s;
  if (!h->f.string(ds, s)) return DFN_ERR_UNSUPPORTED;
  const size_t n = std::min(s.size(), (size_t)cap - 1);
  std::memcpy(buf, s.data(), n);
  buf[n] = 0;
  return DFN_OK;
  H5_READ_END
  DFN_API_END
}

API int dfn_h5_write_f64(const char *path, int32_t n, const char *const *names, const int32_t *ranks,
                         const int64_t *dims, const double *const *data) {
  DFN_API_BEGIN
  return h5_write_f64(path, n, names, ranks, dims, data);
  DFN_API_END
}
